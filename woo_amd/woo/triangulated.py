"""woo.triangulated: triangulated surfaces of the DEM step (py/triangulated.py:11-79)."""
import math

import numpy as np

from . import utils


def _rotation_from_x(ax):
    """orthonormal frame whose first column is the unit vector ax"""
    ax = np.asarray(ax, dtype=float)
    ax = ax / np.linalg.norm(ax)
    t = np.array([0., 0., 1.]) if abs(ax[2]) < 0.9 else np.array([0., 1., 0.])
    e1 = np.cross(t, ax)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(ax, e1)
    return np.stack([ax, e1, e2], axis=1)


def cylinder(A, B, radius, div=20, axDiv=1, capA=False, capB=False, masks=None, wallCaps=False, angVel=0,
             realRot=False, fixed=True, **kw):
    """Triangulated cylinder from A to B as a list of facets (and end walls with wallCaps);
    with angVel the mantle moves through Facet.fakeVel = -radius*angVel*(n x axis), the
    reference's default (realRot=False). Caps come first, the mantle after.

    The mantle is swept between ``div`` points on each base circle (the first and last
    coincide, as in the reference's linspace with endpoint), ``axDiv`` segments along the axis
    (negative: that ratio to the circumferential segment length)."""
    if realRot:
        raise NotImplementedError("realRot (StableCircularOrbit on the nodes) needs moving facet nodes")
    A, B = np.asarray(A, dtype=float), np.asarray(B, dtype=float)
    cylLen = np.linalg.norm(B - A)
    ax = (B - A) / cylLen
    rot = _rotation_from_x(ax)
    thetas = np.linspace(0, 2 * math.pi, num=div, endpoint=True)
    if axDiv < 0:
        axDiv = int(abs(axDiv) * cylLen / (2 * radius * math.pi / div))
    xx = np.linspace(0, cylLen, num=max(axDiv + 1, 2))
    ring = [(radius * math.cos(t), radius * math.sin(t)) for t in thetas]
    pts = [[A + rot @ np.array([x, y, z]) for (y, z) in ring] for x in xx]
    caps = []
    if wallCaps:
        cAx = -1
        for i in (0, 1, 2):
            if abs(ax[i]) > 1 - 1e-6:
                cAx = i
        if cAx < 0:
            raise ValueError("With wallCaps=True, cylinder axis must be aligned with some global axis")
        sign = int(math.copysign(1, ax[cAx]))
        wkw = {k: v for k, v in kw.items() if k in ("mat", "mask", "color", "visible")}
        if capA:
            caps.append(utils.wall(tuple(A), axis=cAx, sense=sign, **wkw))
        if capB:
            caps.append(utils.wall(tuple(B), axis=cAx, sense=-sign, **wkw))
    elif capA or capB:
        raise NotImplementedError("triangulated caps are not supported; use wallCaps=True")
    ff = []
    fkw = {k: v for k, v in kw.items() if k in ("mat", "mask", "halfThick", "color", "visible", "wire")}
    for i in range(len(xx) - 1):
        for j in range(len(ring) - 1):
            a0, a1, b0, b1 = pts[i][j], pts[i][j + 1], pts[i + 1][j], pts[i + 1][j + 1]
            for tri in ((a0, b0, b1), (a0, b1, a1)):
                if np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0])) == 0:
                    continue
                ff.append(utils.facet([tuple(v) for v in tri], fixed=fixed, **fkw))
    if angVel != 0:
        for f in ff:
            f.shape.fakeVel = tuple(-radius * angVel * np.cross(f.shape.getNormal(), ax))
    if masks:
        for f in ff:
            f.mask = masks[0]
        if wallCaps:
            if capA:
                caps[0].mask = masks[1]
            if capB:
                caps[-1].mask = masks[2]
    return caps + ff
