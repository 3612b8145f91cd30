"""Seeded initial conditions for DEM scenes (the role of woo.pack / SpherePack).

``periodic_pack`` fills a periodic cube to a target solid fraction with
polydisperse spheres: radii and centres are drawn from a seeded generator, then
the GPU relaxation pass (wdg_relax) pushes overlapping spheres apart with
slightly inflated radii, so the returned packing is free of overlaps.
"""
import math

import numpy as np

from ._native import Device, SPHERE


def periodic_pack(n, rmin, rmax, phi, seed=2, inflate=0.005, max_iter=4000, device=0, dev=None):
    """Return (pos[n,3], radius[n], L, info) for n spheres with radius ~ U[rmin,rmax]
    in a periodic cube of edge L such that the solid fraction is ``phi``.

    Relaxation stops once the largest overlap of the inflated radii is below
    2*rmin*inflate, which leaves the true radii without any overlap."""
    rng = np.random.default_rng(seed)
    radius = rng.uniform(rmin, rmax, size=n)
    pos, L3, info = box_pack(radius, phi, (1., 1., 1.), rng, inflate, max_iter, device, dev)
    return pos, radius, L3[0], info


def box_pack(radius, phi, aspect, rng, inflate=0.005, max_iter=4000, device=0, dev=None):
    """Dense random packing of spheres of the given radii in a periodic box of the given aspect
    ratio, sized so that the solid fraction is ``phi``. Returns (pos, (Lx, Ly, Lz), info)."""
    n = len(radius)
    vol = float(np.sum((4 / 3.) * math.pi * radius ** 3))
    a = np.asarray(aspect, dtype=float)
    s = (vol / phi / float(np.prod(a))) ** (1 / 3.)
    L3 = tuple(float(x) for x in s * a)
    pos = rng.uniform(0., 1., size=(n, 3)) * np.array(L3)
    own = dev is None
    d = Device(device) if own else dev
    d.set_field(periodic=True, cell=L3)
    d.set_materials([dict(kind=0, density=1., young=1., ktDivKn=.2, tanPhi=.5, surfEnergy=0., alpha=0.)])
    zeros = np.zeros((n, 3))
    d.upload_particles(shape=np.full(n, SPHERE, np.int32), pos=pos, radius=radius, mass=np.ones(n),
                       inertia=np.ones((n, 3)), vel=zeros)
    tol = 2 * float(np.min(radius)) * inflate
    ov, it = d.relax(max_iter, inflate, tol=tol)
    if not ov <= tol:
        raise RuntimeError(f"packing relaxation did not converge: overlap {ov} > {tol} after {it} sweeps")
    pos = d.get("pos")
    if own:
        d.close()
    return pos, L3, dict(maxOverlap=ov - tol, iterations=it, inflate=inflate)


def drum_fill(n, R, length, rmin_clip, rmax_clip, median, sigma, phi=0.58, seed=4, dev=None):
    """n spheres (radii lognormal(median, sigma) clipped to [rmin_clip, rmax_clip]) in the lower
    part of a drum of (inscribed) radius R and given length (axis x, centred at the origin): a
    dense periodic packing of the bounding box of a circular segment is cut to the spheres that
    clear the mantle and the end planes, and the lowest n are kept, so that no sphere overlaps
    another or the boundary. The segment is sized for n spheres at solid fraction phi (plus 12 %
    for the spheres lost at the boundary). Returns (pos, radius, info) with the filled share of
    the cross-section in info."""
    rng = np.random.default_rng(seed)
    sample = np.clip(np.exp(rng.normal(math.log(median), sigma, size=200000)), rmin_clip, rmax_clip)
    ev = float(np.mean((4 / 3.) * math.pi * sample ** 3))
    fill = min(1.12 * n * ev / (phi * math.pi * R * R * length), 0.95)
    # segment of central angle t holds (t - sin t) / (2 pi) of the disc
    lo, hi = 0., 2 * math.pi
    for _ in range(100):
        t = .5 * (lo + hi)
        lo, hi = (t, hi) if (t - math.sin(t)) / (2 * math.pi) < fill else (lo, t)
    h = R * (1 - math.cos(t / 2))
    W = 2 * math.sqrt(max(R * R - (R - h) ** 2, 0.)) if h < R else 2 * R
    # radii drawn until their volume fills the box (length, W, h) at phi
    target = phi * length * W * h
    radius = np.empty(0)
    vol = 0.
    while vol < target:
        r = np.clip(np.exp(rng.normal(math.log(median), sigma, size=max(n // 4, 1000))), rmin_clip, rmax_clip)
        v = np.cumsum((4 / 3.) * math.pi * r ** 3)
        k = int(np.searchsorted(v, target - vol)) + 1
        radius = np.concatenate([radius, r[:k]])
        vol += float(v[min(k, len(v)) - 1])
    pos, L3, info = box_pack(radius, phi, (length, W, h), rng, dev=dev)
    x = np.mod(pos[:, 0], L3[0]) - L3[0] / 2
    y = np.mod(pos[:, 1], L3[1]) - L3[1] / 2
    z = np.mod(pos[:, 2], L3[2]) - R
    ok = (np.abs(x) <= length / 2 - radius) & (np.hypot(y, z) <= R - radius)
    idx = np.flatnonzero(ok)
    if len(idx) < n:
        raise RuntimeError(f"drum fill: only {len(idx)} of {n} spheres clear the boundary")
    idx = idx[np.argsort(z[idx], kind="stable")][:n]
    info.update(boxSpheres=len(radius), kept=int(ok.sum()), fill=fill, segmentHeight=h, box=L3)
    return np.stack([x[idx], y[idx], z[idx]], 1), radius[idx], info


def check_overlaps(pos, radius, L):
    """Largest overlap of a periodic packing relative to the smaller radius (CPU, for tests);
    negative means the closest pair is separated."""
    from scipy.spatial import cKDTree

    x = np.mod(pos, L)
    x[x >= L] = 0.
    tree = cKDTree(x, boxsize=L)
    pairs = tree.query_pairs(2 * float(np.max(radius)), output_type="ndarray")
    if len(pairs) == 0:
        return -np.inf
    i, j = pairs[:, 0], pairs[:, 1]
    d = x[j] - x[i]
    d -= L * np.round(d / L)
    ov = radius[i] + radius[j] - np.linalg.norm(d, axis=1)
    return float(np.max(ov / np.minimum(radius[i], radius[j])))
