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
    vol = float(np.sum((4 / 3.) * math.pi * radius ** 3))
    L = (vol / phi) ** (1 / 3.)
    pos = rng.uniform(0., L, size=(n, 3))
    own = dev is None
    d = Device(device) if own else dev
    d.set_field(periodic=True, cell=(L, L, L))
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
    return pos, radius, L, dict(maxOverlap=ov - tol, iterations=it, inflate=inflate)


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
