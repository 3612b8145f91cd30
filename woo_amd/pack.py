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
    """Largest relative overlap of a periodic packing (CPU, cell-binned), for tests."""
    n = len(radius)
    h = 2 * float(np.max(radius))
    nc = max(1, int(L // h))
    cell = np.floor(np.mod(pos, L) / (L / nc)).astype(np.int64) % nc
    key = (cell[:, 2] * nc + cell[:, 1]) * nc + cell[:, 0]
    order = np.argsort(key, kind="stable")
    worst = 0.
    buckets = {}
    for i in order:
        buckets.setdefault(int(key[i]), []).append(i)
    for i in range(n):
        cx, cy, cz = cell[i]
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    k = (((cz + dz) % nc) * nc + (cy + dy) % nc) * nc + (cx + dx) % nc
                    for j in buckets.get(int(k), ()):
                        if j <= i:
                            continue
                        d = pos[j] - pos[i]
                        d -= L * np.round(d / L)
                        ov = radius[i] + radius[j] - math.sqrt(float(d @ d))
                        worst = max(worst, ov / min(radius[i], radius[j]))
    return worst
