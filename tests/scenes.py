"""Seeded test scenes, built identically on the CPU oracle and on the HIP engine.

A scene is a plain dict of numpy arrays (the woo.utils.sphere / woo.utils.wall
inputs); mass and inertia follow Sphere::lumpMassInertia (Sphere.cpp:26-33).
"""
import math

import numpy as np

MOVABLE_MASK = 0b0101  # DemField.defaultMovableMask
BOUNDARY_MASK = 0b0011  # DemField.defaultBoundaryMask
DOF_ALL = 63


def sphere_mass_inertia(r, rho):
    r = np.asarray(r, dtype=np.float64)
    m = (4 / 3.) * math.pi * (r * r * r) * rho
    i = 0. + 1. * (2. / 5.) * m * (r * r)
    return m, i


def _rss_positions(rng, n, lo, hi, radii, max_tries=2000000):
    """random sequential placement without overlap (cell-binned rejection)"""
    pos = np.empty((n, 3))
    rmax = float(np.max(radii))
    h = 2 * rmax
    cells = {}
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random placement did not converge")
        r = radii[k]
        c = rng.uniform(lo + r, hi - r, size=3)
        key = tuple((c // h).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in cells.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if np.sum((pos[j] - c) ** 2) < (radii[j] + r) ** 2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pos[k] = c
            cells.setdefault(key, []).append(k)
            k += 1
    return pos


def deposition(n=2000, seed=1, box=0.2, r=5e-3, rho=3000., young=1e7, ktDivKn=.2, tanPhi=.5, dense=False,
               walls_first=False, vel_sigma=0.):
    """Gravity deposition in a box of 6 walls (FrictMat). dense=True packs spheres on a
    jittered lattice with slight overlaps so that contacts exist from the first step."""
    rng = np.random.default_rng(seed)
    radii = np.full(n, r)
    if dense:
        a = 1.98 * r
        m = int(math.ceil(n ** (1 / 3.)))
        g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n].astype(float)
        pos = (g + 0.5) * a + rng.uniform(-0.01 * r, 0.01 * r, size=(n, 3))
        box = m * a
    else:
        pos = _rss_positions(rng, n, 0., box, radii)
    vel = rng.normal(0, vel_sigma, size=(n, 3)) if vel_sigma > 0 else np.zeros((n, 3))
    mass, inertia = sphere_mass_inertia(radii, rho)
    walls = []
    for sense, coord in ((1, 0.), (-1, box)):
        for ax in (0, 1, 2):
            p = np.zeros(3)
            p[ax] = coord
            walls.append(dict(pos=p, axis=ax, sense=sense))
    pwave = r / math.sqrt(young / rho)
    return dict(
        materials=[dict(kind=0, density=rho, young=young, ktDivKn=ktDivKn, tanPhi=tanPhi, surfEnergy=0., alpha=0.)],
        pos=pos, radius=radii, vel=vel, mass=mass, inertia=np.repeat(inertia[:, None], 3, 1),
        walls=walls, walls_first=walls_first, gravity=(0., 0., -9.81), periodic=False, cell=(1., 1., 1.),
        dt=0.5 * pwave, physKind=0, lawKind=0, damping=0.2, box=box)


def periodic_hertz(n=1000, seed=2, rmin=2e-3, rmax=4e-3, rho=3000., young=1e8, tanPhi=.4, poisson=.25, en=.6,
                   spacing=1.0, vel_sigma=0.1):
    """Periodic Hertz box with polydisperse spheres on a jittered cubic lattice;
    spacing is the lattice pitch relative to 2*rmax (below 1 gives initial contacts)."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, size=n)
    m = int(math.ceil(n ** (1 / 3.)))
    a = spacing * 2 * rmax
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n].astype(float)
    pos = (g + 0.5) * a + rng.uniform(-0.02 * rmin, 0.02 * rmin, size=(n, 3))
    L = m * a
    # unwrapped coordinates: shift some particles by whole periods
    pos += L * rng.integers(-1, 2, size=(n, 3)) * (rng.uniform(size=(n, 1)) < 0.1)
    vel = rng.normal(0, vel_sigma, size=(n, 3))
    mass, inertia = sphere_mass_inertia(radii, rho)
    pwave = rmin / math.sqrt(young / rho)
    return dict(
        materials=[dict(kind=1, density=rho, young=young, ktDivKn=.2, tanPhi=tanPhi, surfEnergy=0., alpha=0.)],
        pos=pos, radius=radii, vel=vel, mass=mass, inertia=np.repeat(inertia[:, None], 3, 1),
        walls=[], walls_first=False, gravity=(0., 0., 0.), periodic=True, cell=(L, L, L),
        dt=0.3 * pwave, physKind=1, lawKind=1, damping=0., poisson=poisson, en=en, box=L)


def periodic_bidisperse(n_side=18, r_small=1e-3, r_big=6e-3, n_big=3, seed=4, rho=3000., young=1e8, tanPhi=.4,
                        poisson=.25, en=.6, vel_sigma=0.05):
    """Periodic Hertz box: small spheres on a jittered lattice with a few large spheres put
    in their place. A large sphere has well over a hundred near neighbours (the rare case of
    the broadphase, whose per-sphere staging then overflows)."""
    rng = np.random.default_rng(seed)
    a = 2.02 * r_small
    L = n_side * a
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(float)
    small = (g + 0.5) * a + rng.uniform(-0.01 * r_small, 0.01 * r_small, size=g.shape)
    big = []
    while len(big) < n_big:
        c = rng.uniform(0, L, size=3)
        if all(np.linalg.norm((c - b) - L * np.round((c - b) / L)) > 2 * r_big + 4 * a for b in big):
            big.append(c)
    keep = np.ones(len(small), bool)
    shell = []
    n_shell = 140
    k = np.arange(n_shell) + 0.5
    # Fibonacci points: small spheres on the large sphere's surface, slightly overlapping it
    u = np.stack([np.cos(np.pi * (1 + 5 ** 0.5) * k) * np.sqrt(1 - (1 - 2 * k / n_shell) ** 2),
                  np.sin(np.pi * (1 + 5 ** 0.5) * k) * np.sqrt(1 - (1 - 2 * k / n_shell) ** 2),
                  1 - 2 * k / n_shell], -1)
    for b in big:
        d = small - b
        d -= L * np.round(d / L)
        keep &= np.linalg.norm(d, axis=1) > r_big + 3.2 * r_small
        shell.append(b + (r_big + 0.995 * r_small) * u)
    pos = np.concatenate([small[keep], np.concatenate(shell), np.array(big)])
    radii = np.concatenate([np.full(int(keep.sum()) + n_shell * n_big, r_small), np.full(n_big, r_big)])
    perm = rng.permutation(len(radii))
    pos, radii = pos[perm], radii[perm]
    vel = rng.normal(0, vel_sigma, size=pos.shape)
    mass, inertia = sphere_mass_inertia(radii, rho)
    pwave = r_small / math.sqrt(young / rho)
    return dict(
        materials=[dict(kind=1, density=rho, young=young, ktDivKn=.2, tanPhi=tanPhi, surfEnergy=0., alpha=0.)],
        pos=pos, radius=radii, vel=vel, mass=mass, inertia=np.repeat(inertia[:, None], 3, 1),
        walls=[], walls_first=False, gravity=(0., 0., 0.), periodic=True, cell=(L, L, L),
        dt=0.3 * pwave, physKind=1, lawKind=1, damping=0., poisson=poisson, en=en, box=L)


def drum(n=2000, seed=11, R=0.04, length=0.03, div=64, omega=2., rmin=1.99e-3, rmax=2e-3, pitch=3.99e-3, rho=3000.,
         young=1e8, tanPhi=.4, poisson=.25, en=.6, vel_sigma=0.05, caps=True, halfThick=0.):
    """Rotating drum (woo.triangulated.cylinder with fakeVel, triangulated.py:11-79): axis x,
    2*div facets spanning the length, fakeVel = -R*omega*(n x axis), end walls; spheres on a
    lattice in the lower part, the outermost pushed onto the facets so that facet contacts
    exist from the first step. Hertz law, gravity -z."""
    rng = np.random.default_rng(seed)
    ax = np.array([1., 0., 0.])
    th = np.linspace(0, 2 * math.pi, div + 1)
    facets = []
    for k in range(div):
        y0, z0, y1, z1 = R * math.cos(th[k]), R * math.sin(th[k]), R * math.cos(th[k + 1]), R * math.sin(th[k + 1])
        A0, A1 = (-length / 2, y0, z0), (-length / 2, y1, z1)
        B0, B1 = (length / 2, y0, z0), (length / 2, y1, z1)
        for tri in ((A0, B0, B1), (A0, B1, A1)):
            v = np.array(tri)
            nrm = np.cross(v[1] - v[0], v[2] - v[0])
            nrm /= np.linalg.norm(nrm)
            facets.append(dict(verts=v, fakeVel=-R * omega * np.cross(nrm, ax), halfThick=halfThick))
    Rin = R * math.cos(math.pi / div)
    a = pitch
    xs = np.arange(-length / 2 + rmax * 1.05, length / 2 - rmax * 1.05 + 1e-12, a)
    ys = np.arange(-Rin, Rin + 1e-12, a)
    g = np.array([(x, y, z) for z in ys for y in ys for x in xs if math.hypot(y, z) < Rin - 0.6 * rmax])
    g = g[np.argsort(g[:, 2], kind="stable")][:n]
    n = len(g)
    radii = np.exp(rng.uniform(math.log(rmin), math.log(rmax), size=n))
    pos = g + rng.uniform(-0.02 * rmin, 0.02 * rmin, size=(n, 3))
    # spheres near the mantle moved out until they overlap the plane of their facet by 0.1 % of r
    rho_ax = np.hypot(pos[:, 1], pos[:, 2])
    phi = np.arctan2(pos[:, 2], pos[:, 1]) % (2 * math.pi)
    dphi = phi - (np.floor(phi / (2 * math.pi / div)) + 0.5) * (2 * math.pi / div)
    out = rho_ax * np.cos(dphi) > Rin - 2 * radii - halfThick
    scale = (Rin - 0.999 * radii[out] - halfThick) / np.cos(dphi[out]) / rho_ax[out]
    pos[out, 1] *= scale
    pos[out, 2] *= scale
    vel = rng.normal(0, vel_sigma, size=(n, 3))
    mass, inertia = sphere_mass_inertia(radii, rho)
    walls = []
    if caps:
        walls = [dict(pos=np.array([-length / 2, 0., 0.]), axis=0, sense=1),
                 dict(pos=np.array([length / 2, 0., 0.]), axis=0, sense=-1)]
    pwave = rmin / math.sqrt(young / rho)
    return dict(
        materials=[dict(kind=1, density=rho, young=young, ktDivKn=.2, tanPhi=tanPhi, surfEnergy=0., alpha=0.)],
        pos=pos, radius=radii, vel=vel, mass=mass, inertia=np.repeat(inertia[:, None], 3, 1),
        walls=walls, facets=facets, walls_first=True, gravity=(0., 0., -9.81), periodic=False, cell=(1., 1., 1.),
        dt=0.3 * pwave, physKind=1, lawKind=1, damping=0., poisson=poisson, en=en, box=2 * R)


def facet_box(n=800, seed=12, **kw):
    """deposition() with each of the 6 walls replaced by two facets of the box face"""
    sc = deposition(n=n, seed=seed, **kw)
    b = sc["box"]
    c = [np.array([x, y, z]) for z in (0., b) for y in (0., b) for x in (0., b)]  # corner index x + 2y + 4z
    faces = [(0, 2, 6, 4), (1, 3, 7, 5), (0, 1, 5, 4), (2, 3, 7, 6), (0, 1, 3, 2), (4, 5, 7, 6)]
    facets = []
    for q in faces:
        v = [c[i] for i in q]
        facets.append(dict(verts=np.array([v[0], v[1], v[2]])))
        facets.append(dict(verts=np.array([v[0], v[2], v[3]])))
    sc["walls"] = []
    sc["facets"] = facets
    return sc


def _particle_order(sc):
    """list of ('s', i) / ('w', j) / ('f', k) in id order"""
    n = len(sc["radius"])
    sph = [("s", i) for i in range(n)]
    wal = [("w", j) for j in range(len(sc["walls"]))] + [("f", k) for k in range(len(sc.get("facets", [])))]
    return wal + sph if sc["walls_first"] else sph + wal


def n_boundary(sc):
    return len(sc["walls"]) + len(sc.get("facets", []))


def build_oracle(sc, **params):
    from oracle import Oracle

    p = dict(dt=sc["dt"], gravity=sc["gravity"], periodic=int(sc["periodic"]), cell=sc["cell"],
             physKind=sc["physKind"], lawKind=sc["lawKind"], damping=sc["damping"],
             poisson=sc.get("poisson", .2), en=sc.get("en", float("nan")), deterministic=1)
    p.update(params)
    o = Oracle(**p)
    for m in sc["materials"]:
        o.material(**m)
    for kind, i in _particle_order(sc):
        if kind == "s":
            o.sphere(sc["pos"][i], sc["radius"][i], 0, False, MOVABLE_MASK)
        elif kind == "w":
            w = sc["walls"][i]
            o.wall(w["pos"], w["axis"], w["sense"], 0, BOUNDARY_MASK)
        else:
            f = sc["facets"][i]
            fid = o.facet(f["verts"], f.get("fakeVel"), f.get("halfThick", 0.), 0, BOUNDARY_MASK)
            if f.get("n21lim") is not None:
                o.set_facet_n21lim(fid, f["n21lim"])
    order = _particle_order(sc)
    o.set("vel", _gather(sc, order, "vel", 3))
    o.set("mass", _gather(sc, order, "mass", 1))
    o.set("inertia", _gather(sc, order, "inertia", 3))
    return o


def _gather(sc, order, key, w):
    out = np.zeros((len(order), w)) if w > 1 else np.zeros(len(order))
    for k, (kind, i) in enumerate(order):
        if kind == "s":
            out[k] = sc[key][i]
    return out


def gpu_arrays(sc):
    order = _particle_order(sc)
    n = len(order)
    shape = np.array([{"s": 0, "w": 1, "f": 2}[k] for k, _ in order], dtype=np.int32)
    pos = np.zeros((n, 3))
    fverts = np.zeros((n, 9))
    fvel = np.zeros((n, 3))
    fht = np.zeros(n)
    fn21 = np.full((n, 3), np.nan)
    radius = np.zeros(n)
    axis = np.zeros(n, dtype=np.int32)
    sense = np.zeros(n, dtype=np.int32)
    mask = np.zeros(n, dtype=np.uint32)
    flags = np.zeros(n, dtype=np.uint32)
    integ = np.zeros(n, dtype=np.uint8)
    for k, (kind, i) in enumerate(order):
        if kind == "s":
            pos[k] = sc["pos"][i]
            radius[k] = sc["radius"][i]
            mask[k] = MOVABLE_MASK
            integ[k] = 1
        elif kind == "w":
            w = sc["walls"][i]
            pos[k] = w["pos"]
            axis[k] = w["axis"]
            sense[k] = w["sense"]
            mask[k] = BOUNDARY_MASK
            flags[k] = DOF_ALL
        else:
            f = sc["facets"][i]
            fverts[k] = np.asarray(f["verts"], dtype=float).reshape(9)
            pos[k] = fverts[k, :3]
            if f.get("fakeVel") is not None:
                fvel[k] = f["fakeVel"]
            fht[k] = f.get("halfThick", 0.)
            if f.get("n21lim") is not None:
                fn21[k] = f["n21lim"]
            mask[k] = BOUNDARY_MASK
            flags[k] = DOF_ALL
    out = dict(shape=shape, pos=pos, radius=radius, wallAxis=axis, wallSense=sense, material=np.zeros(n, np.int32),
               mask=mask, flags=flags, integrated=integ, vel=_gather(sc, order, "vel", 3),
               mass=_gather(sc, order, "mass", 1), inertia=_gather(sc, order, "inertia", 3))
    if sc.get("facets"):
        out.update(facetVerts=fverts, fakeVel=fvel, halfThick=fht, facetN21lim=fn21)
    return out


def smooth_facets(sc):
    """Facet.n21lim of every facet of the scene, by the oracle's Facet::computeNeighborAngles
    (facets sharing vertices are edge neighbours)"""
    o = build_oracle(sc)
    for k, (kind, i) in enumerate(_particle_order(sc)):
        if kind == "f":
            sc["facets"][i]["n21lim"] = None
    o.facet_neighbor_angles()
    for k, (kind, i) in enumerate(_particle_order(sc)):
        if kind == "f":
            sc["facets"][i]["n21lim"] = o.facet_n21lim(k)
    return sc


def build_gpu(sc, device=0):
    from woo_amd._native import Device

    d = Device(device)
    d.set_field(gravity=sc["gravity"], periodic=sc["periodic"], cell=sc["cell"])
    d.set_materials(sc["materials"])
    d.upload_particles(**gpu_arrays(sc))
    return d


def contact_kw(sc):
    return dict(physKind=sc["physKind"], lawKind=sc["lawKind"], poisson=sc.get("poisson", .2),
                en=sc.get("en", float("nan")))


class GpuRunner:
    """Drives the HIP engine through the C ABI one engine at a time (Scene::doOneStep)."""

    def __init__(self, sc, verletDist=-.05, device=0):
        self.sc = sc
        self.d = build_gpu(sc, device)
        self.step = 0
        self.dt = sc["dt"]
        self.vd = verletDist

    def run(self, engines, steps=1):
        for _ in range(steps):
            for e in engines:
                if e == "Collider":
                    self.d.collide(self.vd)
                elif e == "ContactLoop":
                    self.d.contact_loop(self.step, self.dt, **contact_kw(self.sc))
                elif e == "Leapfrog":
                    self.d.leapfrog(self.dt, self.sc["damping"], True)
                elif e == "ForceResetter":
                    self.d.force_reset()
                elif e == "IntraForce":
                    self.d.intra_force()
                else:
                    raise ValueError(e)
            self.step += 1


def translated(sc, offset):
    """the scene moved by offset (positions of spheres and walls)"""
    out = dict(sc)
    off = np.asarray(offset, dtype=float)
    out["pos"] = sc["pos"] + off
    out["walls"] = [dict(w, pos=np.asarray(w["pos"], dtype=float) + off) for w in sc["walls"]]
    return out


def batch_arrays(scenes):
    """one upload holding several independent scenes: each scene's particles in its own id
    order, scene after scene; material k belongs to scene k; returns (arrays, materials, dt)"""
    parts = [gpu_arrays(sc) for sc in scenes]
    keys = parts[0].keys()
    out = {}
    for k in keys:
        out[k] = np.concatenate([p[k] for p in parts])
    out["material"] = np.concatenate([np.full(len(p["shape"]), i, np.int32) for i, p in enumerate(parts)])
    out["scene"] = np.concatenate([np.full(len(p["shape"]), i, np.int32) for i, p in enumerate(parts)])
    mats = [sc["materials"][0] for sc in scenes]
    return out, mats, np.array([sc["dt"] for sc in scenes])


def build_batch(scenes, device=0):
    from woo_amd._native import Device

    d = Device(device)
    sc0 = scenes[0]
    d.set_field(gravity=sc0["gravity"], periodic=False)
    arrays, mats, dts = batch_arrays(scenes)
    d.set_materials(mats)
    d.upload_particles(**arrays)
    d.set_scenes(dts)
    return d, arrays
