"""Learned coarse step: a message-passing network on the contact graph (hand-written BF16
MFMA kernels in csrc/wdg_mpnn.hip, driven through the C ABI).

The network maps the current state of a granular packing (contact graph with edges between
spheres closer than 1.1 (ri + rj), node inputs velocity / radius / overlap with the floor,
edge inputs relative position / overlap / relative velocity / radii) to each sphere's mean
acceleration over a coarse step of 10 DEM steps. ``Surrogate.advance`` moves the spheres
over that step; ``run_with_fallback`` checks the learned steps against DEM re-simulation
every ``check_every`` coarse steps and falls back to the DEM result when the position error
exceeds the bound.

Scales: lengths by R0 (2.5 mm), velocities by V0 (1 m/s), accelerations by G (9.81 m/s^2).
"""
import ctypes as C
import math

import numpy as np

from ._native import lib as _lib, dptr, iptr

R0, V0, G = 2.5e-3, 1.0, 9.81
CUTOFF = 1.1
NODE_IN, EDGE_IN = 5, 9
# overlaps are fractions 1e-3..1e-2 of the contact distance: scaled so that BF16 resolves them
OV_SCALE, OV_CLIP = 100., 20.


def _L():
    L = _lib()
    if not getattr(L, "_mpnn_ready", False):
        P = C.POINTER
        V = C.c_void_p
        L.wdg_mpnn_last_error.restype = C.c_char_p
        L.wdg_mpnn_create.argtypes = [C.c_int, C.c_uint64, P(V)]
        L.wdg_mpnn_destroy.argtypes = [V]
        L.wdg_mpnn_num_params.argtypes = [V, P(C.c_int64)]
        L.wdg_mpnn_get_params.argtypes = [V, P(C.c_float), P(C.c_float)]
        L.wdg_mpnn_set_params.argtypes = [V, P(C.c_float)]
        L.wdg_mpnn_set_graph.argtypes = [V, C.c_int64, C.c_int64, P(C.c_float), P(C.c_int32), P(C.c_int32),
                                         P(C.c_int32), P(C.c_float), P(C.c_float)]
        L.wdg_mpnn_forward.argtypes = [V, P(C.c_float)]
        L.wdg_mpnn_grad.argtypes = [V, P(C.c_double)]
        L.wdg_mpnn_adam.argtypes = [V, C.c_double, C.c_double, C.c_double, C.c_double]
        L.wdg_mpnn_graph_from_ctx.argtypes = [V, V, C.c_double, C.c_double, C.c_double, C.c_int, P(C.c_int64),
                                              P(C.c_int64), P(C.c_int)]
        L.wdg_mpnn_advance.argtypes = [V, V, C.c_double]
        L._mpnn_ready = True
    return L


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"woo_amd mpnn error {rc}: {_L().wdg_mpnn_last_error().decode()}")


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


class Graph:
    """contact graph in CSR by receiver: rowOff[N+1], send[E], rev[E] (opposite edge),
    node inputs x[N,5], edge inputs ef[E,9] (float32)"""

    def __init__(self, rowOff, send, rev, x, ef):
        self.rowOff = np.ascontiguousarray(rowOff, dtype=np.int32)
        self.send = np.ascontiguousarray(send, dtype=np.int32)
        self.rev = np.ascontiguousarray(rev, dtype=np.int32)
        self.x = np.ascontiguousarray(x, dtype=np.float32)
        self.ef = np.ascontiguousarray(ef, dtype=np.float32)

    @property
    def N(self):
        return len(self.rowOff) - 1

    @property
    def E(self):
        return len(self.send)


def contact_graph(pos, vel, radius, floor_z=0.):
    """Graph of spheres closer than CUTOFF (ri + rj), from positions/velocities/radii (host)."""
    from scipy.spatial import cKDTree

    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    rmax = float(np.max(radius)) if n else 0.
    pairs = cKDTree(pos).query_pairs(CUTOFF * 2 * rmax, output_type="ndarray") if n else np.zeros((0, 2), int)
    if len(pairs):
        i, j = pairs[:, 0], pairs[:, 1]
        d = np.linalg.norm(pos[j] - pos[i], axis=1)
        keep = d < CUTOFF * (radius[i] + radius[j])
        i, j = i[keep], j[keep]
    else:
        i = j = np.zeros(0, dtype=np.int64)
    recv = np.concatenate([i, j])
    send = np.concatenate([j, i])
    order = np.lexsort((send, recv))
    recv, send = recv[order], send[order]
    E = len(recv)
    rowOff = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowOff, recv + 1, 1)
    rowOff = np.cumsum(rowOff)
    # opposite edge: (send, recv) located by its sort key
    key = recv.astype(np.int64) * n + send
    rkey = send.astype(np.int64) * n + recv
    rev = np.searchsorted(key, rkey)
    x = node_inputs(vel, radius, pos[:, 2] - floor_z)
    rs = radius[recv] + radius[send]
    dx = (pos[send] - pos[recv]) / rs[:, None]
    dist = np.sqrt((dx[:, 0] * dx[:, 0] + dx[:, 1] * dx[:, 1]) + dx[:, 2] * dx[:, 2])
    ef = np.concatenate([dx, ((1. - dist) * OV_SCALE)[:, None], (vel[send] - vel[recv]) / V0,
                         (radius[recv] / R0)[:, None], (radius[send] / R0)[:, None]], axis=1)
    return Graph(rowOff, send, rev, x, ef)


def node_inputs(vel, radius, height):
    """velocity / V0, radius / R0, overlap with the floor (1 - height / r) * OV_SCALE clipped"""
    radius = np.asarray(radius)
    ov = np.clip((1. - np.asarray(height) / radius) * OV_SCALE, -OV_CLIP, OV_CLIP)
    return np.concatenate([np.asarray(vel) / V0, (radius / R0)[:, None], ov[:, None]], axis=1)


class Network:
    """the device network: parameters, Adam state, the current graph"""

    def __init__(self, seed=1, device=0):
        self.L = _L()
        self.h = C.c_void_p()
        _chk(self.L.wdg_mpnn_create(device, seed, C.byref(self.h)))
        n = C.c_int64()
        _chk(self.L.wdg_mpnn_num_params(self.h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if self.h:
            self.L.wdg_mpnn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        _chk(self.L.wdg_mpnn_get_params(self.h, _fptr(p), None))
        return p

    def grads(self):
        g = np.empty(self.nparams, dtype=np.float32)
        _chk(self.L.wdg_mpnn_get_params(self.h, None, _fptr(g)))
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _chk(self.L.wdg_mpnn_set_params(self.h, _fptr(p)))

    def set_graph(self, g, target=None):
        t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
        _chk(self.L.wdg_mpnn_set_graph(self.h, g.N, g.E, _fptr(g.x), iptr(g.rowOff), iptr(g.send), iptr(g.rev),
                                       _fptr(g.ef), _fptr(t)))

    def predict(self, g):
        """mean acceleration over the coarse step (m/s^2), per node"""
        self.set_graph(g)
        self.N = g.N
        out = np.empty((g.N, 3), dtype=np.float32)
        _chk(self.L.wdg_mpnn_forward(self.h, _fptr(out)))
        return out.astype(np.float64) * G

    def loss_grad(self, g, target_acc):
        """loss against target accelerations (m/s^2) and the gradients on the device"""
        self.set_graph(g, np.asarray(target_acc) / G)
        loss = C.c_double()
        _chk(self.L.wdg_mpnn_grad(self.h, C.byref(loss)))
        return loss.value

    def graph_from(self, dev, floorZ=0., maxDisp=0., force=False):
        """graph of the spheres of an engine context (woo_amd._native.Device) on the device;
        returns (nodes, edges, rebuilt)"""
        n, e, rb = C.c_int64(), C.c_int64(), C.c_int()
        _chk(self.L.wdg_mpnn_graph_from_ctx(self.h, dev.h, CUTOFF, floorZ, maxDisp, int(force), C.byref(n),
                                            C.byref(e), C.byref(rb)))
        self.N = n.value
        return n.value, e.value, bool(rb.value)

    def forward(self, out=True):
        """forward pass on the current graph; accelerations (m/s^2) when out"""
        o = np.empty((self.N, 3), dtype=np.float32) if out else None
        _chk(self.L.wdg_mpnn_forward(self.h, _fptr(o)))
        return None if o is None else o.astype(np.float64) * G

    def advance(self, dev, dT):
        """the context's spheres over the coarse step dT with the last forward pass"""
        _chk(self.L.wdg_mpnn_advance(self.h, dev.h, dT))

    def adam(self, lr=1e-3, b1=.9, b2=.999, eps=1e-8):
        _chk(self.L.wdg_mpnn_adam(self.h, lr, b1, b2, eps))

    def train(self, samples, epochs=10, lr=1e-3, seed=0, log=None):
        """samples: list of (Graph, target accelerations); one Adam step per sample, shuffled
        each epoch; returns the mean loss of every epoch"""
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            tot = 0.
            for k in rng.permutation(len(samples)):
                g, t = samples[k]
                tot += self.loss_grad(g, t)
                self.adam(lr)
            hist.append(tot / len(samples))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5f}")
        return hist


# ---------------------------------------------------------------- scenes and data
RHO, YOUNG, TANPHI, POISSON, EN = 3000., 1e8, .4, .25, .6
COARSE = 10  # DEM steps per coarse step
VERLET = -.16  # relative Verlet distance: covers the edge cutoff 1.1 (ri + rj) for r in [2, 3] mm


def column(width, height, rmin=2e-3, rmax=3e-3, seed=0, offset=(0., 0., 0.)):
    """spheres (radius uniform in [rmin, rmax]) on a jittered lattice filling a column of the
    given width (x, y) and height (z) standing on the floor z = offset z; returns (pos, radius)"""
    rng = np.random.default_rng(seed)
    a = 2 * rmax * 1.002
    m = max(int(width // a), 1)
    k = max(int(height // a), 1)
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(k), indexing="ij"), -1).reshape(-1, 3)
    pos = (g + 0.5) * a + rng.uniform(-1, 1, size=g.shape) * 0.0005 * a
    pos[:, 2] += rmax * 0.002
    radius = rng.uniform(rmin, rmax, size=len(pos))
    return pos + np.asarray(offset), radius


def engine_arrays(pos, radius, floor_z=0., scene=None, n_walls_first=True):
    """one floor wall then the spheres, for Device.upload_particles"""
    from ._native import SPHERE, WALL

    n = len(radius)
    shape = np.concatenate([[WALL], np.full(n, SPHERE)]).astype(np.int32)
    P = np.concatenate([[[0., 0., floor_z]], pos])
    rad = np.concatenate([[0.], radius])
    mass = np.concatenate([[0.], (4 / 3.) * math.pi * radius ** 3 * RHO])
    inertia = np.repeat((0. + 1. * (2. / 5.) * mass * rad * rad)[:, None], 3, 1)
    axis = np.full(n + 1, 2, np.int32)
    sense = np.zeros(n + 1, np.int32)
    sense[0] = 1
    flags = np.zeros(n + 1, np.uint32)
    flags[0] = 63
    integ = np.ones(n + 1, np.uint8)
    integ[0] = 0
    mask = np.full(n + 1, 5, np.uint32)
    mask[0] = 3
    out = dict(shape=shape, pos=P, radius=rad, wallAxis=axis, wallSense=sense, mask=mask, flags=flags,
               integrated=integ, mass=mass, inertia=inertia, vel=np.zeros((n + 1, 3)))
    return out


def material():
    return dict(kind=1, density=RHO, young=YOUNG, ktDivKn=.2, tanPhi=TANPHI, surfEnergy=0., alpha=0.)


def contact_params():
    from ._native import LAW_HERTZDMT, PHYS_HERTZ

    return dict(physKind=PHYS_HERTZ, lawKind=LAW_HERTZDMT, poisson=POISSON, en=EN, trsfRenorm=100)


def dem_dt(rmin=2e-3):
    return 0.3 * rmin / math.sqrt(YOUNG / RHO)


ENGINES = ["Collider", "ContactLoop", "Leapfrog"]


def dem_run(dev, steps, step, dt):
    return dev.run(ENGINES, steps, step, dt, collide=dict(verletDist=VERLET, forceInit=0), contact=contact_params(),
                   leapfrog=dict(damping=0.), dynDtPeriod=0)[0]


def column_samples(n_columns=500, width=0.03, height=0.06, snapshots=20, every=1500, seed=100, device=0, log=None):
    """Training data from DEM: n_columns seeded columns released onto their floors, run side by
    side as independent scenes of one context; at every ``every`` steps a snapshot (graph of each
    column) and, COARSE steps later, the mean acceleration over the coarse step as target.
    Returns a list of (Graph, target accelerations) per snapshot, holding all columns."""
    from ._native import Device

    cols = []
    for c in range(n_columns):
        off = (c % 25) * (width + 0.05), (c // 25) * (width + 0.05), 0.
        cols.append(column(width, height, seed=seed + c, offset=off))
    dev = Device(device)
    parts = [engine_arrays(p, r) for p, r in cols]
    arr = {k: np.concatenate([q[k] for q in parts]) for k in parts[0]}
    arr["scene"] = np.concatenate([np.full(len(q["shape"]), i, np.int32) for i, q in enumerate(parts)])
    dt = dem_dt()
    dev.set_field(gravity=(0., 0., -9.81), periodic=False)
    dev.set_materials([material()])
    dev.upload_particles(**arr)
    dev.set_scenes(np.full(n_columns, dt))
    sph = arr["shape"] == 0
    rad = arr["radius"][sph]
    samples, step = [], 0
    for k in range(snapshots):
        step = dem_run(dev, every if k else 1, step, dt)
        P0, V0_ = dev.get("pos")[sph], dev.get("vel")[sph]
        step = dem_run(dev, COARSE, step, dt)
        V1 = dev.get("vel")[sph]
        acc = (V1 - V0_) / (COARSE * dt)
        g = contact_graph(P0, V0_, rad, floor_z=0.)
        samples.append((g, acc, P0, V0_))
        if log:
            log(f"snapshot {k}: step {step}, edges {g.E}, |a| mean {np.mean(np.linalg.norm(acc, axis=1)):.2f}")
    dev.close()
    return samples, rad


# ---------------------------------------------------------------- column collapse
def collapse_arrays(width=0.2706, height=0.6, seed=4):
    """the granular column: spheres (r uniform in [2, 3] mm) on a lattice filling a
    width x width x height column on the floor (about 200,000 spheres at the default size)"""
    pos, rad = column(width, height, seed=seed)
    return engine_arrays(pos, rad), rad


def upload(dev, arr, state=None):
    """all particles (fresh contact history), optionally with a saved (pos, vel, angVel, ori)"""
    a = dict(arr)
    if state is not None:
        a["pos"], a["vel"], a["angVel"], a["ori"] = state
    dev.set_field(gravity=(0., 0., -9.81), periodic=False)
    dev.set_materials([material()])
    dev.upload_particles(**a)


def snapshot(dev):
    return dev.get("pos"), dev.get("vel"), dev.get("angVel"), dev.get("ori")


def run_learned(dev, net, arr, coarse_steps, dt, check_every=50, bound=0.2e-3, log=None):
    """Coarse steps of COARSE DEM steps by the network, on the context dev. Every check_every
    coarse steps the last coarse step is re-simulated by DEM from its start; when the largest
    position difference exceeds ``bound`` the segment since the previous check is redone by
    DEM (fallback). The state after a check is the DEM one. Returns statistics."""
    skin = 0.  # the graph reuses the slot list while no sphere moved farther than this
    # the context's Verlet distance covers the edge cutoff with room for this displacement
    rmin = float(np.min(arr["radius"][arr["shape"] == 0]))
    skin = abs(VERLET) * rmin - 0.05 * 2 * float(np.max(arr["radius"]))
    stats = dict(checks=0, fallbacks=0, errors=[], learned_coarse_steps=0, dem_steps=0)
    step, k = 0, 0
    sph = arr["shape"] == 0
    while k < coarse_steps:
        seg = min(check_every, coarse_steps - k)
        start = snapshot(dev)
        pre = None
        for j in range(seg):
            if j == seg - 1:
                pre = snapshot(dev)
            net.graph_from(dev, floorZ=0., maxDisp=max(skin, 0.))
            net.forward(out=False)
            net.advance(dev, COARSE * dt)
            stats["learned_coarse_steps"] += 1
        p_sur = dev.get("pos")[sph]
        upload(dev, arr, pre)
        step = dem_run(dev, COARSE, step, dt)
        stats["dem_steps"] += COARSE
        err = float(np.max(np.linalg.norm(dev.get("pos")[sph] - p_sur, axis=1)))
        stats["checks"] += 1
        stats["errors"].append(err)
        if err > bound:
            stats["fallbacks"] += 1
            upload(dev, arr, start)
            step = dem_run(dev, seg * COARSE, step, dt)
            stats["dem_steps"] += seg * COARSE
        if log:
            log(f"coarse {k + seg}/{coarse_steps}: check error {err * 1e3:.4f} mm"
                f"{' -> DEM fallback' if err > bound else ''}")
        k += seg
    return stats


def runout_profile(pos, radius, center, bins=np.linspace(0., 0.6, 61)):
    """run-out (largest horizontal distance of a sphere from the column axis) and the pile's
    height profile: highest sphere top per ring of horizontal distance"""
    rh = np.hypot(pos[:, 0] - center[0], pos[:, 1] - center[1])
    top = pos[:, 2] + radius
    prof = np.zeros(len(bins) - 1)
    k = np.clip(np.digitize(rh, bins) - 1, 0, len(bins) - 2)
    np.maximum.at(prof, k, top)
    return float(np.max(rh + radius)), prof
