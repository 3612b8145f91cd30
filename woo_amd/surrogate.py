"""Learned coarse step: a message-passing network on the contact graph (hand-written BF16
MFMA kernels in csrc/wdg_mpnn.hip, driven through the C ABI).

The network maps the current state of a granular packing (contact graph with edges between
spheres closer than 1.1 (ri + rj); node inputs velocity / radius / overlap with the floor,
edge inputs relative position / overlap / relative velocity / radii, with the contact normal
scaled by the elastic and viscous Hertz terms) to each sphere's mean acceleration over a
coarse step of 10 DEM steps. ``Network.advance`` moves the spheres over that step;
``run_learned`` checks the learned steps against DEM re-simulation every ``check_every``
coarse steps and falls back to DEM when the position error exceeds the bound.

Training data come from the engine itself: ``column_training_set`` runs seeded columns as
independent scenes of one context, builds the graph of each snapshot on the device and keeps
it there with its targets, so that training never moves a graph through the host.

Scales: lengths by R0 (2.5 mm), velocities by V0 (1 m/s), accelerations (outputs and targets)
by A0 = 100 g (981 m/s^2), the order of the contact accelerations, so that the network's
outputs are of order one.
"""
import ctypes as C
import math

import numpy as np

from ._native import lib as _lib, dptr, iptr

R0, V0, G = 2.5e-3, 1.0, 9.81
A0 = 100. * G  # unit of the network's outputs (m/s^2); MPNN_A0 in csrc/wdg_mpnn.hip
CUTOFF = 1.1
NODE_IN, EDGE_IN = 7, 15
# overlaps are fractions 1e-3..1e-2 of the contact distance: scaled so that BF16 resolves them,
# and clipped as inputs of their own (the force inputs use them unclipped)
OV_SCALE, OV_CLIP = 100., 2.


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
        L.wdg_mpnn_set_targets.argtypes = [V, P(C.c_float)]
        L.wdg_mpnn_keep_graph.argtypes = [V, C.c_int32]
        L.wdg_mpnn_use_graph.argtypes = [V, C.c_int32]
        L.wdg_mpnn_get_graph.argtypes = [V, P(C.c_int64), P(C.c_int64), P(C.c_int32), P(C.c_int32), P(C.c_int32),
                                         P(C.c_float), P(C.c_float), P(C.c_int64)]
        L._mpnn_ready = True
    return L


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"woo_amd mpnn error {rc}: {_L().wdg_mpnn_last_error().decode()}")


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


class Graph:
    """contact graph in CSR by receiver: rowOff[N+1], send[E], rev[E] (opposite edge),
    node inputs x[N,7], edge inputs ef[E,15] (float32)"""

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
    vel = np.asarray(vel, dtype=np.float64)
    ef = edge_inputs(pos[recv], pos[send], vel[recv], vel[send], radius[recv], radius[send])
    return Graph(rowOff, send, rev, x, ef)


def edge_inputs(pi, pj, vi, vj, ri, rj):
    """inputs of edges with receivers i and senders j: the expressions of edge_inputs in
    csrc/wdg_mpnn.hip, in the same order (so that host and device graphs agree)"""
    rs = ri + rj
    dx, dy, dz = (pj[:, 0] - pi[:, 0]) / rs, (pj[:, 1] - pi[:, 1]) / rs, (pj[:, 2] - pi[:, 2]) / rs
    dn = np.sqrt((dx * dx + dy * dy) + dz * dz)
    ov = (1. - dn) * OV_SCALE
    wx, wy, wz = (vj[:, 0] - vi[:, 0]) / V0, (vj[:, 1] - vi[:, 1]) / V0, (vj[:, 2] - vi[:, 2]) / V0
    a, b = ri / R0, rj / R0
    pos_dn = dn > 0
    dsafe = np.where(pos_dn, dn, 1.)
    nx, ny, nz = (np.where(pos_dn, d / dsafe, 0.) for d in (dx, dy, dz))
    ovp = np.where(ov > 0, ov, 0.)
    a3, b3 = a * a * a, b * b * b
    hz = ovp * np.sqrt(ovp) * (np.sqrt(a * b) * (a + b) / a3)
    mb, Rb = a3 * b3 / (a3 + b3), a * b / (a + b)
    vn = (wx * nx + wy * ny) + wz * nz
    dp = vn * np.sqrt(np.sqrt(ovp)) * (np.sqrt(np.sqrt(a + b)) * np.sqrt(mb * np.sqrt(Rb)) / a3)
    ovc = np.minimum(np.maximum(ov, -OV_CLIP), OV_CLIP)
    cols = [dx, dy, dz, ovc, wx, wy, wz, a, b, -nx * hz, -ny * hz, -nz * hz, nx * dp, ny * dp, nz * dp]
    return np.stack(cols, axis=1).astype(np.float32)


def node_inputs(vel, radius, height):
    """velocity / V0, radius / R0, overlap with the floor ov = (1 - height / r) * OV_SCALE
    clipped, ov+^(3/2) / a and v_z ov+^(1/4) / a (a = r / R0): node_inputs in csrc/wdg_mpnn.hip"""
    radius = np.asarray(radius, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64)
    a = radius / R0
    ov = (1. - np.asarray(height) / radius) * OV_SCALE
    ovp = np.where(ov > 0, ov, 0.)
    vz = vel[:, 2] / V0
    cols = [vel[:, 0] / V0, vel[:, 1] / V0, vz, a, np.minimum(np.maximum(ov, -OV_CLIP), OV_CLIP),
            ovp * np.sqrt(ovp) / a, vz * np.sqrt(np.sqrt(ovp)) / a]
    return np.stack(cols, axis=1).astype(np.float32)


def contact_mask(g):
    """spheres in contact (a positive floor overlap input or a positive overlap input on one of
    their edges); the learned step moves the others in free flight"""
    recv = np.repeat(np.arange(g.N), np.diff(g.rowOff))
    hit = np.zeros(g.N, dtype=bool)
    np.logical_or.at(hit, recv, g.ef[:, 3] > 0)
    return (g.x[:, 4] > 0) | hit


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
        return out.astype(np.float64) * A0

    def step_acc(self, g, gravity=(0., 0., -G)):
        """the accelerations the learned step applies: the network's for spheres in contact,
        gravity for the others"""
        a = self.predict(g)
        a[~contact_mask(g)] = gravity
        return a

    def loss_grad(self, g, target_acc):
        """loss against target accelerations (m/s^2) and the gradients on the device"""
        self.set_graph(g, np.asarray(target_acc) / A0)
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
        return None if o is None else o.astype(np.float64) * A0

    def advance(self, dev, dT):
        """the context's spheres over the coarse step dT with the last forward pass"""
        _chk(self.L.wdg_mpnn_advance(self.h, dev.h, dT))

    def adam(self, lr=1e-3, b1=.9, b2=.999, eps=1e-8):
        _chk(self.L.wdg_mpnn_adam(self.h, lr, b1, b2, eps))

    def set_targets(self, acc):
        """targets (m/s^2) of the current graph, in its output order"""
        t = np.ascontiguousarray(np.asarray(acc) / A0, dtype=np.float32)
        _chk(self.L.wdg_mpnn_set_targets(self.h, _fptr(t)))

    def keep(self, slot):
        """keep the current graph and its targets on the device"""
        _chk(self.L.wdg_mpnn_keep_graph(self.h, slot))

    def use(self, slot):
        _chk(self.L.wdg_mpnn_use_graph(self.h, slot))

    def graph(self):
        """the current graph as a host Graph (inputs as stored in BF16) and the output row of
        each node"""
        n, e = C.c_int64(), C.c_int64()
        _chk(self.L.wdg_mpnn_get_graph(self.h, C.byref(n), C.byref(e), None, None, None, None, None, None))
        N, E = n.value, e.value
        rowOff, send, rev = np.empty(N + 1, np.int32), np.empty(max(E, 1), np.int32), np.empty(max(E, 1), np.int32)
        x, ef = np.empty((max(N, 1), NODE_IN), np.float32), np.empty((max(E, 1), EDGE_IN), np.float32)
        row = np.empty(max(N, 1), np.int64)
        _chk(self.L.wdg_mpnn_get_graph(self.h, None, None, iptr(rowOff), iptr(send), iptr(rev), _fptr(x), _fptr(ef),
                                       row.ctypes.data_as(C.POINTER(C.c_int64))))
        return Graph(rowOff, send[:E], rev[:E], x[:N], ef[:E]), row[:N]

    def kept_loss(self, slot):
        """loss (mean over nodes of |a - t|^2, in units of A0^2) on a kept graph; also leaves the
        gradients"""
        self.use(slot)
        loss = C.c_double()
        _chk(self.L.wdg_mpnn_grad(self.h, C.byref(loss)))
        return loss.value

    def train_kept(self, slots, epochs=10, lr=1e-3, lr_end=None, seed=0, log=None):
        """Adam over graphs kept on the device, one step per graph, shuffled each epoch; the
        learning rate falls geometrically from lr to lr_end; returns the mean loss per epoch"""
        rng = np.random.default_rng(seed)
        lr_end = lr if lr_end is None else lr_end
        hist = []
        for ep in range(epochs):
            r = lr * (lr_end / lr) ** (ep / max(epochs - 1, 1))
            tot = 0.
            for k in rng.permutation(len(slots)):
                tot += self.kept_loss(slots[k])
                self.adam(r)
            hist.append(tot / len(slots))
            if log:
                log(f"epoch {ep}: lr {r:.2e} loss {hist[-1]:.5f}")
        return hist

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


def column_training_set(net, slot0=0, n_columns=500, width=0.03, height=0.6, snapshots=20, every=1500, seed=100,
                        rollout=0, device=0, log=None):
    """Training data kept on the device: n_columns seeded columns (width x width x height,
    spheres on a jittered lattice, released onto their floors) run side by side as independent
    scenes of one context. At every ``every`` steps the graph of all columns is built on the
    device from the engine's slot list, and COARSE steps later the mean acceleration over the
    coarse step becomes its target; the graph is kept in slot slot0 + k of the network.

    With rollout > 0 the states come from the network itself: from each DEM snapshot the
    network advances the columns by 1..rollout coarse steps (varying with k), and the DEM
    target is computed from where it arrived (contact history restarted), so that training sees
    the drift of its own coarse steps and learns to correct it; the DEM trajectory then resumes
    from its snapshot.
    Returns (slots, ballistic mean squared error per snapshot (in units of A0^2), spheres)."""
    from ._native import Device

    side = int(math.ceil(math.sqrt(n_columns)))
    parts = []
    for c in range(n_columns):
        off = (c % side) * (width + 0.05), (c // side) * (width + 0.05), 0.
        p, r = column(width, height, seed=seed + c, offset=off)
        parts.append(engine_arrays(p, r))
    arr = {k: np.concatenate([q[k] for q in parts]) for k in parts[0]}
    arr["scene"] = np.concatenate([np.full(len(q["shape"]), i, np.int32) for i, q in enumerate(parts)])
    dev = Device(device)
    dt = dem_dt()
    dev.set_field(gravity=(0., 0., -G), periodic=False)
    dev.set_materials([material()])
    dev.upload_particles(**arr)
    dev.set_scenes(np.full(n_columns, dt))
    sph = arr["shape"] == 0
    slots, ball, step = [], [], 0
    skin = 0.25 * float(np.min(arr["radius"][sph]))
    for k in range(snapshots):
        step = dem_run(dev, every if k else 1, step, dt)
        if rollout:
            dev.snapshot(0)
            for j in range(1 + (7 * k) % rollout):
                net.graph_from(dev, floorZ=0., maxDisp=skin)
                net.forward(out=False)
                net.advance(dev, COARSE * dt)
            dev.snapshot(1)
            dev.restore(1)
            dev.collide(VERLET)
        n, e, _ = net.graph_from(dev, floorZ=0., force=True)
        v0 = dev.get("vel")[sph]
        dem_run(dev, COARSE, step, dt)
        acc = (dev.get("vel")[sph] - v0) / (COARSE * dt)
        if rollout:
            dev.restore(0)
        else:
            step += COARSE
        net.set_targets(acc)
        net.keep(slot0 + k)
        slots.append(slot0 + k)
        ball.append(float(np.mean(np.sum((acc - np.array([0., 0., -G])) ** 2, axis=1))) / A0 ** 2)
        if log:
            log(f"snapshot {k}: step {step}, nodes {n}, edges {e}, ballistic rms {math.sqrt(ball[-1]) * A0:.2f} m/s^2")
    dev.close()
    return slots, ball, int(np.sum(sph))


def held_out_rmse(net, slots):
    """root mean squared acceleration error (m/s^2) on kept graphs of the learned step and of the
    ballistic step (free fall for every sphere): the errors of spheres in contact, averaged over
    all spheres (a sphere in free flight moves the same in both). The ballistic step is the
    network with a zero decoder and the free-fall bias."""
    trained = float(np.mean([net.kept_loss(k) for k in slots]))
    p = net.params()
    q = p.copy()
    n = len(p)
    # the decoder is the last H * 3 + 3 parameters: weights then bias
    q[n - 3 - 64 * 3:n - 3] = 0.
    q[n - 3:] = (0., 0., -G / A0)
    net.set_params(q)
    ballistic = float(np.mean([net.kept_loss(k) for k in slots]))
    net.set_params(p)
    return math.sqrt(trained) * A0, math.sqrt(ballistic) * A0


# squat columns of the collapse's aspect ratio (height / width = 0.6 / 0.2706), a fifth of its
# size: (columns, width, height, snapshots, steps between snapshots)
SQUAT = (64, 0.054, 0.12, 20, 2000)


def train_on_columns(net, n_columns=500, epochs=30, rounds=1, rollout=50, snapshots=20, every=1500, seed=100,
                     squat=SQUAT, log=None):
    """Train the network on DEM trajectories of seeded columns (column_training_set): tall narrow
    ones (3 cm x 60 cm, which reach the collapse's speeds) and, with ``squat``, columns of the
    collapse's own aspect ratio (whose spreading front and settling pile the tall ones do not
    show); then for each of ``rounds`` rounds add the states its own coarse steps reach from
    the tall columns' snapshots (rollout) with their DEM targets, and train on the union.
    Returns the loss history, the ballistic loss of the DEM states, the spheres and the graphs."""
    slots, ball, n = column_training_set(net, 0, n_columns=n_columns, snapshots=snapshots, every=every, seed=seed,
                                         log=log)
    if squat:
        nc, w, h, sn, ev = squat
        more, b2, n2 = column_training_set(net, 10000, n_columns=nc, width=w, height=h, snapshots=sn, every=ev,
                                           seed=seed + 50000, log=log)
        slots, ball, n = slots + more, ball + b2, n + n2
    hist = net.train_kept(slots, epochs=epochs, lr=1e-3, lr_end=1e-4, log=log)
    for r in range(rounds):
        more, _, _ = column_training_set(net, 100 * (r + 1), n_columns=n_columns, snapshots=snapshots, every=every,
                                         seed=seed, rollout=rollout, log=log)
        slots = slots + more
        hist += net.train_kept(slots, epochs=max(epochs // 2, 1), lr=3e-4, lr_end=3e-5, seed=r + 1, log=log)
    return hist, ball, n, len(slots)


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


def run_learned(dev, net, radius, coarse_steps, dt, check_every=50, bound=0.2e-3, step=0, energy=None, slack=0.,
                backoff=False, log=None):
    """Coarse steps of COARSE DEM steps by the network, on the context dev (its spheres have the
    given radii). Every check_every coarse steps the last coarse step is redone by DEM from its
    start; when the largest position difference exceeds ``bound``, the segment since the
    previous check is redone by DEM (fallback). The state after a check is the DEM one. The
    states are kept as device snapshots (slots 0-2).

    The check sees one coarse step only, so errors that build up over many segments escape it.
    Two optional guards: with ``energy`` (a function of dev giving the total mechanical energy,
    kinetic plus gravitational), a segment whose total energy rose by more than ``slack`` also
    falls back, since contacts only dissipate; with ``backoff``, a fallback makes the next
    segment plain DEM, the next two after a second fallback in a row, and so on up to 64, the
    count starting again at one after a segment that passed. Returns statistics."""
    # room for motion between two slot-list refreshes of the graph builder
    skin = 0.25 * float(np.min(radius))
    stats = dict(checks=0, fallbacks=0, energy_fallbacks=0, dem_segments=0, errors=[], learned_coarse_steps=0,
                 dem_steps=0)
    k = 0
    dem_left, wait = 0, 1
    while k < coarse_steps:
        seg = min(check_every, coarse_steps - k)
        # the scene's step number at the segment's start: each coarse step stands for COARSE
        # DEM steps, so the check's DEM redo starts at step0 + (seg - 1) COARSE and a fallback
        # at step0 (the trsfRenorm phase of dem_run follows the true step)
        step0 = step
        if dem_left > 0:
            dem_run(dev, seg * COARSE, step0, dt)
            stats["dem_steps"] += seg * COARSE
            stats["dem_segments"] += 1
            dem_left -= 1
            step = step0 + seg * COARSE
            k += seg
            continue
        e0 = energy(dev) if energy else 0.
        dev.snapshot(0)
        for j in range(seg):
            if j == seg - 1:
                dev.snapshot(1)
            net.graph_from(dev, floorZ=0., maxDisp=skin)
            net.forward(out=False)
            net.advance(dev, COARSE * dt)
            stats["learned_coarse_steps"] += 1
        dev.snapshot(2)
        dev.restore(1)
        dem_run(dev, COARSE, step0 + (seg - 1) * COARSE, dt)
        stats["dem_steps"] += COARSE
        err = dev.snapshot_distance(2)
        stats["checks"] += 1
        stats["errors"].append(err)
        fail = err > bound
        if not fail and energy and energy(dev) > e0 + slack:
            fail = True
            stats["energy_fallbacks"] += 1
        if fail:
            stats["fallbacks"] += 1
            dev.restore(0)
            dem_run(dev, seg * COARSE, step0, dt)
            stats["dem_steps"] += seg * COARSE
            if backoff:
                dem_left, wait = wait, min(2 * wait, 64)
        else:
            wait = 1
        step = step0 + seg * COARSE
        if log:
            log(f"coarse {k + seg}/{coarse_steps}: check error {err * 1e3:.4f} mm"
                f"{' -> DEM fallback' if fail else ''}")
        k += seg
    stats["step"] = step
    return stats


def mechanical_energy(dev, mass, g=G):
    """kinetic energy of the context's integrated nodes plus the gravitational energy m g z of
    the spheres (mass per particle, zero for walls)"""
    return dev.kinetic_energy() + g * float(np.dot(mass, dev.get("pos")[:, 2]))


def collapse_to_rest(dev, dt, step=0, chunk=10000, max_steps=500000, log=None):
    """DEM until the collapsed pile is at rest: the kinetic energy, sampled every ``chunk``
    steps, is below 5 % of its peak and fell by less than 5 % over the last 4 samples. Spheres
    without rolling resistance keep rolling on the floor, so the energy levels off above zero
    instead of vanishing. Returns the step reached and the energy history."""
    hist = [(step, dev.kinetic_energy())]
    peak = hist[0][1]
    while step < max_steps:
        step = dem_run(dev, chunk, step, dt)
        ek = dev.kinetic_energy()
        peak = max(peak, ek)
        hist.append((step, ek))
        if log:
            log(f"collapse: step {step} kinetic energy {ek:.4e} J (peak {peak:.4e})")
        if len(hist) > 4 and ek < 0.05 * peak and hist[-5][1] - ek < 0.05 * ek:
            break
    return step, hist


def deposit(pos, radius, center, ring=0.005, rmax=2.5):
    """the pile's height profile (highest sphere top per ring of horizontal distance from the
    column axis, rings of ``ring`` m), its edge (outer radius of the last ring of the
    contiguous pile at least 4 sphere diameters high, so that spheres rolling away alone do not
    count) and the radius holding 99 % of the spheres"""
    rh = np.hypot(pos[:, 0] - center[0], pos[:, 1] - center[1])
    top = pos[:, 2] + radius
    nb = int(rmax / ring)
    prof = np.zeros(nb)
    k = np.clip((rh / ring).astype(np.int64), 0, nb - 1)
    np.maximum.at(prof, k, top)
    high = prof >= 8 * float(np.max(radius))
    edge = 0
    while edge < nb and high[edge]:
        edge += 1
    return prof, edge * ring, float(np.percentile(rh, 99))


def runout_profile(pos, radius, center, bins=np.linspace(0., 0.6, 61)):
    """run-out (largest horizontal distance of a sphere from the column axis) and the pile's
    height profile: highest sphere top per ring of horizontal distance"""
    rh = np.hypot(pos[:, 0] - center[0], pos[:, 1] - center[1])
    top = pos[:, 2] + radius
    prof = np.zeros(len(bins) - 1)
    k = np.clip(np.digitize(rh, bins) - 1, 0, len(bins) - 2)
    np.maximum.at(prof, k, top)
    return float(np.max(rh + radius)), prof
