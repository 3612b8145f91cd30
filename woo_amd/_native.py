"""ctypes binding of the C ABI declared in include/woo_amd.h.

Loading fails loudly when libwoo_amd.so is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwoo_amd.so")

SPHERE, WALL, FACET = 0, 1, 2
FRICTMAT, HERTZMAT = 0, 1
PHYS_FRICT, PHYS_HERTZ = 0, 1
LAW_IDEALELPL, LAW_HERTZDMT = 0, 1
FIELD = dict(pos=0, vel=1, angVel=2, ori=3, force=4, torque=5, mass=6, inertia=7)
FIELD_W = dict(pos=3, vel=3, angVel=3, ori=4, force=3, torque=3, mass=1, inertia=3)
ENG = dict(ForceResetter=1, Collider=2, ContactLoop=3, Leapfrog=4, DynDt=5, IntraForce=6)

# every entry point include/woo_amd.h declares
EXPORTS = [
    "wdg_last_error", "wdg_ctx_create", "wdg_ctx_destroy", "wdg_set_field", "wdg_set_materials",
    "wdg_upload_particles", "wdg_upload", "wdg_download", "wdg_collide", "wdg_contact_loop", "wdg_leapfrog",
    "wdg_force_reset", "wdg_crit_dt", "wdg_contact_count", "wdg_contact_list", "wdg_observables", "wdg_stats",
    "wdg_run", "wdg_relax", "wdg_sync", "wdg_stream", "wdg_stress", "wdg_kinetic_energy",
    "wdg_set_precision", "wdg_set_scenes", "wdg_run_graph", "wdg_tune", "wdg_get_tuning", "wdg_set_tuning",
    "wdg_mpnn_last_error", "wdg_mpnn_create", "wdg_mpnn_destroy", "wdg_mpnn_num_params", "wdg_mpnn_get_params",
    "wdg_mpnn_set_params", "wdg_mpnn_set_graph", "wdg_mpnn_forward", "wdg_mpnn_grad", "wdg_mpnn_adam",
    "wdg_mpnn_graph_from_ctx", "wdg_mpnn_advance", "wdg_snapshot", "wdg_restore", "wdg_snapshot_distance",
    "wdg_mpnn_set_targets", "wdg_mpnn_keep_graph", "wdg_mpnn_use_graph", "wdg_mpnn_get_graph",
    "wdg_set_energy_tracking", "wdg_energy", "wdg_energy_clear", "wdg_restore_contacts", "wdg_energy_set",
    "wdg_kinetic_energy_noalloc", "wdg_intra_force", "wdg_facet_nodes", "wdg_set_update_phys",
    "wdg_get_update_phys", "wdg_get_scene_dt", "wdg_crit_dt_scenes", "wdg_set_eval_stress",
    "wdg_contact_loop_stress", "wdg_energy_scenes", "wdg_energy_scenes_set",
]

# Scene.energy terms, in WDG_EN_* order
ENERGY_TERMS = ("kinetic", "elast", "grav", "nonviscDamp", "plast", "viscN", "viscT")


class Field(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("distFactor", C.c_double), ("loneMask", C.c_uint32),
                ("periodic", C.c_int32), ("cell", C.c_double * 3)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("density", C.c_double), ("young", C.c_double), ("ktDivKn", C.c_double),
                ("tanPhi", C.c_double), ("surfEnergy", C.c_double), ("alpha", C.c_double)]


_P = C.POINTER


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64), ("shape", _P(C.c_int32)), ("radius", _P(C.c_double)),
                ("wallAxis", _P(C.c_int32)), ("wallSense", _P(C.c_int32)), ("material", _P(C.c_int32)),
                ("mask", _P(C.c_uint32)), ("flags", _P(C.c_uint32)), ("integrated", _P(C.c_uint8)),
                ("pos", _P(C.c_double)), ("vel", _P(C.c_double)), ("angVel", _P(C.c_double)),
                ("ori", _P(C.c_double)), ("force", _P(C.c_double)), ("torque", _P(C.c_double)),
                ("mass", _P(C.c_double)), ("inertia", _P(C.c_double)), ("facetVerts", _P(C.c_double)),
                ("fakeVel", _P(C.c_double)), ("halfThick", _P(C.c_double)), ("scene", _P(C.c_int32)),
                ("facetN21lim", _P(C.c_double))]


class Tuning(C.Structure):
    _fields_ = [("cfBlocks", C.c_int32), ("graphSteps", C.c_int32), ("skinSteps", C.c_int32)]


class CollideParams(C.Structure):
    _fields_ = [("verletDist", C.c_double), ("forceInit", C.c_int32)]


class ContactParams(C.Structure):
    _fields_ = [("step", C.c_int64), ("dt", C.c_double), ("physKind", C.c_int32), ("lawKind", C.c_int32),
                ("applyForces", C.c_int32), ("trsfRenorm", C.c_int32), ("poisson", C.c_double), ("en", C.c_double),
                ("noSlip", C.c_int32), ("noBreak", C.c_int32), ("noFrict", C.c_int32), ("noAttraction", C.c_int32),
                ("relRollStiff", C.c_double), ("relTwistStiff", C.c_double), ("rollTanPhi", C.c_double)]


class LeapfrogParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("damping", C.c_double), ("reset", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [("nParticles", C.c_int64), ("nCandidates", C.c_int64), ("nReal", C.c_int64),
                ("nRefresh", C.c_int64), ("nSteps", C.c_int64), ("verletDist", C.c_double),
                ("brute", C.c_int32), ("gridCells", C.c_int32 * 3), ("refreshSeconds", C.c_double)]


class Program(C.Structure):
    _fields_ = [("nEngines", C.c_int32), ("engines", C.c_int32 * 8), ("collide", CollideParams),
                ("contact", ContactParams), ("leapfrog", LeapfrogParams), ("dynDtPeriod", C.c_int32),
                ("maxRelInc", C.c_double), ("dtSafety", C.c_double)]


_lib = None


def lib():
    """Load libwoo_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m woo_amd.build` (there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    V, I = C.c_void_p, C.c_int
    L.wdg_last_error.restype = C.c_char_p
    L.wdg_ctx_create.argtypes = [I, _P(V)]
    L.wdg_ctx_destroy.argtypes = [V]
    L.wdg_set_field.argtypes = [V, _P(Field)]
    L.wdg_set_materials.argtypes = [V, _P(Material), C.c_int32]
    L.wdg_upload_particles.argtypes = [V, _P(Particles)]
    L.wdg_set_scenes.argtypes = [V, C.c_int32, _P(C.c_double)]
    L.wdg_run_graph.argtypes = [V, _P(Program), C.c_int64, _P(C.c_int64), _P(C.c_double)]
    L.wdg_tune.argtypes = [V, _P(Program), _P(C.c_int64), _P(C.c_double), _P(Tuning), _P(C.c_int64)]
    L.wdg_get_tuning.argtypes = [V, _P(Tuning)]
    L.wdg_set_tuning.argtypes = [V, _P(Tuning)]
    L.wdg_upload.argtypes = [V, C.c_int32, _P(C.c_double)]
    L.wdg_download.argtypes = [V, C.c_int32, _P(C.c_double)]
    L.wdg_collide.argtypes = [V, _P(CollideParams)]
    L.wdg_contact_loop.argtypes = [V, _P(ContactParams)]
    L.wdg_leapfrog.argtypes = [V, _P(LeapfrogParams)]
    L.wdg_force_reset.argtypes = [V]
    L.wdg_intra_force.argtypes = [V]
    L.wdg_set_update_phys.argtypes = [V, C.c_int32]
    L.wdg_get_scene_dt.argtypes = [V, _P(C.c_double)]
    L.wdg_set_eval_stress.argtypes = [V, C.c_int32]
    L.wdg_contact_loop_stress.argtypes = [V, _P(C.c_double)]
    L.wdg_crit_dt_scenes.argtypes = [V, _P(C.c_double)]
    L.wdg_get_update_phys.argtypes = [V, _P(C.c_int32)]
    L.wdg_facet_nodes.argtypes = [V, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_double)]
    L.wdg_crit_dt.argtypes = [V, _P(C.c_double)]
    L.wdg_contact_count.argtypes = [V, C.c_int32, _P(C.c_int64)]
    L.wdg_contact_list.argtypes = [V, C.c_int32, _P(C.c_int32), _P(C.c_int32), _P(C.c_double)]
    L.wdg_observables.argtypes = [V, _P(C.c_double), _P(C.c_double), _P(C.c_double)]
    L.wdg_stats.argtypes = [V, _P(Counters)]
    L.wdg_stress.argtypes = [V, C.c_double, _P(C.c_double)]
    L.wdg_kinetic_energy.argtypes = [V, _P(C.c_double)]
    L.wdg_kinetic_energy_noalloc.argtypes = [V, V]
    L.wdg_set_energy_tracking.argtypes = [V, C.c_int32]
    L.wdg_energy.argtypes = [V, _P(C.c_double)]
    L.wdg_energy_clear.argtypes = [V]
    L.wdg_energy_set.argtypes = [V, _P(C.c_double)]
    L.wdg_energy_scenes.argtypes = [V, _P(C.c_double)]
    L.wdg_energy_scenes_set.argtypes = [V, _P(C.c_double)]
    L.wdg_restore_contacts.argtypes = [V, C.c_int64, _P(C.c_int32), _P(C.c_int32), _P(C.c_double), C.c_double]
    L.wdg_set_precision.argtypes = [V, C.c_int32]
    L.wdg_run.argtypes = [V, _P(Program), C.c_int64, _P(C.c_int64), _P(C.c_double)]
    L.wdg_relax.argtypes = [V, C.c_int32, C.c_double, C.c_double, _P(C.c_double), _P(C.c_int32)]
    L.wdg_sync.argtypes = [V]
    L.wdg_stream.argtypes = [V]
    L.wdg_stream.restype = V
    L.wdg_snapshot.argtypes = [V, C.c_int32]
    L.wdg_restore.argtypes = [V, C.c_int32]
    L.wdg_snapshot_distance.argtypes = [V, C.c_int32, _P(C.c_double)]
    _lib = L
    return L


class WdgError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise WdgError(f"woo_amd error {rc}: {lib().wdg_last_error().decode()}")


def dptr(a):
    return a.ctypes.data_as(_P(C.c_double)) if a is not None else None


def iptr(a, t=C.c_int32):
    return a.ctypes.data_as(_P(t)) if a is not None else None


class Device:
    """One engine context on a HIP device (owns all device buffers of a DemField)."""

    def __init__(self, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        check(self.L.wdg_ctx_create(device, C.byref(self.h)))
        self.n = 0
        # what was set up, kept for checkpoint()
        self._field, self._mats, self._static, self._scenes = None, None, None, None
        self._tracking = False

    def close(self):
        if self.h:
            self.L.wdg_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_field(self, gravity=(0, 0, 0), distFactor=-1., loneMask=2, periodic=False, cell=(1, 1, 1)):
        f = Field()
        f.gravity[:] = list(gravity)
        f.distFactor = distFactor
        f.loneMask = loneMask
        f.periodic = int(bool(periodic))
        f.cell[:] = list(cell)
        check(self.L.wdg_set_field(self.h, C.byref(f)))
        self._field = dict(gravity=[float(x) for x in gravity], distFactor=float(distFactor), loneMask=int(loneMask),
                           periodic=bool(periodic), cell=[float(x) for x in cell])

    def set_materials(self, mats):
        arr = (Material * max(1, len(mats)))()
        for i, m in enumerate(mats):
            arr[i] = Material(**m)
        check(self.L.wdg_set_materials(self.h, arr, len(mats)))
        self._mats = [{k: (float(v) if k != "kind" else int(v)) for k, v in m.items()} for m in mats]

    def upload_particles(self, shape, pos, mass, inertia, radius=None, wallAxis=None, wallSense=None, material=None,
                         mask=None, flags=None, integrated=None, vel=None, angVel=None, ori=None, force=None,
                         torque=None, facetVerts=None, fakeVel=None, halfThick=None, scene=None, facetN21lim=None):
        n = len(shape)
        keep = []

        def arr(x, dt, w=None):
            if x is None:
                return None
            a = np.ascontiguousarray(x, dtype=dt)
            if w is not None:
                a = a.reshape(n, w) if w > 1 else a.reshape(n)
            keep.append(a)
            return a

        p = Particles()
        p.n = n
        p.shape = iptr(arr(shape, np.int32))
        p.radius = dptr(arr(radius, np.float64))
        p.wallAxis = iptr(arr(wallAxis, np.int32))
        p.wallSense = iptr(arr(wallSense, np.int32))
        p.material = iptr(arr(material, np.int32))
        p.mask = iptr(arr(mask, np.uint32), C.c_uint32)
        p.flags = iptr(arr(flags, np.uint32), C.c_uint32)
        p.integrated = iptr(arr(integrated, np.uint8), C.c_uint8)
        p.pos = dptr(arr(pos, np.float64, 3))
        p.vel = dptr(arr(vel, np.float64, 3))
        p.angVel = dptr(arr(angVel, np.float64, 3))
        p.ori = dptr(arr(ori, np.float64, 4))
        p.force = dptr(arr(force, np.float64, 3))
        p.torque = dptr(arr(torque, np.float64, 3))
        p.mass = dptr(arr(mass, np.float64))
        p.inertia = dptr(arr(inertia, np.float64, 3))
        p.facetVerts = dptr(arr(facetVerts, np.float64, 9))
        p.fakeVel = dptr(arr(fakeVel, np.float64, 3))
        p.halfThick = dptr(arr(halfThick, np.float64))
        p.scene = iptr(arr(scene, np.int32))
        p.facetN21lim = dptr(arr(facetN21lim, np.float64, 3))
        check(self.L.wdg_upload_particles(self.h, C.byref(p)))
        self.n = n
        static = dict(shape=shape, radius=radius, wallAxis=wallAxis, wallSense=wallSense, material=material, mask=mask,
                      flags=flags, integrated=integrated, facetVerts=facetVerts, fakeVel=fakeVel, halfThick=halfThick,
                      scene=scene, facetN21lim=facetN21lim)
        self._static = {k: np.array(v) for k, v in static.items() if v is not None}

    def get(self, field):
        w = FIELD_W[field]
        out = np.empty((self.n, w) if w > 1 else (self.n,), dtype=np.float64)
        check(self.L.wdg_download(self.h, FIELD[field], dptr(out)))
        return out

    def set(self, field, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(self.L.wdg_upload(self.h, FIELD[field], dptr(a)))

    def collide(self, verletDist=-.05, forceInit=False):
        p = CollideParams(verletDist=verletDist, forceInit=int(forceInit))
        check(self.L.wdg_collide(self.h, C.byref(p)))

    def contact_loop(self, step, dt, physKind=PHYS_FRICT, lawKind=LAW_IDEALELPL, applyForces=True, trsfRenorm=100,
                     poisson=.2, en=float("nan"), noSlip=False, noBreak=False, noFrict=False, noAttraction=True,
                     relRollStiff=0., relTwistStiff=0., rollTanPhi=0.):
        p = ContactParams(step=step, dt=dt, physKind=physKind, lawKind=lawKind, applyForces=int(applyForces),
                          trsfRenorm=trsfRenorm, poisson=poisson, en=en, noSlip=int(noSlip), noBreak=int(noBreak),
                          noFrict=int(noFrict), noAttraction=int(noAttraction), relRollStiff=relRollStiff,
                          relTwistStiff=relTwistStiff, rollTanPhi=rollTanPhi)
        check(self.L.wdg_contact_loop(self.h, C.byref(p)))

    def leapfrog(self, dt, damping=.2, reset=True):
        p = LeapfrogParams(dt=dt, damping=damping, reset=int(reset))
        check(self.L.wdg_leapfrog(self.h, C.byref(p)))

    def force_reset(self):
        check(self.L.wdg_force_reset(self.h))

    def intra_force(self):
        """IntraForce([In2_Facet]): the facets' contact forces spread to their nodes"""
        check(self.L.wdg_intra_force(self.h))

    def facet_nodes(self, ids):
        """node forces [n, 3, 3] and torques [n, 3, 3] (facet, node, component) of the facets ids
        from the last IntraForce"""
        a = np.ascontiguousarray(np.atleast_1d(ids), dtype=np.int64)
        out = np.empty((len(a), 18), dtype=np.float64)
        check(self.L.wdg_facet_nodes(self.h, a.ctypes.data_as(C.POINTER(C.c_int64)), len(a), dptr(out)))
        return out[:, :9].reshape(-1, 3, 3), out[:, 9:].reshape(-1, 3, 3)

    def crit_dt(self):
        d = C.c_double()
        check(self.L.wdg_crit_dt(self.h, C.byref(d)))
        return d.value

    def contacts(self, real_only=False, data=True):
        w = 1 if real_only else 0
        n = C.c_int64()
        check(self.L.wdg_contact_count(self.h, w, C.byref(n)))
        n = n.value
        ids = np.empty((n, 2), dtype=np.int32)
        cd = np.empty((n, 3), dtype=np.int32)
        d = np.empty((n, 16), dtype=np.float64) if data else None
        check(self.L.wdg_contact_list(self.h, w, iptr(ids), iptr(cd), dptr(d)))
        return ids, cd, d

    def restore_contacts(self, ids, cellDist, data, verletDist=-.05):
        """wdg_restore_contacts: a full collider refresh, then the listed contacts are real again with
        their saved state (the arrays of contacts(real_only=True))"""
        ids = np.ascontiguousarray(ids, dtype=np.int32).reshape(-1, 2)
        cd = np.ascontiguousarray(cellDist, dtype=np.int32).reshape(-1, 3) if cellDist is not None else None
        d = np.ascontiguousarray(data, dtype=np.float64).reshape(-1, 16)
        check(self.L.wdg_restore_contacts(self.h, len(ids), iptr(ids), iptr(cd) if cd is not None else None, dptr(d),
                                          float(verletDist)))

    def checkpoint(self):
        """The context's state as a dict of numpy arrays (for np.savez): the set-up (field,
        materials, particle shapes and flags, scenes), the node fields and the real contacts with
        their history. Device.from_checkpoint continues from it bit for bit."""
        import json
        if self._static is None:
            raise WdgError("checkpoint: no particles uploaded through this Device")
        ck = {"meta": np.array(json.dumps(dict(field=self._field, materials=self._mats)))}
        for k, v in self._static.items():
            ck["p_" + k] = v
        for f in ("pos", "vel", "angVel", "ori", "force", "torque", "mass", "inertia"):
            ck["p_" + f] = self.get(f)
        if self._scenes is not None:
            ck["scenes_dt"] = self._scenes
        ids, cd, data = self.contacts(real_only=True)
        ck.update(c_ids=ids, c_cellDist=cd, c_data=data)
        # Scene.trackEnergy and Scene.energy
        if self._scenes is not None and len(self._scenes) > 1:
            ck["energy_scenes"] = self.energy_scenes()
        else:
            ck["energy"] = np.array([self.energy()[k] for k in ENERGY_TERMS])
        ck["track_energy"] = np.array(self._tracking)
        return ck

    @classmethod
    def from_checkpoint(cls, ck, device=0, verletDist=-.05):
        """a new context holding the state of checkpoint() (a dict or an np.load result)"""
        import json
        meta = json.loads(str(ck["meta"]))
        d = cls(device)
        if meta["field"] is not None:
            d.set_field(**meta["field"])
        if meta["materials"] is not None:
            d.set_materials(meta["materials"])
        d.upload_particles(**{k[2:]: np.asarray(ck[k]) for k in ck if k.startswith("p_")})
        if "scenes_dt" in ck:
            d.set_scenes(np.asarray(ck["scenes_dt"]))
        d.restore_contacts(ck["c_ids"], ck["c_cellDist"], ck["c_data"], verletDist)
        # tracking first: switching it on for the first time clears the terms
        if "track_energy" in ck:
            d.set_energy_tracking(bool(ck["track_energy"]))
        if "energy" in ck:
            d.energy_set(np.asarray(ck["energy"]))
        if "energy_scenes" in ck:
            d.energy_scenes_set(np.asarray(ck["energy_scenes"]))
        return d

    def observables(self, lo, hi):
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        out = np.empty(3, dtype=np.float64)
        check(self.L.wdg_observables(self.h, dptr(lo), dptr(hi), dptr(out)))
        return dict(coordNumber=out[0], coordNumberSkipFree=out[1], solidFraction=out[2])

    def stress(self, volume=-1.):
        """Cauchy stress (3x3) from the real contacts; volume <= 0: the periodic cell's."""
        out = np.empty(9, dtype=np.float64)
        check(self.L.wdg_stress(self.h, volume, dptr(out)))
        return out.reshape(3, 3)

    def set_precision(self, bits):
        """64: plain FP64 engine; 32: contact geometry and law in FP32 (restarts contacts)."""
        check(self.L.wdg_set_precision(self.h, int(bits)))

    def kinetic_energy(self):
        e = C.c_double()
        check(self.L.wdg_kinetic_energy(self.h, C.byref(e)))
        return e.value

    def kinetic_energy_noalloc(self, dOut):
        """wdg_kinetic_energy into 8 bytes of device memory at address dOut, enqueued on the
        engine's stream without any device allocation (no wait)"""
        check(self.L.wdg_kinetic_energy_noalloc(self.h, C.c_void_p(dOut)))

    def set_energy_tracking(self, on=True):
        """Scene.trackEnergy: wdg_run adds the EnergyTracker terms while on"""
        check(self.L.wdg_set_energy_tracking(self.h, int(bool(on))))
        self._tracking = bool(on)

    def energy(self):
        """Scene.energy as a dict of the terms (ENERGY_TERMS)"""
        out = np.empty(7, dtype=np.float64)
        check(self.L.wdg_energy(self.h, dptr(out)))
        return dict(zip(ENERGY_TERMS, out.tolist()))

    def energy_scenes(self):
        """Scene.energy of each scene of a batched context of several scenes: an (nScenes, 7)
        array, columns in ENERGY_TERMS order (energy() is their sum over the scenes)"""
        ns = len(self._scenes) if self._scenes is not None else 0
        out = np.empty((max(ns, 1), 7), dtype=np.float64)
        check(self.L.wdg_energy_scenes(self.h, dptr(out)))
        return out

    def energy_scenes_set(self, terms):
        """EnergyTracker.set of each scene from an (nScenes, 7) array"""
        ns = len(self._scenes) if self._scenes is not None else 0
        a = np.ascontiguousarray(terms, dtype=np.float64).reshape(ns, 7)
        check(self.L.wdg_energy_scenes_set(self.h, dptr(a)))

    def energy_clear(self):
        check(self.L.wdg_energy_clear(self.h))

    def energy_set(self, terms):
        """the terms from a dict of ENERGY_TERMS (missing ones zero) or a sequence in that order"""
        if isinstance(terms, dict):
            terms = [terms.get(k, 0.) for k in ENERGY_TERMS]
        a = np.ascontiguousarray(terms, dtype=np.float64).reshape(7)
        check(self.L.wdg_energy_set(self.h, dptr(a)))

    def stats(self):
        s = Counters()
        check(self.L.wdg_stats(self.h, C.byref(s)))
        return {k: (list(getattr(s, k)) if k == "gridCells" else getattr(s, k)) for k, _ in Counters._fields_}

    def program(self, engines, dt, collide=None, contact=None, leapfrog=None, dynDtPeriod=100, maxRelInc=1e-4,
                dtSafety=.9):
        prog = Program()
        prog.nEngines = len(engines)
        for i, e in enumerate(engines):
            prog.engines[i] = ENG[e] if isinstance(e, str) else e
        prog.collide = CollideParams(**(collide or dict(verletDist=-.05, forceInit=0)))
        cp = dict(step=0, dt=dt, physKind=PHYS_FRICT, lawKind=LAW_IDEALELPL, applyForces=1, trsfRenorm=100,
                  poisson=.2, en=float("nan"), noSlip=0, noBreak=0, noFrict=0, noAttraction=1, relRollStiff=0.,
                  relTwistStiff=0., rollTanPhi=0.)
        cp.update(contact or {})
        prog.contact = ContactParams(**cp)
        lp = dict(dt=dt, damping=.2, reset=1)
        lp.update(leapfrog or {})
        prog.leapfrog = LeapfrogParams(**lp)
        prog.dynDtPeriod = dynDtPeriod
        prog.maxRelInc = maxRelInc
        prog.dtSafety = dtSafety
        return prog

    def run(self, engines, nSteps, step, dt, graph=False, **kw):
        """wdg_run (or wdg_run_graph with graph=True); kw: collide, contact, leapfrog, dynDtPeriod, ..."""
        prog = self.program(engines, dt, **kw)
        st = C.c_int64(step)
        d = C.c_double(dt)
        fn = self.L.wdg_run_graph if graph else self.L.wdg_run
        check(fn(self.h, C.byref(prog), nSteps, C.byref(st), C.byref(d)))
        return st.value, d.value

    def set_scenes(self, dt):
        """independent scenes, dt of each (empty: back to one scene)"""
        a = np.ascontiguousarray(dt, dtype=np.float64)
        check(self.L.wdg_set_scenes(self.h, len(a), dptr(a) if len(a) else None))
        self._scenes = a.copy()

    def set_eval_stress(self, on):
        """ContactLoop.evalStress: the contact loop's own stress in periodic cells"""
        check(self.L.wdg_set_eval_stress(self.h, int(bool(on))))

    def contact_loop_stress(self):
        """ContactLoop.stress of the last contact loop (3x3; zero when not evaluated)"""
        out = np.zeros(9)
        check(self.L.wdg_contact_loop_stress(self.h, dptr(out)))
        return out.reshape(3, 3)

    def scene_dt(self):
        """current dt of each batched scene"""
        out = np.empty(0 if self._scenes is None else len(self._scenes), dtype=np.float64)
        check(self.L.wdg_get_scene_dt(self.h, dptr(out) if len(out) else None))
        return out

    def crit_dt_scenes(self):
        """DynDt's critical timestep of each batched scene"""
        out = np.empty(0 if self._scenes is None else len(self._scenes), dtype=np.float64)
        check(self.L.wdg_crit_dt_scenes(self.h, dptr(out) if len(out) else None))
        return out

    def tune(self, engines, step, dt, **kw):
        """autotuner; returns (tuning dict, new step, steps taken)"""
        prog = self.program(engines, dt, **kw)
        st = C.c_int64(step)
        d = C.c_double(dt)
        t = Tuning()
        n = C.c_int64()
        check(self.L.wdg_tune(self.h, C.byref(prog), C.byref(st), C.byref(d), C.byref(t), C.byref(n)))
        return dict(cfBlocks=t.cfBlocks, graphSteps=t.graphSteps, skinSteps=t.skinSteps), st.value, n.value

    def tuning(self):
        t = Tuning()
        check(self.L.wdg_get_tuning(self.h, C.byref(t)))
        return dict(cfBlocks=t.cfBlocks, graphSteps=t.graphSteps, skinSteps=t.skinSteps)

    UPDATE_PHYS = ("never", "always", "once")

    @property
    def update_phys(self):
        """ContactLoop.updatePhys (ContactLoop.hpp:119): 'never', 'always' or 'once'."""
        m = C.c_int32()
        check(self.L.wdg_get_update_phys(self.h, C.byref(m)))
        return self.UPDATE_PHYS[m.value]

    @update_phys.setter
    def update_phys(self, mode):
        check(self.L.wdg_set_update_phys(self.h, self.UPDATE_PHYS.index(mode) if isinstance(mode, str) else int(mode)))

    def set_tuning(self, cfBlocks, graphSteps, skinSteps=0):
        t = Tuning(cfBlocks, graphSteps, skinSteps)
        check(self.L.wdg_set_tuning(self.h, C.byref(t)))

    def relax(self, maxIter=2000, inflate=0.01, tol=0.):
        ov = C.c_double()
        it = C.c_int32()
        check(self.L.wdg_relax(self.h, maxIter, inflate, tol, C.byref(ov), C.byref(it)))
        return ov.value, it.value

    def sync(self):
        check(self.L.wdg_sync(self.h))

    def snapshot(self, slot):
        """keep the kinematic state (pos, vel, angVel, ori, force, torque) and, in the FP64 contact
        form, the contact history on the device (slot 0..3)"""
        check(self.L.wdg_snapshot(self.h, slot))

    def restore(self, slot):
        """load a snapshot back: stepping on gives the trajectory the saved state would have had
        (the contact history restarts only in the FP32 form or for a snapshot without contacts)"""
        check(self.L.wdg_restore(self.h, slot))

    def snapshot_distance(self, slot):
        """largest distance of a sphere from its position in the snapshot"""
        d = C.c_double()
        check(self.L.wdg_snapshot_distance(self.h, slot, C.byref(d)))
        return d.value

    def stream(self):
        return self.L.wdg_stream(self.h)
