"""ctypes front end of the CPU oracle (test infrastructure only)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
liboracle_path = os.path.join(_HERE, "_build", "liboracle.so")

ENG = dict(ForceResetter=1, Collider=2, ContactLoop=3, Leapfrog=4, DynDt=5, IntraForce=6)
FIELD = dict(pos=0, vel=1, angVel=2, ori=3, force=4, torque=5, mass=6, inertia=7)
_FIELD_W = dict(pos=3, vel=3, angVel=3, ori=4, force=3, torque=3, mass=1, inertia=3)


class OrcParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("dtSafety", C.c_double), ("trsfRenorm", C.c_int), ("deterministic", C.c_int),
        ("nThreads", C.c_int),
        ("gravity", C.c_double * 3), ("distFactor", C.c_double), ("loneMask", C.c_uint32),
        ("periodic", C.c_int), ("cell", C.c_double * 3),
        ("verletDist", C.c_double), ("colliderMode", C.c_int),
        ("physKind", C.c_int), ("lawKind", C.c_int), ("applyForces", C.c_int),
        ("poisson", C.c_double), ("en", C.c_double),
        ("noSlip", C.c_int), ("noBreak", C.c_int), ("noFrict", C.c_int), ("noAttraction", C.c_int),
        ("damping", C.c_double), ("reset", C.c_int),
        ("dynDtPeriod", C.c_int), ("maxRelInc", C.c_double), ("trackEnergy", C.c_int),
        ("updatePhys", C.c_int),
        ("relRollStiff", C.c_double), ("relTwistStiff", C.c_double), ("rollTanPhi", C.c_double),
        ("evalStress", C.c_int),
    ]

# Scene.energy terms, in the order orc_energy and wdg_energy give them
ENERGY_TERMS = ("kinetic", "elast", "grav", "nonviscDamp", "plast", "viscN", "viscT")


def build_oracle():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def native_library():
    """The oracle compiled with -march=native for the host it runs on (the CPU baseline's build;
    the tests use the portable one): built by the Makefile's `native` target, a few seconds, under
    a name tagged with this host's CPU model and flags, so that a build from another host is never
    loaded here."""
    import hashlib

    ident = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith(("model name", "flags")):
                ident += line
            if line.strip() == "":
                break
    except OSError:
        pass
    tag = hashlib.sha1(ident.encode()).hexdigest()[:12]
    subprocess.check_call(["make", "-s", "-C", _HERE, "native", f"NATIVE_TAG={tag}"])
    return os.path.join(_HERE, "_build", f"liboracle_native_{tag}.so")


_libs = {}


def _load(path=None):
    path = path or liboracle_path
    if path in _libs:
        return _libs[path]
    if path == liboracle_path and not os.path.exists(liboracle_path):
        build_oracle()
    L = C.CDLL(path)
    P, D, I = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.orc_default_params.argtypes = [C.POINTER(OrcParams)]
    L.orc_create.argtypes = [C.POINTER(OrcParams)]; L.orc_create.restype = P
    L.orc_destroy.argtypes = [P]
    L.orc_set_params.argtypes = [P, C.POINTER(OrcParams)]
    L.orc_add_material.argtypes = [P, C.c_int] + [C.c_double] * 6; L.orc_add_material.restype = C.c_int
    L.orc_set_material.argtypes = [P, C.c_int, C.c_int] + [C.c_double] * 6; L.orc_set_material.restype = C.c_int
    L.orc_get_params.argtypes = [P, C.POINTER(OrcParams)]; L.orc_get_params.restype = C.c_int
    L.orc_add_sphere.argtypes = [P, D, C.c_double, C.c_int, C.c_int, C.c_uint32]; L.orc_add_sphere.restype = C.c_int
    L.orc_add_wall.argtypes = [P, D, C.c_int, C.c_int, C.c_int, C.c_uint32]; L.orc_add_wall.restype = C.c_int
    L.orc_add_facet.argtypes = [P, D, D, C.c_double, C.c_int, C.c_uint32]; L.orc_add_facet.restype = C.c_int
    L.orc_add_spheres.argtypes = [P, C.c_int64, D, D, C.c_int, C.c_uint32]; L.orc_add_spheres.restype = C.c_int
    L.orc_num_particles.argtypes = [P]; L.orc_num_particles.restype = C.c_int
    L.orc_get.argtypes = [P, C.c_int, D]; L.orc_get.restype = C.c_int
    L.orc_set.argtypes = [P, C.c_int, D]; L.orc_set.restype = C.c_int
    L.orc_run.argtypes = [P, I, C.c_int, C.c_int]; L.orc_run.restype = C.c_int
    L.orc_step.argtypes = [P]; L.orc_step.restype = C.c_int64
    L.orc_dt.argtypes = [P]; L.orc_dt.restype = C.c_double
    L.orc_verlet_dist.argtypes = [P]; L.orc_verlet_dist.restype = C.c_double
    L.orc_num_refresh.argtypes = [P]; L.orc_num_refresh.restype = C.c_int
    L.orc_num_contacts.argtypes = [P, C.c_int]; L.orc_num_contacts.restype = C.c_int64
    L.orc_get_contacts.argtypes = [P, C.c_int, I, I, D]; L.orc_get_contacts.restype = C.c_int64
    L.orc_coord_number.argtypes = [P, C.c_int]; L.orc_coord_number.restype = C.c_double
    L.orc_solid_fraction.argtypes = [P, D, D]; L.orc_solid_fraction.restype = C.c_double
    L.orc_kinetic_energy.argtypes = [P]; L.orc_kinetic_energy.restype = C.c_double
    L.orc_energy.argtypes = [P, C.POINTER(C.c_double)]
    L.orc_set_flags.argtypes = [P, C.c_int64, C.c_uint32]
    L.orc_stress.argtypes = [P, C.c_double, D]; L.orc_stress.restype = C.c_int
    L.orc_cl_stress.argtypes = [P, D]; L.orc_cl_stress.restype = C.c_int
    L.orc_set_facet_n21lim.argtypes = [P, C.c_int64, D]; L.orc_set_facet_n21lim.restype = C.c_int
    L.orc_facet_n21lim.argtypes = [P, C.c_int64, D]; L.orc_facet_n21lim.restype = C.c_int
    L.orc_facet_neighbor_angles.argtypes = [P]; L.orc_facet_neighbor_angles.restype = C.c_int
    L.orc_pwave_dt.argtypes = [P]; L.orc_pwave_dt.restype = C.c_double
    L.orc_crit_dt.argtypes = [P]; L.orc_crit_dt.restype = C.c_double
    L.orc_facet_nodes.argtypes = [P, C.c_int64, D]; L.orc_facet_nodes.restype = C.c_int
    L.orc_brute_aabb_pairs.argtypes = [P, I, C.c_int64]; L.orc_brute_aabb_pairs.restype = C.c_int64
    L.orc_last_error.restype = C.c_char_p
    _libs[path] = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Oracle:
    """One oracle scene. ``params`` are OrcParams field overrides."""

    def __init__(self, lib=None, **params):
        L = _load(lib)
        self.p = OrcParams()
        L.orc_default_params(C.byref(self.p))
        self.p.en = math.nan
        self.set_params(**params, _apply=False)
        self.h = L.orc_create(C.byref(self.p))
        self.L = L

    def set_params(self, _apply=True, **kw):
        if _apply:
            # from the scene's own values (DynDt changes dt, 'once' turns into 'never')
            self.L.orc_get_params(self.h, C.byref(self.p))
        for k, v in kw.items():
            if k in ("gravity", "cell"):
                getattr(self.p, k)[:] = list(v)
            else:
                setattr(self.p, k, v)
        if _apply:
            self.L.orc_set_params(self.h, C.byref(self.p))

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def material(self, kind=0, density=1e3, young=1e7, ktDivKn=.2, tanPhi=.5, surfEnergy=0., alpha=0.):
        return self.L.orc_add_material(self.h, kind, density, young, ktDivKn, tanPhi, surfEnergy, alpha)

    def set_material(self, m, kind=0, density=1e3, young=1e7, ktDivKn=.2, tanPhi=.5, surfEnergy=0., alpha=0.):
        """Material m edited in place (existing contacts keep their phys unless updatePhys)."""
        if self.L.orc_set_material(self.h, m, kind, density, young, ktDivKn, tanPhi, surfEnergy, alpha) != 0:
            raise IndexError(f"no material {m}")

    UPDATE_PHYS = ("never", "always", "once")

    @property
    def update_phys(self):
        """ContactLoop.updatePhys: 'never', 'always' or 'once' ('once' reads 'never' after a loop)."""
        p = OrcParams()
        self.L.orc_get_params(self.h, C.byref(p))
        self.p.updatePhys = p.updatePhys
        return self.UPDATE_PHYS[p.updatePhys]

    @update_phys.setter
    def update_phys(self, mode):
        self.set_params(updatePhys=self.UPDATE_PHYS.index(mode) if isinstance(mode, str) else int(mode))

    def sphere(self, pos, radius, mat=0, fixed=False, mask=0b0101):
        a = np.ascontiguousarray(pos, dtype=np.float64)
        return self.L.orc_add_sphere(self.h, _dp(a), float(radius), mat, int(bool(fixed)), mask)

    def spheres(self, pos, radius, mat=0, mask=0b0101):
        p = np.ascontiguousarray(pos, dtype=np.float64)
        r = np.ascontiguousarray(radius, dtype=np.float64)
        return self.L.orc_add_spheres(self.h, len(r), _dp(p), _dp(r), mat, mask)

    def wall(self, pos, axis, sense=0, mat=0, mask=0b0011):
        a = np.ascontiguousarray(pos, dtype=np.float64)
        return self.L.orc_add_wall(self.h, _dp(a), axis, sense, mat, mask)

    def facet(self, verts, fakeVel=None, halfThick=0., mat=0, mask=0b0011):
        v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(9))
        fv = None if fakeVel is None else np.ascontiguousarray(np.asarray(fakeVel, dtype=np.float64).reshape(3))
        return self.L.orc_add_facet(self.h, _dp(v), None if fv is None else _dp(fv), float(halfThick), mat, mask)

    def set_flags(self, i, flags):
        if self.L.orc_set_flags(self.h, int(i), int(flags)) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())

    @property
    def n(self):
        return self.L.orc_num_particles(self.h)

    def get(self, field):
        w = _FIELD_W[field]
        out = np.empty((self.n, w) if w > 1 else (self.n,), dtype=np.float64)
        self.L.orc_get(self.h, FIELD[field], _dp(out))
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.L.orc_set(self.h, FIELD[field], _dp(a))

    def run(self, engines, steps=1):
        e = np.array([ENG[x] if isinstance(x, str) else x for x in engines], dtype=np.int32)
        rc = self.L.orc_run(self.h, _ip(e), len(e), steps)
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())

    @property
    def step(self):
        return self.L.orc_step(self.h)

    @property
    def dt(self):
        return self.L.orc_dt(self.h)

    @property
    def verletDist(self):
        return self.L.orc_verlet_dist(self.h)

    @property
    def nRefresh(self):
        return self.L.orc_num_refresh(self.h)

    def contacts(self, real_only=False, data=True):
        """Return (ids[n,2], cellDist[n,3], data[n,16]) sorted by (min id, max id)."""
        w = 1 if real_only else 0
        n = self.L.orc_num_contacts(self.h, w)
        ids = np.empty((n, 2), dtype=np.int32)
        cd = np.empty((n, 3), dtype=np.int32)
        d = np.empty((n, 16), dtype=np.float64) if data else None
        self.L.orc_get_contacts(self.h, w, _ip(ids), _ip(cd), _dp(d) if data else None)
        return ids, cd, d

    def brute_aabb_pairs(self):
        n = self.L.orc_brute_aabb_pairs(self.h, None, 0)
        ids = np.empty((n, 2), dtype=np.int32)
        self.L.orc_brute_aabb_pairs(self.h, _ip(ids), n)
        return ids

    def facet_nodes(self, i):
        """Facet i's node forces (3x3, node by row) and torques (3x3) from the last IntraForce."""
        out = np.empty(18, dtype=np.float64)
        if self.L.orc_facet_nodes(self.h, int(i), _dp(out)) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return out[:9].reshape(3, 3), out[9:].reshape(3, 3)

    def coord_number(self, skipFree=False):
        return self.L.orc_coord_number(self.h, int(skipFree))

    def solid_fraction(self, lo, hi):
        a, b = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
        return self.L.orc_solid_fraction(self.h, _dp(a), _dp(b))

    def kinetic_energy(self):
        return self.L.orc_kinetic_energy(self.h)

    def energy(self):
        """Scene.energy as a dict (trackEnergy=1 in the parameters)"""
        out = (C.c_double * 7)()
        self.L.orc_energy(self.h, out)
        return dict(zip(ENERGY_TERMS, list(out)))

    def stress(self, volume=-1.):
        """Cauchy stress tensor (3x3) from the real contacts (DemFuncs::stressStiffness)."""
        out = np.zeros(9)
        if self.L.orc_stress(self.h, volume, _dp(out)) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return out.reshape(3, 3)

    def facet_neighbor_angles(self):
        """Facet::computeNeighborAngles for every facet (shared node = same vertex coordinates)."""
        self.L.orc_facet_neighbor_angles(self.h)

    def facet_n21lim(self, i):
        out = np.empty(3)
        if self.L.orc_facet_n21lim(self.h, int(i), _dp(out)) != 0:
            raise IndexError(f"particle {i} is not a facet")
        return out

    def set_facet_n21lim(self, i, v):
        a = np.ascontiguousarray(v, dtype=np.float64)
        if self.L.orc_set_facet_n21lim(self.h, int(i), _dp(a)) != 0:
            raise IndexError(f"particle {i} is not a facet")

    def cl_stress(self):
        """ContactLoop.stress (3x3) of the last contact loop (evalStress in a periodic cell)."""
        out = np.zeros(9)
        self.L.orc_cl_stress(self.h, _dp(out))
        return out.reshape(3, 3)

    def pwave_dt(self):
        return self.L.orc_pwave_dt(self.h)

    def crit_dt(self):
        return self.L.orc_crit_dt(self.h)
