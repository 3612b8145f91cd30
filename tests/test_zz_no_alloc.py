"""wdg_kinetic_energy_noalloc (the kinetic energy without any device allocation) against
wdg_kinetic_energy through the C ABI: bit-identical on two calls in a row on one GPU, the
second on a smaller scene, so a stale partial left in the shared scratch array would show.
"""
import ctypes as C

import numpy as np
import pytest

from tests.scenes import GpuRunner, deposition

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), n) == 0
        return p

    def read_double(self, p):
        out = np.zeros(1, np.float64)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, 8, 2) == 0  # device to host
        return float(out[0])


def _scene(n, seed, steps):
    g = GpuRunner(deposition(n=n, seed=seed, vel_sigma=0.3))
    g.run(["Collider", "ContactLoop", "Leapfrog"], steps)
    return g


def test_noalloc_kinetic_energy_equals_allocating_entry_point():
    hip = _Hip()
    out = hip.malloc(8)
    try:
        for n, seed, steps in ((3000, 31, 30), (700, 32, 12)):
            g = _scene(n, seed, steps)
            ref = g.d.kinetic_energy()
            g.d.kinetic_energy_noalloc(out.value)
            g.d.sync()
            got = hip.read_double(out)
            print(f"n={n}: kinetic energy {ref!r} vs no-allocation {got!r}", flush=True)
            assert ref > 0
            assert got == ref, (n, ref, got)
            g.d.close()
    finally:
        hip.L.hipFree(out)
