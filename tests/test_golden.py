"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle):
the oracle must keep reproducing them bit for bit (CPU), and the HIP engine must match them
(GPU) to the parity bars of the engine: identical real-contact sets, positions within 1e-9 r,
forces within 1e-12 of the force scale (observed: identical)."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import STEPS, scenes
from tests.scenes import GpuRunner, build_oracle, n_boundary

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["c1_deposition", "c1_dense", "hertz_periodic_216", "drum_facets"]


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False))


def _check(get, contacts, g, sc, step, exact):
    r = np.concatenate([sc["radius"], np.ones(n_boundary(sc))]) if not sc["walls_first"] else np.concatenate(
        [np.ones(n_boundary(sc)), sc["radius"]])
    ids = contacts()
    np.testing.assert_array_equal(ids, g[f"real_{step}"])
    for f in ("pos", "vel", "angVel", "force", "torque"):
        a, b = get(f), g[f"{f}_{step}"]
        if exact:
            np.testing.assert_array_equal(a, b, err_msg=f"{f} at step {step}")
        elif f == "pos":
            assert np.max(np.linalg.norm(a - b, axis=1) / r) <= 1e-9, (f, step)
        else:
            scale = max(np.max(np.abs(b)), 1e-300)
            assert np.max(np.abs(a - b)) <= 1e-12 * scale + 1e-300, (f, step, np.max(np.abs(a - b)) / scale)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    sc, engines = scenes()[name]
    g = load(name)
    assert list(g["engines"]) == engines
    o = build_oracle(sc)
    done = 0
    for s in STEPS:
        o.run(engines, s - done)
        done = s
        _check(o.get, lambda: o.contacts(real_only=True)[0].astype(np.int32), g, sc, s, exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(name):
    sc, engines = scenes()[name]
    g = load(name)
    run = GpuRunner(sc)
    done = 0
    for s in STEPS:
        run.run(engines, s - done)
        done = s
        _check(run.d.get, lambda: run.d.contacts(real_only=True, data=False)[0].astype(np.int32), g, sc, s,
               exact=False)
