"""The learned coarse step with its DEM fallback (woo_amd.surrogate.run_learned) on a small
granular column. Every check_every coarse steps the last coarse step is redone by DEM; when the
redo differs from the learned state by more than the bound, the whole segment is redone by DEM
from its start. So:
- with the bound at 0 every segment falls back, and the run is plain DEM bit for bit (the
  segment's start is restored with its contact history);
- with a bound that never fires, the state after each check is the DEM redo of the last coarse
  step from the learned state, which the test rebuilds by hand on a second context."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SETTLE = 200  # DEM steps before the first segment, so that contacts and their history exist


def _setup(seed=8):
    from woo_amd._native import Device
    from woo_amd.surrogate import Network, collapse_arrays, dem_dt, dem_run, upload

    arr, rad = collapse_arrays(width=0.03, height=0.06, seed=seed)
    dt = dem_dt()
    net = Network(seed=3)
    # a network whose output differs from free fall (the untrained decoder is zero)
    p = net.params()
    p += np.random.default_rng(5).normal(0., 0.02, size=p.shape).astype(np.float32)
    net.set_params(p)
    d = Device(0)
    upload(d, arr)
    step = dem_run(d, SETTLE, 0, dt)
    return d, net, arr, rad, dt, step


def test_forced_fallback_is_plain_dem():
    from woo_amd._native import Device
    from woo_amd.surrogate import COARSE, dem_run, run_learned, upload

    d1, net, arr, rad, dt, step = _setup()
    print("spheres", int(np.sum(arr["shape"] == 0)), "real contacts", d1.stats()["nReal"], flush=True)
    st = run_learned(d1, net, rad, coarse_steps=12, dt=dt, check_every=5, bound=0., step=step)
    print(st, flush=True)
    assert st["checks"] == 3 and st["fallbacks"] == 3
    assert st["step"] == step + 12 * COARSE
    assert min(st["errors"]) > 0.
    d0 = Device(0)
    upload(d0, arr)
    s0 = dem_run(d0, SETTLE + 12 * COARSE, 0, dt)
    assert s0 == st["step"]
    for f in ("pos", "vel", "angVel", "ori"):
        np.testing.assert_array_equal(d1.get(f), d0.get(f), err_msg=f)
    assert d1.stats()["nReal"] == d0.stats()["nReal"]


def test_check_state_is_dem_redo_of_last_coarse_step():
    from woo_amd.surrogate import COARSE, dem_run, run_learned

    d1, net, arr, rad, dt, step = _setup()
    d0, _, _, _, _, _ = _setup()
    seg, nseg = 4, 2
    st = run_learned(d1, net, rad, coarse_steps=seg * nseg, dt=dt, check_every=seg, bound=1e9, step=step)
    print(st, flush=True)
    assert st["checks"] == nseg and st["fallbacks"] == 0
    assert st["learned_coarse_steps"] == seg * nseg
    # by hand: seg - 1 learned coarse steps, then the last one by DEM, per segment
    skin = 0.25 * float(np.min(rad))
    s = step
    for _ in range(nseg):
        for _ in range(seg - 1):
            net.graph_from(d0, floorZ=0., maxDisp=skin)
            net.forward(out=False)
            net.advance(d0, COARSE * dt)
        dem_run(d0, COARSE, s + (seg - 1) * COARSE, dt)
        s += seg * COARSE
    assert st["step"] == s
    for f in ("pos", "vel", "angVel"):
        np.testing.assert_array_equal(d1.get(f), d0.get(f), err_msg=f)
    # the learned state differs from plain DEM (so the checks compared something)
    assert max(st["errors"]) > 0.
