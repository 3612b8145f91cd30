"""ContactLoop.evalStress / ContactLoop.stress (ContactLoop.cpp:117,123,216-229,243-244) on the
HIP engine against the oracle: in a periodic cell each contact loop sums F branch^T over its
real contacts and divides by the cell volume. The sums run in another order than the oracle's,
so the bar is 1e-10 of the largest component; the oracle itself is pinned by DemFuncs' stress
(the symmetric part) and the contact torques (the antisymmetric part)."""
import numpy as np
import pytest

from tests.scenes import build_gpu, build_oracle, contact_kw, deposition, periodic_hertz

pytestmark = pytest.mark.gpu

ENG = ["Collider", "ContactLoop", "Leapfrog"]


@pytest.mark.parametrize("graph", [False, True])
def test_contact_loop_stress_matches_oracle(graph):
    sc = periodic_hertz(n=1200, spacing=0.92, seed=35, vel_sigma=1.)
    o = build_oracle(sc, evalStress=1)
    d = build_gpu(sc)
    d.set_eval_stress(True)
    kw = dict(contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0, graph=graph)
    step = 0
    for k in range(3):
        o.run(ENG, 30)
        step, _ = d.run(ENG, 30, step, sc["dt"], **kw)
        so, sg = o.cl_stress(), d.contact_loop_stress()
        scale = float(np.max(np.abs(so)))
        print(f"step {step}: max |stress| {scale:.4e}, difference {np.max(np.abs(sg - so)):.3e}", flush=True)
        assert scale > 0
        np.testing.assert_allclose(sg, so, rtol=0, atol=1e-10 * scale)
    # not symmetric: the tangential forces' moments are in it
    assert np.max(np.abs(sg - sg.T)) > 1e-8 * scale
    d.set_eval_stress(False)
    d.run(ENG, 1, step, sc["dt"], **kw)
    assert not d.contact_loop_stress().any()
    d.close()


def test_contact_loop_stress_is_zero_in_aperiodic_scenes():
    sc = deposition(n=400, seed=9, dense=True, vel_sigma=.05)
    d = build_gpu(sc)
    d.set_eval_stress(True)
    d.run(ENG, 5, 0, sc["dt"], contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    assert d.stats()["nReal"] > 0
    assert not d.contact_loop_stress().any()
    d.close()
