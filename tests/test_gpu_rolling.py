"""Rolling and twisting resistance of Law2_L6Geom_FrictPhys_IdealElPl (IdealElPl.cpp:93-123) on
the HIP engine against the oracle.

The torque is this step's dt kr angVel (kr = kn charLen), capped at |Fn| rollTanPhi charLen, and
is applied with the force (Contact::getForceTorqueBranch, applyForceUninodal: xc x F + T). Kicked
dense packings in a box of walls and in a box of facets (whose node forces and torques, by
In2_Facet, take w_i T as well), stepped by the plain loop and by the graph replay. Bars as in the
other parity tests: identical real-contact sets, positions within 1e-9 r, angular velocities
within 1e-9 of their largest magnitude, facet node sums within 1e-12 relative.
"""
import numpy as np
import pytest

from tests.scenes import _particle_order, build_gpu, build_oracle, contact_kw, deposition, facet_box

pytestmark = pytest.mark.gpu

ROLL = dict(relRollStiff=1., relTwistStiff=.5, rollTanPhi=.1)


def _radii(sc):
    return np.array([sc["radius"][i] if k == "s" else 1. for k, i in _particle_order(sc)])


def _compare(o, d, sc, tag):
    io = o.contacts(real_only=True, data=False)[0]
    ig = d.contacts(real_only=True, data=False)[0]
    np.testing.assert_array_equal(io, ig, err_msg=tag)
    err = float(np.max(np.linalg.norm(d.get("pos") - o.get("pos"), axis=1) / _radii(sc)))
    wo, wg = o.get("angVel"), d.get("angVel")
    werr = float(np.max(np.abs(wg - wo))) / max(float(np.max(np.abs(wo))), 1e-300)
    print(f"{tag}: {len(io)} real contacts, position error {err:.3e}, angVel error {werr:.3e}", flush=True)
    assert err <= 1e-9, tag
    assert werr <= 1e-9, tag
    return len(io)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("kind", ["walls", "facets"])
def test_rolling_resistance_matches_oracle(kind, graph):
    if kind == "walls":
        sc = deposition(n=1500, seed=7, dense=True, vel_sigma=.05)
        eng = ["Collider", "ContactLoop", "Leapfrog"]
    else:
        sc = facet_box(n=800, seed=12, dense=True, vel_sigma=.05)
        eng = ["Collider", "ContactLoop", "IntraForce", "Leapfrog"]
    # angular velocities to resist
    rng = np.random.default_rng(3)
    sc["angVel"] = rng.normal(0, 20., size=(len(sc["radius"]), 3))
    o = build_oracle(sc, **ROLL)
    d = build_gpu(sc)
    order = _particle_order(sc)
    w = np.array([sc["angVel"][i] if k == "s" else np.zeros(3) for k, i in order])
    o.set("angVel", w)
    d.set("angVel", w)
    run_kw = dict(contact=dict(contact_kw(sc), **ROLL), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0,
                  graph=graph)
    step = 0
    for k in range(3):
        o.run(eng, 40)
        step, _ = d.run(eng, 40, step, sc["dt"], **run_kw)
        n = _compare(o, d, sc, f"{kind} graph={graph} step {step}")
    assert n > 0
    if kind == "facets":
        ids = [i for i, (kd, _) in enumerate(order) if kd == "f"]
        F, T = d.facet_nodes(ids)
        for j, i in enumerate(ids):
            fo, to = o.facet_nodes(i)
            scale = max(float(np.max(np.abs(fo))), 1e-300)
            np.testing.assert_allclose(F[j], fo, rtol=0, atol=1e-12 * scale)
            tscale = max(float(np.max(np.abs(to))), 1e-300)
            np.testing.assert_allclose(T[j], to, rtol=0, atol=1e-12 * tscale)
    d.close()


def test_rolling_resistance_changes_the_motion_and_needs_fp64():
    # the same scene without the resistance ends elsewhere; the FP32 form refuses it
    sc = deposition(n=500, seed=8, dense=True, vel_sigma=.05)
    rng = np.random.default_rng(4)
    w = np.vstack([rng.normal(0, 20., size=(len(sc["radius"]), 3)), np.zeros((len(sc["walls"]), 3))])
    ends = []
    for roll in (ROLL, {}):
        d = build_gpu(sc)
        d.set("angVel", w)
        d.run(["Collider", "ContactLoop", "Leapfrog"], 60, 0, sc["dt"], contact=dict(contact_kw(sc), **roll),
              leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
        ends.append(d.get("angVel"))
        d.close()
    assert np.max(np.abs(ends[0] - ends[1])) > 1e-3 * np.max(np.abs(ends[1]))
    d = build_gpu(sc)
    d.set_precision(32)
    with pytest.raises(Exception, match="FP64"):
        d.run(["Collider", "ContactLoop", "Leapfrog"], 2, 0, sc["dt"], contact=dict(contact_kw(sc), **ROLL),
              leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    d.close()


def test_snapshot_restore_with_rolling_resistance():
    # the phys torque has no history (it is this step's), so a restored snapshot steps on as the
    # first continuation did, bit for bit
    sc = deposition(n=600, seed=10, dense=True, vel_sigma=.05)
    rng = np.random.default_rng(5)
    w = np.vstack([rng.normal(0, 20., size=(len(sc["radius"]), 3)), np.zeros((len(sc["walls"]), 3))])
    d = build_gpu(sc)
    d.set("angVel", w)
    kw = dict(contact=dict(contact_kw(sc), **ROLL), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    step, _ = d.run(["Collider", "ContactLoop", "Leapfrog"], 30, 0, sc["dt"], **kw)
    d.snapshot(2)
    d.run(["Collider", "ContactLoop", "Leapfrog"], 30, step, sc["dt"], **kw)
    p1, w1 = d.get("pos"), d.get("angVel")
    d.restore(2)
    d.run(["Collider", "ContactLoop", "Leapfrog"], 30, step, sc["dt"], **kw)
    np.testing.assert_array_equal(d.get("pos"), p1)
    np.testing.assert_array_equal(d.get("angVel"), w1)
    d.close()
