"""Batched independent scenes (one context, per-scene dt and material, pairs only within a
scene) and the graph-replayed step loop with its autotuner. Each scene of a batch must
evolve exactly as the same scene run alone (same coordinates), and the graph runner exactly
as the plain engine loop: both are checked bit for bit."""
import numpy as np
import pytest

from tests.scenes import GpuRunner, build_batch, contact_kw, deposition, periodic_hertz, translated

pytestmark = pytest.mark.gpu

ENG = ["Collider", "ContactLoop", "Leapfrog"]


def scenes_c4_like(k=5, seed=61):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(150), np.log(900))))
        E = float(np.exp(rng.uniform(np.log(1e6), np.log(1e8))))
        tp = float(rng.uniform(.1, .9))
        sc = deposition(n=n, seed=seed + i, dense=True, vel_sigma=0.05, young=E, tanPhi=tp)
        out.append(translated(sc, (0.5 * i, 0.02 * i, -0.01 * i)))
    return out


def batch_run(scenes, steps, graph=False, tuning=None):
    d, arr = build_batch(scenes)
    if tuning:
        d.set_tuning(**tuning)
    sc = scenes[0]
    d.run(ENG, steps, 0, sc["dt"], graph=graph, contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]),
          dynDtPeriod=0)
    return d, arr


def slices(scenes):
    o, out = 0, []
    for sc in scenes:
        k = len(sc["radius"]) + len(sc["walls"])
        out.append(slice(o, o + k))
        o += k
    return out


def test_batched_scenes_equal_serial_runs():
    scenes = scenes_c4_like()
    steps = 150
    d, _ = batch_run(scenes, steps)
    P, V, F = d.get("pos"), d.get("vel"), d.get("force")
    st = d.stats()
    print("batched: particles", st["nParticles"], "real", st["nReal"], "refreshes", st["nRefresh"])
    assert st["nReal"] > 500
    for sc, sl in zip(scenes, slices(scenes)):
        g = GpuRunner(sc)
        g.run(ENG, steps)
        np.testing.assert_array_equal(P[sl], g.d.get("pos"))
        np.testing.assert_array_equal(V[sl], g.d.get("vel"))
        ns = len(sc["radius"])
        np.testing.assert_array_equal(F[sl][:ns], g.d.get("force")[:ns])


def test_scenes_do_not_interact():
    # two copies of one scene at the same place, as two scenes: each moves as the scene alone
    sc = deposition(n=400, seed=62, dense=True, vel_sigma=0.05)
    d, _ = batch_run([sc, sc], 60)
    g = GpuRunner(sc)
    g.run(ENG, 60)
    P = d.get("pos")
    k = len(sc["radius"]) + len(sc["walls"])
    np.testing.assert_array_equal(P[:k], g.d.get("pos"))
    np.testing.assert_array_equal(P[k:], g.d.get("pos"))


@pytest.mark.parametrize("tuning", [dict(cfBlocks=4, graphSteps=8), dict(cfBlocks=2, graphSteps=16),
                                    dict(cfBlocks=8, graphSteps=2)])
def test_graph_replay_equals_plain_loop_batched(tuning):
    scenes = scenes_c4_like(k=4, seed=63)
    steps = 300
    d0, _ = batch_run(scenes, steps)
    d1, _ = batch_run(scenes, steps, graph=True, tuning=tuning)
    s0, s1 = d0.stats(), d1.stats()
    print("refreshes plain", s0["nRefresh"], "graph", s1["nRefresh"], "real", s0["nReal"], s1["nReal"])
    for f in ("pos", "vel", "angVel", "ori"):
        np.testing.assert_array_equal(d1.get(f), d0.get(f))
    assert s0["nRefresh"] > 2 and s1["nRefresh"] == s0["nRefresh"]


def test_graph_replay_equals_plain_loop_periodic_hertz():
    sc = periodic_hertz(n=1500, spacing=0.92, seed=64)
    kw = dict(contact=contact_kw(sc), leapfrog=dict(damping=0.), dynDtPeriod=0)
    g0, g1 = GpuRunner(sc), GpuRunner(sc)
    g0.d.run(ENG, 250, 0, sc["dt"], **kw)
    g1.d.run(ENG, 250, 0, sc["dt"], graph=True, **kw)
    for f in ("pos", "vel", "angVel", "force"):
        np.testing.assert_array_equal(g1.d.get(f), g0.d.get(f))


def test_autotuner_picks_a_candidate_and_keeps_results_exact():
    scenes = scenes_c4_like(k=4, seed=65)
    sc = scenes[0]
    kw = dict(contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    d1, _ = build_batch(scenes)
    t, step, taken = d1.tune(ENG, 0, sc["dt"], **kw)
    print("tuned", t, "steps taken", taken)
    assert t["cfBlocks"] in (1, 2, 4) and t["graphSteps"] in (4, 8, 16)
    assert step == taken
    step, _ = d1.run(ENG, 100, step, sc["dt"], graph=True, **kw)
    d0, _ = build_batch(scenes)
    d0.run(ENG, step, 0, sc["dt"], **kw)
    np.testing.assert_array_equal(d1.get("pos"), d0.get("pos"))
