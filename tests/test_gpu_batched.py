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


def batch_run(scenes, steps, graph=False, tuning=None, energy=False):
    d, arr = build_batch(scenes)
    if energy:
        d.set_energy_tracking(True)
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
    assert t["cfBlocks"] in (1, 2, 4) and t["graphSteps"] in (4, 8, 16) and t["skinSteps"] in (0, 8, 16, 32)
    assert step == taken
    step, _ = d1.run(ENG, 100, step, sc["dt"], graph=True, **kw)
    d0, _ = build_batch(scenes)
    d0.run(ENG, step, 0, sc["dt"], **kw)
    np.testing.assert_array_equal(d1.get("pos"), d0.get("pos"))


def scenes_c4_sample(k=16, seed=71, nmin=500, nmax=5000):
    """C4-shaped scenes: N log-uniform in [nmin, nmax] (the smallest and largest forced in),
    E log-uniform in [1e6, 1e8] Pa, tanPhi uniform in [0.1, 0.9], each with its own box and
    dt = 0.5 pWaveDt; most start as a dense lattice (contacts from the first step), some as
    a loose random placement in a box of edge 0.2 (N / 2000)^(1/3)"""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=k)).astype(int)
    ns[0], ns[1] = nmin, nmax
    out = []
    for i, n in enumerate(ns):
        E = float(np.exp(rng.uniform(np.log(1e6), np.log(1e8))))
        tp = float(rng.uniform(.1, .9))
        if i % 4 == 3:
            sc = deposition(n=int(n), seed=seed + i, box=0.2 * (n / 2000.) ** (1 / 3.), young=E, tanPhi=tp,
                            vel_sigma=0.3)
        else:
            sc = deposition(n=int(n), seed=seed + i, dense=True, vel_sigma=0.08, young=E, tanPhi=tp)
        out.append(translated(sc, (0.3 * (i % 4), 0.3 * (i // 4), 0.)))
    return out


@pytest.mark.parametrize("skin", [0, 16])
def test_c4_batch_against_oracle_per_scene(skin):
    """16 C4-shaped scenes (500-5000 spheres, own E, tanPhi and dt) stepped together, graph
    replayed, through several refreshes; each scene against the oracle running it alone:
    real-contact sets equal, positions within 1e-9 r (Scene::doOneStep per scene,
    Scene.cpp:366-467). With per-scene skins (skinSteps 16) the candidate pairs differ from
    the oracle's, the real contacts and the trajectories do not. Energy is tracked: each
    scene's Scene.energy terms within 1e-9 of its largest term."""
    from tests.scenes import build_oracle
    from woo_amd._native import ENERGY_TERMS

    scenes = scenes_c4_sample()
    steps = 240
    d, _ = batch_run(scenes, steps, graph=True, tuning=dict(cfBlocks=2, graphSteps=8, skinSteps=skin), energy=True)
    E = d.energy_scenes()
    st = d.stats()
    print("batch: particles", st["nParticles"], "real", st["nReal"], "refreshes", st["nRefresh"], flush=True)
    assert st["nRefresh"] >= 3, "the window must hold at least two refreshes after the first"
    P = d.get("pos")
    ids = d.contacts(real_only=True, data=False)[0]
    total_real = 0
    for i, (sc, sl) in enumerate(zip(scenes, slices(scenes))):
        o = build_oracle(sc, trackEnergy=1)
        o.run(ENG, steps)
        eo = o.energy()
        ref = np.array([eo.get(k, 0.) for k in ENERGY_TERMS])
        np.testing.assert_allclose(E[i], ref, rtol=0, atol=1e-9 * float(np.max(np.abs(ref))),
                                   err_msg=f"scene {i}: energy")
        r = sc["radius"][0]
        err = float(np.max(np.abs(P[sl] - o.get("pos")))) / r
        mine = ids[(ids[:, 0] >= sl.start) & (ids[:, 0] < sl.stop)] - sl.start
        theirs = o.contacts(real_only=True, data=False)[0]
        total_real += len(theirs)
        print(f"scene {i}: N={len(sc['radius'])} E={sc['materials'][0]['young']:.3g} "
              f"dt={sc['dt']:.3g} real={len(theirs)} pos err/r={err:.2e}", flush=True)
        # both lists are in role order (pA, pB), sorted by (min id, max id)
        assert np.array_equal(mine, theirs), f"scene {i}: real contacts differ"
        assert err <= 1e-9, f"scene {i}: positions differ by {err} r"
    assert total_real > 10000


def test_graph_replay_follows_field_and_material_changes():
    # gravity is captured by value in a step graph and the material table by address: changing
    # either recaptures, so the replay stays the plain loop bit for bit
    sc = deposition(n=600, seed=66, dense=True, vel_sigma=0.05)
    kw = dict(contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    out = []
    for graph in (False, True):
        g = GpuRunner(sc)
        g.d.set_tuning(cfBlocks=2, graphSteps=8)
        step, _ = g.d.run(ENG, 40, 0, sc["dt"], graph=graph, **kw)
        g.d.set_field(gravity=(0., 3., -5.), periodic=False)
        step, _ = g.d.run(ENG, 40, step, sc["dt"], graph=graph, **kw)
        # 70 materials: the device table is reallocated; material 0 stays the scene's
        mats = [dict(sc["materials"][0], tanPhi=0.2)] + [sc["materials"][0]] * 69
        g.d.set_materials(mats)
        g.d.run(ENG, 40, step, sc["dt"], graph=graph, **kw)
        out.append({f: g.d.get(f) for f in ("pos", "vel", "angVel")})
    for f in out[0]:
        np.testing.assert_array_equal(out[1][f], out[0][f], err_msg=f)


def test_batched_dyndt_per_scene_equals_serial_runs():
    # DynDt in a batched context: each scene's DynDt::run (DynDt.cpp:90-113) on its own critical
    # timestep and dt, exactly as the scene run alone (the scenes' stiffnesses differ, so do their
    # timesteps)
    eng = ENG + ["DynDt"]
    scenes = scenes_c4_like(k=4, seed=66)
    for sc in scenes:
        sc["dt"] *= 0.3  # below the critical one, so that DynDt raises it step by step
    steps = 240
    kw = dict(leapfrog=dict(damping=scenes[0]["damping"]), dynDtPeriod=20, maxRelInc=1e-2, dtSafety=.7)
    d, _ = build_batch(scenes)
    d.run(eng, steps, 0, scenes[0]["dt"], contact=contact_kw(scenes[0]), **kw)
    dts = d.scene_dt()
    crit = d.crit_dt_scenes()
    P = d.get("pos")
    print("scene dt", dts, "critical", crit)
    for q, (sc, sl) in enumerate(zip(scenes, slices(scenes))):
        g = GpuRunner(sc)
        _, dt1 = g.d.run(eng, steps, 0, sc["dt"], contact=contact_kw(sc), **kw)
        assert dts[q] == dt1 and dt1 != sc["dt"]
        assert crit[q] == g.d.crit_dt()
        np.testing.assert_array_equal(P[sl], g.d.get("pos"))
    assert len(set(dts.tolist())) == len(scenes)


@pytest.mark.parametrize("graph", [False, True])
def test_batched_energy_per_scene_equals_serial_runs(graph):
    # EnergyTracker per scene of a batch (Scene.energy, EnergyTracker.hpp): each scene's terms as
    # when the scene is stepped alone. The additions go in another order (per particle, then per
    # scene, against per workgroup), so the terms agree to rounding: within 1e-11 of the scene's
    # largest term
    from woo_amd._native import ENERGY_TERMS
    scenes = scenes_c4_like(k=4, seed=67)
    steps = 160
    sc0 = scenes[0]
    kw = dict(contact=contact_kw(sc0), leapfrog=dict(damping=sc0["damping"]), dynDtPeriod=0)
    d, _ = build_batch(scenes)
    d.set_energy_tracking(True)
    d.run(ENG, steps, 0, sc0["dt"], graph=graph, **kw)
    E = d.energy_scenes()
    tot = d.energy()
    assert E.shape == (len(scenes), 7)
    for t, name in enumerate(ENERGY_TERMS):
        assert tot[name] == pytest.approx(float(np.sum(E[:, t])), rel=1e-14, abs=1e-300)
    P = d.get("pos")
    for q, (sc, sl) in enumerate(zip(scenes, slices(scenes))):
        g = GpuRunner(sc)
        g.d.set_energy_tracking(True)
        g.run(ENG, steps)
        np.testing.assert_array_equal(P[sl], g.d.get("pos"))
        e1 = g.d.energy()
        ref = np.array([e1[k] for k in ENERGY_TERMS])
        scale = float(np.max(np.abs(ref)))
        print(f"scene {q}: batched {E[q]} alone {ref}", flush=True)
        assert scale > 0
        np.testing.assert_allclose(E[q], ref, rtol=0, atol=1e-11 * scale)
        # the dissipation terms of the law and the damping are there
        for k in ("kinetic", "elast", "plast", "nonviscDamp"):
            assert e1[k] != 0, k
    # checkpoint of the batch keeps each scene's terms
    ck = d.checkpoint()
    d2 = type(d).from_checkpoint(ck)
    np.testing.assert_array_equal(d2.energy_scenes(), E)
    d2.close()
    d.close()


def test_batched_energy_set_clear_and_refusals():
    # the per-scene terms start at zero, can be set (a checkpoint) and cleared; the one-scene
    # setter is refused in a batched context
    scenes = scenes_c4_like(k=3, seed=68)
    d, _ = build_batch(scenes)
    d.set_energy_tracking(True)
    assert np.all(d.energy_scenes() == 0)
    with pytest.raises(Exception):
        d.energy_set([0.] * 7)
    d.energy_scenes_set(np.ones((3, 7)))
    np.testing.assert_array_equal(d.energy_scenes(), np.ones((3, 7)))
    sc = scenes[0]
    d.run(ENG, 5, 0, sc["dt"], contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    E = d.energy_scenes()
    assert np.all(E[:, 4] >= 1.)  # plast accumulates on the value set
    d.energy_clear()
    assert np.all(d.energy_scenes() == 0)
    d.close()


def test_batched_energy_with_the_autotuner():
    # the autotuner's trial replays capture the per-scene energy kernels as well; the terms stay
    # finite, every scene moves, and the total is the sum over the scenes
    from woo_amd._native import ENERGY_TERMS
    scenes = scenes_c4_like(k=3, seed=69)
    d, _ = build_batch(scenes)
    d.set_energy_tracking(True)
    sc = scenes[0]
    kw = dict(contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    tun, step, taken = d.tune(ENG, 0, sc["dt"], **kw)
    d.run(ENG, 24, step, sc["dt"], graph=True, **kw)
    E = d.energy_scenes()
    assert np.all(np.isfinite(E)) and np.all(E[:, 0] > 0)
    tot = d.energy()
    for t, name in enumerate(ENERGY_TERMS):
        assert tot[name] == pytest.approx(float(np.sum(E[:, t])), rel=1e-14, abs=1e-300)
    d.close()
