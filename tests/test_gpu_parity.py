"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Bars (integer/index work exact, FP64 physics per the engine's spec):
  * contact-pair sets (potential and real) identical, cellDist identical;
  * per-particle forces: max_i |F_i^gpu - F_i^cpu| / max(|F_i^cpu|, S_i) <= 1e-12, where
    S_i is the sum of contact-force magnitudes on particle i (guards equilibrium cancellation);
  * 100-step trajectories: max_i |x_i^gpu - x_i^cpu| / r_i <= 1e-9.
Both sides evaluate the same expressions in the same order without FMA, so the
observed differences are expected to be zero; the tests print them.
"""
import math

import numpy as np
import pytest

from tests.scenes import (GpuRunner, build_oracle, deposition, drum, facet_box, n_boundary, periodic_bidisperse,
                          smooth_facets,
                          periodic_hertz)

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-9


def _sets(ids):
    return set(map(tuple, ids.tolist()))


def compare_contact_sets(o, g, real_only):
    io, co, do = o.contacts(real_only=real_only)
    ig, cg, dg = g.d.contacts(real_only=real_only)
    assert io.shape == ig.shape, (len(_sets(io) - _sets(ig)), len(_sets(ig) - _sets(io)))
    np.testing.assert_array_equal(io, ig)
    np.testing.assert_array_equal(co, cg)
    return do, dg


def force_metric(o, g):
    Fo, Fg = o.get("force"), g.d.get("force")
    To, Tg = o.get("torque"), g.d.get("torque")
    ids, _, data = o.contacts(real_only=True)
    S = np.zeros(len(Fo))
    if len(ids):
        mag = np.linalg.norm(data[:, 9:12], axis=1)
        np.add.at(S, ids[:, 0], mag)
        np.add.at(S, ids[:, 1], mag)
    den = np.maximum(np.linalg.norm(Fo, axis=1), S)
    den[den == 0] = 1.
    ef = np.max(np.linalg.norm(Fg - Fo, axis=1) / den) if len(Fo) else 0.
    # torque scale: S_i times the largest radius
    rmax = 5e-3
    dent = np.maximum(np.linalg.norm(To, axis=1), S * rmax)
    dent[dent == 0] = 1.
    et = np.max(np.linalg.norm(Tg - To, axis=1) / dent) if len(To) else 0.
    return ef, et, int(np.sum(np.any(Fg != Fo, axis=1)))


def traj_metric(o, g, sc, nw_first):
    Po, Pg = o.get("pos"), g.d.get("pos")
    nb = n_boundary(sc)
    r = np.concatenate([np.ones(nb), sc["radius"]]) if nw_first else np.concatenate([sc["radius"], np.ones(nb)])
    return np.max(np.linalg.norm(Pg - Po, axis=1) / r), int(np.sum(np.any(Pg != Po, axis=1)))


ENG = ["Collider", "ContactLoop", "Leapfrog"]


@pytest.mark.parametrize("n,dense,seed", [(400, True, 1), (3000, False, 2), (1200, True, 3)])
def test_broadphase_pairs_aperiodic(n, dense, seed):
    sc = deposition(n=n, dense=dense, seed=seed)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider"])
    g.run(["Collider"])
    compare_contact_sets(o, g, real_only=False)
    # the potential set is exactly the brute-force enlarged-AABB overlap set
    assert _sets(o.brute_aabb_pairs()) == _sets(g.d.contacts(data=False)[0])


def test_sparse_scene_hashed_cells():
    # a few spheres far from the rest make the bounding box so large that a dense cell grid
    # would need far more cells than spheres: the cells are hashed into buckets instead; the
    # pair set, the contacts and the trajectories stay those of the oracle
    sc = deposition(n=2000, dense=True, seed=7)
    far = [(40., 0., 0.), (0., -25., 0.), (-30., 20., 0.), (12., 12., 0.), (0., 3., 0.)]
    for k, d in enumerate(far):
        sc["pos"][k] += np.array(d)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider"])
    g.run(["Collider"])
    compare_contact_sets(o, g, real_only=False)
    assert _sets(o.brute_aabb_pairs()) == _sets(g.d.contacts(data=False)[0])
    st = g.d.stats()
    cells = float(np.prod(st["gridCells"]))
    print("grid", st["gridCells"], "cells", cells, "spheres", len(sc["radius"]))
    assert cells > 8 * len(sc["radius"])
    o, g, worst = _run_both(sc, ENG, 100, check_every=20)
    print("hashed-cell run: worst position error", worst, "refreshes", g.d.stats()["nRefresh"])
    assert worst <= TRAJ_TOL


@pytest.mark.parametrize("n,spacing,seed", [(300, 1.0, 4), (2000, 1.01, 5)])
def test_broadphase_pairs_periodic(n, spacing, seed):
    sc = periodic_hertz(n=n, spacing=spacing, seed=seed)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider"])
    g.run(["Collider"])
    compare_contact_sets(o, g, real_only=False)
    st = g.d.stats()
    print("grid", st["gridCells"], "brute", st["brute"], "candidates", st["nCandidates"])


def test_walls_first_ids():
    sc = deposition(n=600, dense=True, seed=6, walls_first=True)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider", "ContactLoop"])
    g.run(["Collider", "ContactLoop"])
    compare_contact_sets(o, g, real_only=True)
    ef, et, ndiff = force_metric(o, g)
    print("walls-first force err", ef, et, ndiff)
    assert ef <= FORCE_TOL and et <= FORCE_TOL


def test_first_step_forces_frict():
    sc = deposition(n=1000, dense=True, seed=7, vel_sigma=0.02)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider", "ContactLoop"])
    g.run(["Collider", "ContactLoop"])
    do, dg = compare_contact_sets(o, g, real_only=True)
    assert len(do) > 500
    np.testing.assert_allclose(dg, do, rtol=1e-13, atol=0)
    ef, et, ndiff = force_metric(o, g)
    print("frict step-1 force err", ef, et, "particles differing:", ndiff)
    assert ef <= FORCE_TOL and et <= FORCE_TOL


def test_first_step_forces_hertz_periodic():
    sc = periodic_hertz(n=1000, spacing=0.88, seed=8)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider", "ContactLoop"])
    g.run(["Collider", "ContactLoop"])
    do, dg = compare_contact_sets(o, g, real_only=True)
    assert len(do) > 250
    np.testing.assert_allclose(dg, do, rtol=1e-13, atol=0)
    ef, et, ndiff = force_metric(o, g)
    print("hertz step-1 force err", ef, et, "particles differing:", ndiff)
    assert ef <= FORCE_TOL and et <= FORCE_TOL


def _run_both(sc, engines, steps, check_every=10):
    o = build_oracle(sc)
    g = GpuRunner(sc)
    worst = 0.
    for k in range(0, steps, check_every):
        o.run(engines, check_every)
        g.run(engines, check_every)
        compare_contact_sets(o, g, real_only=True)
        e, _ = traj_metric(o, g, sc, sc["walls_first"])
        worst = max(worst, e)
    return o, g, worst


def test_trajectory_100_frict_deposition():
    sc = deposition(n=800, dense=True, seed=9, vel_sigma=0.05)
    o, g, worst = _run_both(sc, ENG, 100)
    e, nd = traj_metric(o, g, sc, False)
    print("frict 100-step position err", worst, "particles differing:", nd, "refresh", o.nRefresh,
          g.d.stats()["nRefresh"])
    assert worst <= TRAJ_TOL
    np.testing.assert_allclose(g.d.get("vel"), o.get("vel"), rtol=0, atol=1e-9 * 0.05 + 1e-15)


def test_trajectory_100_hertz_periodic():
    sc = periodic_hertz(n=1500, spacing=0.92, seed=10)
    o, g, worst = _run_both(sc, ENG, 100)
    print("hertz 100-step position err", worst, "refresh", o.nRefresh, g.d.stats()["nRefresh"])
    assert worst <= TRAJ_TOL


def test_trajectory_high_degree_spheres():
    # large spheres among small ones: more neighbours per sphere than the broadphase stages,
    # so pair generation runs again in place and the adjacency takes the general path
    sc = periodic_bidisperse()
    big = np.flatnonzero(sc["radius"] > 2e-3)
    o, g, worst = _run_both(sc, ENG, 40)
    ids = o.contacts(real_only=False, data=False)[0]
    deg = max(int(np.sum(np.any(ids == b, axis=1))) for b in big)
    real = o.contacts(real_only=True, data=False)[0]
    print("spheres", len(sc["radius"]), "candidate pairs of the busiest large sphere", deg, "real contacts",
          len(real), "worst", worst)
    assert deg > 100
    assert worst <= TRAJ_TOL


def test_minimal_engine_order_with_forceresetter():
    # DemField.minimalEngines order (utils.py:122-129): Leapfrog first, forces gathered lazily
    sc = deposition(n=700, dense=True, seed=11, vel_sigma=0.05)
    eng = ["Leapfrog", "Collider", "ContactLoop"]
    o, g, worst = _run_both(sc, eng, 60)
    assert worst <= TRAJ_TOL
    eng2 = ["ForceResetter", "Collider", "ContactLoop", "Leapfrog"]
    sc2 = deposition(n=500, dense=True, seed=12, vel_sigma=0.05)
    o2 = build_oracle(sc2, reset=0)
    g2 = GpuRunner(sc2)
    o2.run(eng2, 40)
    for _ in range(40):
        g2.d.force_reset()
        g2.d.collide(-.05)
        g2.d.contact_loop(g2.step, g2.dt)
        g2.d.leapfrog(g2.dt, sc2["damping"], reset=False)
        g2.step += 1
    e, _ = traj_metric(o2, g2, sc2, False)
    assert e <= TRAJ_TOL


def test_force_state_changes_mid_run():
    # the integrator skips loading forces that hold the reset value; every change of what
    # that value is (gravity, mass) or of the forces themselves must be seen by the next step
    sc = deposition(n=600, dense=True, seed=21, vel_sigma=0.05)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(ENG, 15)
    g.run(ENG, 15)
    grav = (0.5, -0.25, -4.0)
    o.set_params(gravity=grav)
    g.d.set_field(gravity=grav, periodic=sc["periodic"], cell=sc["cell"])
    o.run(ENG, 15)
    g.run(ENG, 15)
    e, _ = traj_metric(o, g, sc, sc["walls_first"])
    assert e <= TRAJ_TOL
    rng = np.random.default_rng(5)
    m = o.get("mass")
    F = rng.normal(0., 1., size=(len(m), 3)) * (m * 9.81)[:, None]
    o.set("force", F)
    g.d.set("force", F)
    o.run(ENG, 10)
    g.run(ENG, 10)
    e, _ = traj_metric(o, g, sc, sc["walls_first"])
    assert e <= TRAJ_TOL
    o.set("mass", m * 1.25)
    g.d.set("mass", m * 1.25)
    o.run(ENG, 10)
    g.run(ENG, 10)
    compare_contact_sets(o, g, real_only=True)
    e, n_diff = traj_metric(o, g, sc, sc["walls_first"])
    print("positions differing", n_diff, "worst", e)
    assert e <= TRAJ_TOL
    # spheres sum in the canonical order: identical; the walls' thousands of contacts are
    # summed by a parallel tree, equal up to rounding
    ns = len(sc["radius"])
    np.testing.assert_array_equal(o.get("force")[:ns], g.d.get("force")[:ns])
    np.testing.assert_allclose(g.d.get("force")[ns:], o.get("force")[ns:], rtol=1e-12, atol=0)


def test_stress_and_kinetic_energy_match_oracle():
    # DemFuncs::stressStiffness and the kinetic energy on the same state; the GPU sums in
    # slot / block order, so agreement is to summation roundoff
    for sc, vol in ((periodic_hertz(n=1500, spacing=0.92, seed=41), -1.),
                    (deposition(n=800, dense=True, seed=42, vel_sigma=0.05), None)):
        o = build_oracle(sc)
        g = GpuRunner(sc)
        o.run(ENG, 30)
        g.run(ENG, 30)
        v = vol if vol is not None else float(sc["box"]) ** 3
        so, sg = o.stress(v), g.d.stress(v)
        scale = np.max(np.abs(so))
        print("stress", so.diagonal(), "max diff / max", np.max(np.abs(sg - so)) / scale)
        assert scale > 0
        np.testing.assert_allclose(sg, so, rtol=0, atol=1e-11 * scale)
        eo, eg = o.kinetic_energy(), g.d.kinetic_energy()
        assert eg == pytest.approx(eo, rel=1e-12)


def test_crit_dt_matches():
    for sc in (deposition(n=600, dense=True, seed=13), periodic_hertz(n=800, spacing=0.97, seed=14)):
        o = build_oracle(sc)
        g = GpuRunner(sc)
        o.run(ENG, 5)
        g.run(ENG, 5)
        co, cg = o.crit_dt(), g.d.crit_dt()
        print("critDt", co, cg)
        assert math.isfinite(co) and cg == pytest.approx(co, rel=1e-13)


def test_observables_match():
    sc = deposition(n=900, dense=True, seed=15)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider", "ContactLoop"])
    g.run(["Collider", "ContactLoop"])
    lo, hi = (0, 0, 0), (sc["box"],) * 3
    ob = g.d.observables(lo, hi)
    assert ob["coordNumber"] == pytest.approx(o.coord_number(False), rel=1e-14)
    assert ob["coordNumberSkipFree"] == pytest.approx(o.coord_number(True), rel=1e-14)
    assert ob["solidFraction"] == pytest.approx(o.solid_fraction(lo, hi), rel=1e-12)


def test_program_loop_matches_stepwise():
    sc = deposition(n=500, dense=True, seed=16, vel_sigma=0.05)
    g1 = GpuRunner(sc)
    g1.run(ENG, 30)
    g2 = GpuRunner(sc)
    step, dt = g2.d.run(ENG, 30, 0, sc["dt"], contact=dict(physKind=0, lawKind=0),
                        leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    assert step == 30
    np.testing.assert_array_equal(g1.d.get("pos"), g2.d.get("pos"))


def test_dyndt_program_matches_oracle():
    sc = deposition(n=500, dense=True, seed=17, vel_sigma=0.05)
    eng = ["Leapfrog", "Collider", "ContactLoop", "DynDt"]
    o = build_oracle(sc, dynDtPeriod=10, maxRelInc=1e-2, dtSafety=.9)
    o.run(eng, 35)
    g = GpuRunner(sc)
    step, dt = g.d.run(eng, 35, 0, sc["dt"], contact=dict(physKind=0, lawKind=0), leapfrog=dict(damping=.2),
                       dynDtPeriod=10, maxRelInc=1e-2, dtSafety=.9)
    print("dt oracle", o.dt, "gpu", dt)
    assert dt == pytest.approx(o.dt, rel=1e-13)
    e, _ = traj_metric(o, g, sc, False)
    assert e <= TRAJ_TOL


def test_periodic_double_image_raises():
    # PBC: a pair touching through two images is an error (test_pbc.py:151-160)
    sc = periodic_hertz(n=2, spacing=1.0, seed=18)
    sc["pos"] = np.array([[.3, .5, .5], [.8, .9, .5]])
    sc["radius"] = np.array([.4, .2])
    sc["vel"] = np.zeros((2, 3))
    sc["cell"] = (1., 1., 1.)
    g = GpuRunner(sc, verletDist=0.)
    with pytest.raises(RuntimeError):
        g.run(["Collider"])


def test_periodic_celldist_cases():
    # cellDist follows the shift of the second particle (test_pbc.py:162-172)
    for shift2 in [(0, 0, 0), (1, 2, 3), (6, 8, -4)]:
        sc = periodic_hertz(n=2, spacing=1.0, seed=19)
        sc["pos"] = np.array([[.5, .5, .5], [.8, .6, .5]]) + np.array([[0, 0, 0], shift2])
        sc["radius"] = np.array([.2, .2])
        sc["vel"] = np.zeros((2, 3))
        sc["cell"] = (1., 1., 1.)
        g = GpuRunner(sc, verletDist=0.)
        g.run(["Collider"])
        ids, cd, _ = g.d.contacts()
        assert ids.tolist() == [[0, 1]]
        assert cd.tolist() == [[-s for s in shift2]]


# ---------------------------------------------------------------- facets
def test_broadphase_pairs_drum_facets():
    # facet AABBs (vertices +- halfThick, Verlet-enlarged) against sphere AABBs: the potential
    # set equals the oracle's and the brute-force overlap set
    sc = drum(n=1500, seed=13, halfThick=2e-4)
    o = build_oracle(sc)
    g = GpuRunner(sc)
    o.run(["Collider"])
    g.run(["Collider"])
    compare_contact_sets(o, g, real_only=False)
    ids = g.d.contacts(data=False)[0]
    nb = n_boundary(sc)
    print("candidates", len(ids), "with a facet", int(np.sum((ids[:, 0] >= 2) & (ids[:, 0] < nb))))
    assert int(np.sum((ids[:, 0] >= 2) & (ids[:, 0] < nb))) > 100
    assert _sets(o.brute_aabb_pairs()) == _sets(ids)


@pytest.mark.parametrize("halfThick", [0., 3e-4])
def test_trajectory_drum_facets(halfThick):
    # rotating drum (fakeVel) with end walls, Hertz: sets, trajectories and forces as the oracle
    sc = drum(n=2000, seed=11, halfThick=halfThick)
    o, g, worst = _run_both(sc, ENG, 200, check_every=20)
    ids = o.contacts(real_only=True, data=False)[0]
    nb = n_boundary(sc)
    nf = int(np.sum((ids[:, 0] >= 2) & (ids[:, 0] < nb)))
    ef, et, nd = force_metric(o, g)
    print("drum real contacts", len(ids), "facet contacts", nf, "worst position err", worst, "force", ef, et)
    assert nf > 20
    assert worst <= TRAJ_TOL
    assert ef <= FORCE_TOL and et <= FORCE_TOL
    # facets take no force from the contact loop
    F = g.d.get("force")
    assert np.all(F[2:nb] == 0.)


def test_trajectory_drum_smoothed_facet_edges():
    # Facet.n21lim from Facet::computeNeighborAngles on the drum's mantle: the edge and vertex
    # contacts' normals held within the neighbours' (Facet::adjustBoundaryContactGeom), as the
    # oracle's
    sc = smooth_facets(drum(n=2000, seed=11))
    nb = n_boundary(sc)
    assert sum(int(np.isfinite(f["n21lim"]).sum()) for f in sc["facets"]) > 100
    o, g, worst = _run_both(sc, ENG, 200, check_every=20)
    ids = o.contacts(real_only=True, data=False)[0]
    nf = int(np.sum((ids[:, 0] >= 2) & (ids[:, 0] < nb)))
    print("smoothed drum: facet contacts", nf, "worst position err", worst)
    assert nf > 20
    assert worst <= TRAJ_TOL
    # the smoothing changes the motion: the same drum without n21lim ends elsewhere
    o2 = build_oracle(drum(n=2000, seed=11))
    o2.run(ENG, 200)
    assert np.max(np.abs(o2.get("pos") - o.get("pos"))) > 1e-12
    # the FP32 contact form smooths the same way: 10 steps stay close to the FP64 run
    g64, g32 = GpuRunner(sc), GpuRunner(sc)
    g32.d.set_precision(32)
    g64.run(ENG, 10)
    g32.run(ENG, 10)
    dev = np.max(np.linalg.norm(g32.d.get("pos") - g64.d.get("pos"), axis=1)) / float(np.min(sc["radius"]))
    assert dev < 1e-3


def _drum_torque(sc, first, F, T):
    """moment about the drum axis (x through the origin) of the facet node forces and torques"""
    M = 0.
    for k, f in enumerate(sc["facets"]):
        for i in range(3):
            M += np.cross(f["verts"][i], F[k, i])[0] + T[k, i, 0]
    return M


@pytest.mark.parametrize("graph", [False, True])
def test_drum_facet_node_forces_match_oracle(graph):
    # IntraForce (In2_Facet) in the rotating drum: the force and torque of every facet node, and
    # the drum torque they add up to, as the oracle's, stepped engine by engine and replayed from
    # captured graphs
    sc = drum(n=2000, seed=11)
    eng = ["Collider", "ContactLoop", "IntraForce", "Leapfrog"]
    o = build_oracle(sc)
    first, nb = len(sc["walls"]), n_boundary(sc)
    ids = np.arange(first, nb)
    if graph:
        from tests.scenes import build_gpu, contact_kw
        d = build_gpu(sc)
        step = 0
    else:
        g = GpuRunner(sc)
        d = g.d
    worst = 0.
    for k in range(4):
        o.run(eng, 50)
        if graph:
            step, _ = d.run(eng, 50, step, sc["dt"], graph=True, contact=contact_kw(sc),
                            leapfrog=dict(damping=sc["damping"]))
        else:
            g.run(eng, 50)
        Fg, Tg = d.facet_nodes(ids)
        Fo = np.array([o.facet_nodes(i)[0] for i in ids])
        To = np.array([o.facet_nodes(i)[1] for i in ids])
        scale = np.abs(Fo).max()
        err = max(np.abs(Fg - Fo).max(), np.abs(Tg - To).max() / 0.04) / scale  # torques over the drum radius
        worst = max(worst, err)
        Mo, Mg = _drum_torque(sc, first, Fo, To), _drum_torque(sc, first, Fg, Tg)
        print(f"step {50 * (k + 1)}: node force error {err:.3e} (of the largest), drum torque {Mg:.9e} vs {Mo:.9e}, "
              f"facets touched {int(np.sum(np.abs(Fo).sum((1, 2)) > 0))}", flush=True)
        assert Mg == pytest.approx(Mo, rel=1e-12, abs=1e-12 * scale * .04)
    assert worst <= FORCE_TOL
    with pytest.raises(RuntimeError):
        d.facet_nodes([nb])  # the first sphere


def test_facet_box_deposition_matches_oracle():
    sc = facet_box(n=800, dense=True, seed=9, vel_sigma=0.05)
    o, g, worst = _run_both(sc, ENG, 100)
    ids = o.contacts(real_only=True, data=False)[0]
    nf = int(np.sum(ids[:, 0] >= len(sc["radius"])))  # (pA, pB): the facet is pA
    print("facet box: real", len(ids), "facet contacts", nf, "worst", worst)
    assert nf > 10
    assert worst <= TRAJ_TOL
    ef, et, _ = force_metric(o, g)
    assert ef <= FORCE_TOL and et <= FORCE_TOL
