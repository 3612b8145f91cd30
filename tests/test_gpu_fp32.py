"""FP32 contact form (contact geometry, physics and law in float; positions, the
position/velocity update and the force sums in FP64) against the plain FP64 engine, which
is bit-identical to the CPU oracle (test_gpu_parity). The bar is the one stated for the
precision form: macroscopic observables within 1 %; short trajectories close to FP64."""
import numpy as np
import pytest

from tests.scenes import GpuRunner, deposition, periodic_hertz

pytestmark = pytest.mark.gpu

ENG = ["Collider", "ContactLoop", "Leapfrog"]


def observables(g, sc):
    d = g.d
    box = sc["box"]
    if sc["periodic"]:
        sub = ((.25 * box,) * 3, (.75 * box,) * 3)
        vol = -1.
    else:
        sub = ((0.,) * 3, (box,) * 3)
        vol = box ** 3
    ob = d.observables(*sub)
    return dict(Ek=d.kinetic_energy(), Z=ob["coordNumber"], phi=ob["solidFraction"], stress=d.stress(vol))


def run_pair(sc, steps):
    g64 = GpuRunner(sc)
    g32 = GpuRunner(sc)
    g32.d.set_precision(32)
    g64.run(ENG, steps)
    g32.run(ENG, steps)
    return g64, g32


def radii(sc):
    return sc["radius"] if not sc["walls"] else np.concatenate([sc["radius"], np.ones(len(sc["walls"]))])


@pytest.mark.parametrize("which", ["hertz_periodic", "frict_deposition"])
def test_fp32_form_observables_within_one_percent(which):
    # the scenes are chaotic (collisional gas / settling), so single-particle trajectories of
    # the two forms separate; the trajectories are compared early and the observables as
    # averages over a window of the later run
    if which == "hertz_periodic":
        sc = periodic_hertz(n=4000, spacing=0.92, seed=51)
    else:
        sc = deposition(n=3000, dense=True, seed=52, vel_sigma=0.05)
    g64, g32 = run_pair(sc, 10)
    err = np.max(np.linalg.norm(g32.d.get("pos") - g64.d.get("pos"), axis=1) / radii(sc))
    print("max position deviation / r after 10 steps", err)
    # Until some contact crosses a branch of the law (Hertz: viscous shear term only below
    # the Coulomb cap) or of the damping (sign of f.v), the two forms differ by float
    # roundoff; a branch crossed on one side only changes that contact's force by O(eta v),
    # so the damped scene is checked loosely and later times only through observables.
    assert err < (1e-6 if sc["damping"] == 0 else 1e-3)
    acc = {k: [[], []] for k in ("Ek", "Z", "phi", "p")}
    g64.run(ENG, 20)
    g32.run(ENG, 20)
    for w in range(30):
        for g in (g64, g32):
            g.run(ENG, 10)
        for j, g in enumerate((g64, g32)):
            o = observables(g, sc)
            for k in ("Ek", "Z", "phi"):
                acc[k][j].append(o[k])
            acc["p"][j].append(np.trace(o["stress"]) / 3)
    for k, (a64, a32) in acc.items():
        m64, m32 = float(np.mean(a64)), float(np.mean(a32))
        print(which, k, m64, m32, "rel", abs(m32 - m64) / abs(m64))
        assert abs(m32 - m64) <= 0.01 * abs(m64), k


def test_fp32_contact_forces_close_to_fp64():
    # one contact loop from the same state: float geometry and law against FP64
    sc = periodic_hertz(n=1000, spacing=0.88, seed=53)
    g64, g32 = GpuRunner(sc), GpuRunner(sc)
    g32.d.set_precision(32)
    for g in (g64, g32):
        g.run(["Collider", "ContactLoop"])
    i64, _, d64 = g64.d.contacts(real_only=True)
    i32, _, d32 = g32.d.contacts(real_only=True)
    # contacts far from the touching threshold are the same set
    s64 = {tuple(x) for x, u in zip(i64.tolist(), d64[:, 12]) if u < -1e-6 * 2e-3}
    s32 = {tuple(x) for x in i32.tolist()}
    assert s64 <= s32
    f64, f32 = g64.d.get("force"), g32.d.get("force")
    scale = np.max(np.linalg.norm(f64, axis=1))
    rel = np.max(np.linalg.norm(f32 - f64, axis=1)) / scale
    print("force deviation / max force", rel, "contacts", len(i64), len(i32))
    assert rel < 1e-4


def test_precision_switch_restarts_contacts():
    sc = deposition(n=400, dense=True, seed=54, vel_sigma=0.05)
    g = GpuRunner(sc)
    g.run(ENG, 20)
    n64 = g.d.stats()["nReal"]
    g.d.set_precision(32)
    g.run(ENG, 20)
    g.d.set_precision(64)
    g.run(ENG, 20)
    st = g.d.stats()
    assert st["nReal"] > 0 and n64 > 0
    with pytest.raises(Exception):
        g.d.set_precision(16)


def test_fp32_drum_facet_contacts_close_to_fp64():
    # facet-sphere geometry in float on positions relative to the facet's vertex 0
    from tests.scenes import drum, n_boundary

    sc = drum(n=2000, seed=11)
    g64, g32 = GpuRunner(sc), GpuRunner(sc)
    g32.d.set_precision(32)
    for g in (g64, g32):
        g.run(ENG, 20)
    for g in (g64, g32):
        g.run(["Collider", "ContactLoop"])
    nb = n_boundary(sc)
    i64, _, d64 = g64.d.contacts(real_only=True)
    i32, _, d32 = g32.d.contacts(real_only=True)
    k64 = {tuple(x): j for j, x in enumerate(i64.tolist())}
    rows = [(k64[tuple(x)], j) for j, x in enumerate(i32.tolist()) if tuple(x) in k64 and 2 <= x[0] < nb]
    assert len(rows) > 20
    a = np.array([r[0] for r in rows])
    b = np.array([r[1] for r in rows])
    # normal, contact point and force of the facet contacts
    nerr = np.max(np.abs(d32[b, 0:3] - d64[a, 0:3]))
    perr = np.max(np.abs(d32[b, 6:9] - d64[a, 6:9]))
    fscale = np.max(np.abs(d64[a, 9]))
    ferr = np.max(np.abs(d32[b, 9:12] - d64[a, 9:12])) / fscale
    print("facet contacts", len(rows), "normal err", nerr, "contact point err", perr, "force err / max", ferr)
    assert nerr < 1e-4 and perr < 1e-6 and ferr < 1e-2
    # trajectories stay close over the short run
    dp = np.max(np.linalg.norm(g32.d.get("pos")[nb:] - g64.d.get("pos")[nb:], axis=1) / sc["radius"])
    print("max position deviation / r", dp)
    assert dp < 1e-3
