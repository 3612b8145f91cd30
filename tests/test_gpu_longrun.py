"""Long runs of the HIP engine against the CPU oracle (same seeded inputs, same step count).

* the first workload (2,000 spheres deposited under gravity in a box of 6 walls, FrictMat,
  IdealElPl, damping 0.2, dt = 0.5 pWaveDt) for its full 2,000 steps;
* the benchmark workload (1,000,000 polydisperse Hertz spheres, periodic, phi = 0.55) for 200
  steps, which take the broadphase through several refreshes.

Parity bars as in the short tests: real-contact sets identical, positions within 1e-9 r; the
observables (kinetic energy, coordination number, stress) within 1e-10 relative. The engine
sums every particle's forces in the order of the oracle's deterministic mode, so the runs are
observed to stay identical bit for bit.
"""
import time

import numpy as np
import pytest

from tests.scenes import build_gpu, build_oracle, contact_kw, deposition

pytestmark = pytest.mark.gpu

ENG = ["Collider", "ContactLoop", "Leapfrog"]


def _compare(o, d, r, tag):
    io = o.contacts(real_only=True, data=False)[0]
    ig = d.contacts(real_only=True, data=False)[0]
    np.testing.assert_array_equal(io, ig, err_msg=tag)
    err = float(np.max(np.linalg.norm(d.get("pos") - o.get("pos"), axis=1) / r))
    ek_o, ek_g = o.kinetic_energy(), d.kinetic_energy()
    print(f"{tag}: {len(io)} real contacts, position error {err:.3e}, kinetic energy {ek_g:.9e} vs {ek_o:.9e}",
          flush=True)
    assert err <= 1e-9, tag
    assert abs(ek_g - ek_o) <= 1e-10 * abs(ek_o), tag
    return len(io)


def test_first_workload_full_2000_steps():
    sc = deposition(n=2000, seed=1)
    o = build_oracle(sc)
    d = build_gpu(sc)
    r = np.concatenate([sc["radius"], np.ones(6)])
    step = 0
    t0 = time.perf_counter()
    for k in range(4):
        o.run(ENG, 500)
        step, _ = d.run(ENG, 500, step, sc["dt"], contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]),
                        dynDtPeriod=0)
        n = _compare(o, d, r, f"step {step}")
    st = d.stats()
    print(f"2000 steps in {time.perf_counter() - t0:.1f} s, refreshes {st['nRefresh']}")
    # 2000 steps of 4.3e-5 s: the spheres have fallen about 4 cm and the first ones touch
    assert n > 0 and st["nRefresh"] > 10
    d.close()


def test_benchmark_workload_200_steps_through_refreshes():
    from oracle import Oracle
    from tests.test_gpu_fullsize import CP, EN, POISSON, RHO, TANPHI, YOUNG, gpu_device, workload

    w = workload(1_000_000)
    d = gpu_device(w)
    o = Oracle(dt=w["dt"], gravity=(0, 0, 0), periodic=1, cell=(w["L"],) * 3, physKind=1, lawKind=1, damping=0.,
               poisson=POISSON, en=EN, deterministic=1)
    o.material(kind=1, density=RHO, young=YOUNG, ktDivKn=.2, tanPhi=TANPHI)
    o.spheres(w["pos"], w["radius"], 0)
    o.set("vel", w["vel"])
    o.set("mass", w["mass"])
    o.set("inertia", w["inertia"])
    L = w["L"]
    step = 0
    t0 = time.perf_counter()
    for k in range(4):
        o.run(ENG, 50)
        step, _ = d.run(ENG, 50, step, w["dt"], contact=CP, leapfrog=dict(damping=0.), dynDtPeriod=0)
        _compare(o, d, w["radius"], f"step {step}")
        zo, zg = o.coord_number(False), d.observables((-1e9,) * 3, (1e9,) * 3)["coordNumber"]
        assert zg == pytest.approx(zo, rel=1e-10)
        so, sg = o.stress(L ** 3), d.stress(-1.)
        assert np.max(np.abs(sg - so)) <= 1e-10 * np.max(np.abs(so))
    st = d.stats()
    print(f"200 steps in {time.perf_counter() - t0:.1f} s, GPU refreshes {st['nRefresh']}, oracle {o.nRefresh}")
    assert st["nRefresh"] >= 3  # the first refresh and at least two more inside the run
    d.close()
