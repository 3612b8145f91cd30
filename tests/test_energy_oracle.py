"""Pins of the oracle's energy tracking (Scene.energy, EnergyTracker) against closed forms.

The terms follow Leapfrog.cpp:53-98 (kinetic, grav, nonviscDamp), IdealElPl.cpp:62-92 and
Hertz.cpp:181-206 (elast, plast, viscN, viscT). Each check is a balance or a closed form that a
dropped term, a wrong sign or a value taken at the wrong half step fails.
"""
import math

import numpy as np
import pytest

from oracle import Oracle
from tests.scenes import build_oracle, deposition, periodic_hertz

ENG = ["Collider", "ContactLoop", "Leapfrog"]


def total(e):
    return sum(e.values())


@pytest.mark.parametrize("damping", [0., .2])
def test_free_fall_balance_is_exact(damping):
    # falling from rest under force m g: the kinetic energy at the on-step velocity, the gravity
    # work and the damping dissipation sum to a constant, step by step, to roundoff
    g, dt, r = -9.81, 1e-4, .01
    o = Oracle(damping=damping, gravity=(0, 0, g), dt=dt, trackEnergy=1)
    m = o.material()
    o.sphere((0, 0, 1.), r, m)
    o.run(["Leapfrog"], 1)  # gravity enters at the force reset: step 0 integrates with zero force
    o.set("vel", np.array([[0, 0, 0.]]))
    o.run(["Leapfrog"], 1)
    e0 = total(o.energy())
    mass = o.get("mass")[0]
    for _ in range(20):
        o.run(["Leapfrog"], 7)
        e = o.energy()
        assert total(e) == pytest.approx(e0, rel=1e-12, abs=1e-12 * mass * 9.81)
        # kinetic at the on-step velocity v(t) = v(t+dt/2) - a dt/2
        vz = o.get("vel")[0, 2]
        geff = g * (1 - damping)
        assert e["kinetic"] == pytest.approx(.5 * mass * (vz - .5 * dt * geff) ** 2, rel=1e-12)
        # gravity work: minus the potential energy released, m g dz over the steps tracked
        assert e["grav"] < 0
        if damping:
            # |v||F| d dt against -(g.v) m dt with F = m g: the ratio is the damping coefficient
            assert e["nonviscDamp"] == pytest.approx(-damping * e["grav"], rel=1e-12)
        else:
            assert e["nonviscDamp"] == 0.
    assert e["elast"] == e["plast"] == e["viscN"] == e["viscT"] == 0.


def test_rotational_kinetic_energy():
    dt = 1e-3
    w = np.array([0.3, -1.2, 2.0])
    o = Oracle(damping=0., gravity=(0, 0, 0), dt=dt, trackEnergy=1)
    m = o.material()
    o.sphere((0, 0, 0), .01, m)
    o.set("angVel", w[None, :])
    o.set("vel", np.array([[.1, 0, 0]]))
    o.run(["Leapfrog"], 5)
    mass, inertia = o.get("mass")[0], o.get("inertia")[0]
    e = o.energy()
    assert e["kinetic"] == pytest.approx(.5 * mass * .01 + .5 * float(inertia @ (w * w)), rel=1e-14)
    assert e["grav"] == 0. and e["nonviscDamp"] == 0.


def test_linear_bounce_conserves_kinetic_plus_elastic():
    # IdealElPl head-on bounce on a wall without damping: kinetic + elastic stays at the initial
    # kinetic energy through the contact (leapfrog, O(dt^2)), and nothing is dissipated
    r, rho, E, v0 = .01, 2000., 1e6, .1
    o = Oracle(damping=0., gravity=(0, 0, 0), trackEnergy=1)
    m = o.material(density=rho, young=E)
    o.wall((0, 0, 0), 2, 1, m)
    o.sphere((0, 0, r + 1e-6), r, m)
    o.set("vel", np.array([[0, 0, 0], [0, 0, -v0]]))
    o.set_params(dt=.05 * o.pwave_dt())
    k0 = .5 * o.get("mass")[1] * v0 * v0
    peak, started = 0., False
    for _ in range(100000):
        o.run(ENG, 1)
        e = o.energy()
        real = len(o.contacts(real_only=True, data=False)[0])
        started |= real > 0
        peak = max(peak, e["elast"])
        assert e["kinetic"] + e["elast"] == pytest.approx(k0, rel=5e-3)
        if started and not real:
            break
    assert peak == pytest.approx(k0, rel=5e-3)  # all the kinetic energy is stored at the turn
    assert e["kinetic"] == pytest.approx(k0, rel=1e-3) and e["elast"] == 0.
    assert e["plast"] == e["viscN"] == e["viscT"] == e["nonviscDamp"] == 0.


@pytest.mark.parametrize("en", [.3, .7])
def test_hertz_impact_balance(en):
    # Hertz head-on impact with viscous damping: the kinetic energy lost equals the normal viscous
    # dissipation, and the restitution coefficient follows from the energy left
    r, rho, E, nu, v0 = .05, 1e3, 1e8, .2, .5
    o = Oracle(physKind=1, lawKind=1, poisson=nu, en=en, damping=0., gravity=(0, 0, 0), trackEnergy=1)
    m = o.material(kind=1, density=rho, young=E, ktDivKn=.2, tanPhi=.5)
    o.sphere((0, 0, 0), r, m)
    o.sphere((0, 2 * r + 1e-5, 0), r, m)
    o.set("vel", np.array([[0, v0 / 2, 0], [0, -v0 / 2, 0]]))
    o.set_params(dt=.04 * o.pwave_dt())
    mass = o.get("mass")[0]
    k0 = 2 * .5 * mass * (v0 / 2) ** 2
    started = False
    for _ in range(200000):
        o.run(ENG, 1)
        real = len(o.contacts(real_only=True, data=False)[0])
        started |= real > 0
        e = o.energy()
        assert total(e) == pytest.approx(k0, rel=5e-3)  # first order in dt
        if started and not real:
            break
    v = o.get("vel")
    rest = (v[1, 1] - v[0, 1]) / v0
    assert e["viscN"] > 0 and e["plast"] < 1e-20 * k0  # head-on: shear only from roundoff
    assert e["kinetic"] == pytest.approx(k0 * rest * rest, rel=1e-6)
    assert e["viscN"] + e["viscT"] == pytest.approx(k0 * (1 - rest * rest), rel=5e-3)


def test_sliding_dissipates_two_sevenths():
    # a sphere sliding on a wall rolls at 5/7 of its speed: the kinetic energy drops to 5/7 of
    # its initial value, the rest is the slip dissipation (plast)
    r, rho, E, tanPhi, v0 = .01, 2500., 1e7, .4, .2
    o = Oracle(damping=0., gravity=(0, 0, -9.81), trackEnergy=1)
    m = o.material(density=rho, young=E, ktDivKn=.2, tanPhi=tanPhi)
    o.wall((0, 0, 0), 2, 1, m)
    mass = (4 / 3.) * math.pi * r ** 3 * rho
    kn = E * math.pi * r / 2
    o.sphere((0, 0, r - mass * 9.81 / kn), r, m)
    o.set("vel", np.array([[0, 0, 0], [v0, 0, 0]]))
    o.set_params(dt=.2 * o.pwave_dt())
    k0 = .5 * mass * v0 * v0
    o.run(ENG, int(3 * 2 / 7. * v0 / (tanPhi * 9.81) / o.dt))
    e = o.energy()
    assert e["plast"] == pytest.approx(2 / 7. * k0, rel=.02)
    assert e["kinetic"] == pytest.approx(5 / 7. * k0, rel=.02)
    assert total(e) - e["elast"] - e["grav"] == pytest.approx(k0, rel=.01)


def test_periodic_scene_is_not_tracked_at_step_zero():
    # Leapfrog.cpp:132: a periodic scene starts tracking at step 1
    sc = periodic_hertz(n=64, spacing=1.3, seed=3, vel_sigma=0.2, en=.6)
    o = build_oracle(sc, trackEnergy=1)
    o.run(ENG, 1)
    assert o.energy()["kinetic"] == 0.
    o.run(ENG, 1)
    assert o.energy()["kinetic"] == pytest.approx(o.kinetic_energy(), rel=1e-2)


def periodic_run(en, frac, chunks=10, chunk=100):
    sc = periodic_hertz(n=216, spacing=1.05, seed=5, vel_sigma=0.3, en=en)
    o = build_oracle(sc, trackEnergy=1)
    o.set_params(dt=o.dt * frac)
    o.run(ENG, 2)
    e0 = total(o.energy())
    err = 0.
    for _ in range(chunks):
        o.run(ENG, int(chunk / frac))
        err = max(err, abs(total(o.energy()) - e0) / e0)
    return err, {k: v / e0 for k, v in o.energy().items()}


def test_periodic_hertz_balance_converges():
    # elastic Hertz with friction, no gravity or damping: kinetic + elastic + slip dissipation stay
    # constant up to the integrator's error, which halves with the time step (a missing or
    # misplaced term would leave an error that does not shrink)
    err1, e1 = periodic_run(float("nan"), 1.)
    err2, e2 = periodic_run(float("nan"), .5)
    assert err1 < 3e-3 and err2 < .6 * err1
    assert e1["plast"] > 1e-2 and e1["viscN"] == e1["viscT"] == 0.
    assert e1["grav"] == 0. and e1["nonviscDamp"] == 0.


def test_periodic_hertz_viscous_terms():
    # with viscosity (en=.6) the kinetic energy and the normal viscous dissipation converge in dt,
    # while the slip dissipation does not: Hertz.cpp:87 maps Ft onto the stored shear force and
    # Hertz.cpp:198 adds eta*velT to it once per step, so the stored force grows as dt shrinks
    _, e1 = periodic_run(.6, 1., chunks=4)
    _, e2 = periodic_run(.6, .5, chunks=4)
    assert e1["viscN"] > 0 and e1["viscT"] > 0 and e1["plast"] > 0
    assert e2["kinetic"] == pytest.approx(e1["kinetic"], rel=5e-3)
    assert e2["viscN"] == pytest.approx(e1["viscN"], rel=2e-2)
    assert e2["plast"] > 1.5 * e1["plast"]


def test_deposition_balance():
    # spheres falling into a box with damping: the gravity work released is found as kinetic and
    # elastic energy and as dissipation
    sc = deposition(n=400, seed=6, vel_sigma=0.)
    o = build_oracle(sc, trackEnergy=1, damping=.2)
    o.run(ENG, 1)
    e0 = total(o.energy())
    for _ in range(8):
        o.run(ENG, 250)
    e = o.energy()
    assert -e["grav"] > 0 and e["nonviscDamp"] > 0 and e["plast"] > 0
    assert total(e) - e0 == pytest.approx(0., abs=2e-2 * -e["grav"])


@pytest.mark.parametrize("skip, zero, nonzero", [(512, "grav", "nonviscDamp"), (2048, "nonviscDamp", "grav")])
def test_skip_flags(skip, zero, nonzero):
    # TestLeapfrog.testGravitySkip / testDampingSkip (py/tests/test_energy.py:68-89): a sphere
    # dropped on a wall with damping .4; the skipped term is never added
    o = Oracle(damping=.4, gravity=(0, 0, -10), trackEnergy=1)
    m = o.material()
    o.wall((0, 0, 0), 2, 1, m)
    o.sphere((0, 0, .1), .08, m)
    o.set_flags(1, skip)
    o.set("vel", np.array([[0, 0, 0], [0, 0, -1.]]))
    o.set_params(dt=.5 * o.pwave_dt())
    o.run(ENG, 400)
    e = o.energy()
    assert e[zero] == 0. and e[nonzero] != 0.


def test_energy_skip_flag():
    # DemData.energySkip: the particle's kinetic energy and damping dissipation are left out
    o = Oracle(damping=.2, gravity=(0, 0, -9.81), dt=1e-4, trackEnergy=1)
    m = o.material()
    o.sphere((0, 0, 1.), .01, m)
    o.sphere((1, 0, 1.), .01, m)
    o.set_flags(1, 256)
    o.set("vel", np.array([[0, 0, -1.], [0, 0, -1.]]))
    o.run(["Leapfrog"], 10)
    one = Oracle(damping=.2, gravity=(0, 0, -9.81), dt=1e-4, trackEnergy=1)
    one.sphere((0, 0, 1.), .01, one.material())
    one.set("vel", np.array([[0, 0, -1.]]))
    one.run(["Leapfrog"], 10)
    e, e1 = o.energy(), one.energy()
    assert e["kinetic"] == e1["kinetic"] and e["nonviscDamp"] == e1["nonviscDamp"]
    assert e["grav"] == pytest.approx(2 * e1["grav"], rel=1e-14)


def test_tracking_off_leaves_zero_and_same_trajectory():
    sc = deposition(n=300, seed=7, vel_sigma=0.05)
    a = build_oracle(sc, trackEnergy=1, damping=.2)
    b = build_oracle(sc, damping=.2)
    a.run(ENG, 200)
    b.run(ENG, 200)
    np.testing.assert_array_equal(a.get("pos"), b.get("pos"))
    assert all(v == 0. for v in b.energy().values())
