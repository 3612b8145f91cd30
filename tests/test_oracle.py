"""Pins of the CPU oracle against closed forms and the reference's own test cases.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign, a swapped operand) fails it; none of them retypes the
oracle's formula.
"""
import math

import numpy as np
import pytest

from oracle import Oracle
from tests.scenes import deposition, periodic_hertz, build_oracle

ENG = ["Collider", "ContactLoop", "Leapfrog"]


def two_spheres(v0, en=0., r=.05, rho=1e3, E=1e8, nu=.2, dtfrac=.04, gap=1e-5):
    """Hertz head-on impact (py/tests/test_hertz.py:13-47 setup, relative approach speed v0)."""
    o = Oracle(physKind=1, lawKind=1, poisson=nu, en=en, damping=0., gravity=(0, 0, 0))
    m = o.material(kind=1, density=rho, young=E, ktDivKn=.2, tanPhi=.5)
    o.sphere((0, 0, 0), r, m)
    o.sphere((0, 2 * r + gap, 0), r, m)
    o.set("vel", np.array([[0, v0 / 2, 0], [0, -v0 / 2, 0]]))
    o.set_params(dt=dtfrac * o.pwave_dt())
    return o


def collide_until_separated(o, max_steps=200000):
    n_in, started = 0, False
    for _ in range(max_steps):
        o.run(ENG, 1)
        real = len(o.contacts(real_only=True, data=False)[0])
        if real:
            started = True
            n_in += 1
        elif started:
            return n_in
    raise AssertionError("collision did not end")


@pytest.mark.parametrize("v0", [1., 1e-1, 1e-2, 1e-3])
def test_hertz_elastic_collision_time(v0):
    # TestHertz.testElasticCollisionTime: t = 2.214 (rho/E*)^(2/5) d v0^(-1/5) within 1 %
    o = two_spheres(v0)
    E, nu, r, rho = 1e8, .2, .05, 1e3
    Eeff = (E * E) / (E * (1 - nu ** 2) + E * (1 - nu ** 2))
    tau = 2.214 * (rho / Eeff) ** (2 / 5.) * ((2 * r) / v0 ** (1 / 5.))
    t = collide_until_separated(o) * o.dt
    assert t == pytest.approx(tau, rel=0.01)
    # elastic (en outside (0,1)): the velocities are exchanged
    v = o.get("vel")
    assert (v[1, 1] - v[0, 1]) / v0 == pytest.approx(1., rel=1e-4)
    # head-on: no tangential force arises (only quaternion-rotation roundoff of the normal force)
    assert abs(v[0, 0]) + abs(v[0, 2]) + abs(v[1, 0]) + abs(v[1, 2]) < 1e-14 * v0


@pytest.mark.parametrize("en", [.3, .6, .9])
def test_hertz_restitution(en):
    # Antypov & Elliott (2011) eq. (10) chooses the viscosity so that the restitution
    # coefficient equals en for the unclipped model (noAttraction=False)
    o = two_spheres(0.5, en=en)
    o.set_params(noAttraction=0)
    collide_until_separated(o)
    v = o.get("vel")
    e = (v[1, 1] - v[0, 1]) / 0.5
    assert e == pytest.approx(en, rel=2e-3)
    # with the attraction clip (the default) less energy is dissipated at the end of the impact
    o = two_spheres(0.5, en=en)
    collide_until_separated(o)
    v = o.get("vel")
    e2 = (v[1, 1] - v[0, 1]) / 0.5
    assert en < e2 < en + 0.15 * (1 - en)


def test_linear_contact_period_pins_kn():
    # elastic bounce on a wall (linear spring, no damping): contact lasts half a period pi*sqrt(m/kn);
    # for a wall contact lens=(r,r), contA=pi r^2, so kn = E*pi*r/2 (FrictMat.cpp:37)
    for E in (1e6, 4e6):
        r, rho = .01, 2000.
        o = Oracle(damping=0., gravity=(0, 0, 0))
        m = o.material(density=rho, young=E, ktDivKn=.2, tanPhi=.5)
        o.wall((0, 0, 0), 2, 1, m)
        o.sphere((0, 0, r + 1e-6), r, m)
        o.set("vel", np.array([[0, 0, 0], [0, 0, -0.1]]))
        o.set_params(dt=.01 * o.pwave_dt())
        steps = collide_until_separated(o)
        mass = o.get("mass")[1]
        kn = E * math.pi * r / 2
        assert steps * o.dt == pytest.approx(math.pi * math.sqrt(mass / kn), rel=2e-3)
        assert o.get("vel")[1, 2] == pytest.approx(0.1, rel=1e-6)


def _kn_of_contact(o):
    ids, _, data = o.contacts(real_only=True)
    assert len(ids) == 1
    return float(data[0, 13])


@pytest.mark.parametrize("up", ["never", "always", "once"])
def test_update_phys_never_always_once(up):
    # ContactLoop.updatePhys, the reference's case (py/tests/test_demfield.py:24-50): two
    # overlapping spheres, FrictMat young E0 -> E1 -> E0 between single steps; FrictPhys.kn is
    # proportional to young (FrictMat.cpp:37)
    E0, E1 = 1e3, 1e6
    mat = dict(kind=0, density=1e3, young=E0, ktDivKn=.2, tanPhi=.5)
    o = Oracle(dt=1e-8, damping=0., gravity=(0, 0, 0))
    m = o.material(**mat)
    o.sphere((0, 0, 0), .6, m)
    o.sphere((0, 0, 1), .6, m)
    o.run(ENG, 1)
    kn0 = _kn_of_contact(o)
    o.update_phys = up
    o.set_material(m, **dict(mat, young=E1))
    o.run(ENG, 1)
    kn1 = _kn_of_contact(o)
    if up in ("always", "once"):
        assert kn1 == pytest.approx(kn0 * E1 / E0, rel=1e-12)
    else:
        assert kn1 == kn0
    o.set_material(m, **mat)
    o.run(ENG, 1)
    if up == "always":
        assert o.update_phys == "always"
        assert _kn_of_contact(o) == pytest.approx(kn0, rel=1e-12)
    else:
        assert o.update_phys == "never"
        assert _kn_of_contact(o) == kn1


def _rolling_on_floor(relRollStiff, rollTanPhi, v0=.1, steps=3000, r=.01, rho=2000., E=1e7, dtf=.1):
    """A sphere rolling without slip on a floor wall under gravity (friction angle large enough),
    with rolling resistance; returns the horizontal speed sampled every step after a settling
    period, the time step and the contact's lens sum."""
    o = Oracle(damping=0., gravity=(0, 0, -9.81), relRollStiff=relRollStiff, rollTanPhi=rollTanPhi)
    m = o.material(density=rho, young=E, ktDivKn=.2, tanPhi=.8)
    o.wall((0, 0, 0), 2, 1, m)
    mass = 4 / 3. * math.pi * r ** 3 * rho
    kn = E * math.pi * r / 2
    o.sphere((0, 0, r - mass * 9.81 / kn), r, m)
    o.set("vel", np.array([[0, 0, 0], [v0, 0, 0]]))
    o.set("angVel", np.array([[0, 0, 0], [0, v0 / r, 0]]))
    o.set_params(dt=dtf * o.pwave_dt())
    o.run(ENG, 200)
    v = []
    for _ in range(steps):
        o.run(ENG, 1)
        v.append(o.get("vel")[1, 0])
    _, _, data = o.contacts(real_only=True)
    return np.array(v), o.dt, mass, kn


def test_rolling_resistance_at_its_cap_decelerates_at_five_sevenths_of_the_cap():
    # IdealElPl.cpp:93-123: a saturated rolling torque M = |Fn| rollTanPhi charLen opposes the
    # rotation; rolling without slip, (I + m r^2) alpha = -M with I = 2/5 m r^2 and Fn = m g, so
    # the speed falls at a = (5/7) g rollTanPhi charLen / r, charLen = 2 r (wall contact lens
    # r + uN/2 each, the overlap is 1e-5 r here)
    r, tp = .01, .05
    v, dt, mass, kn = _rolling_on_floor(relRollStiff=1., rollTanPhi=tp, r=r)
    a = np.polyfit(np.arange(len(v)) * dt, v, 1)[0]
    assert a == pytest.approx(-(5 / 7.) * 9.81 * tp * 2 * r / r, rel=5e-3)
    # without resistance it rolls on
    v0, _, _, _ = _rolling_on_floor(relRollStiff=0., rollTanPhi=tp, r=r, steps=500)
    assert abs(v0[-1] - v0[0]) < 1e-4 * v0[0]


def test_rolling_resistance_below_its_cap_damps_exponentially():
    # below the cap the torque is dt kr omega with kr = kn charLen (IdealElPl.cpp:97,109;
    # relRollStiff only switches it on); kn charLen = E A (FrictMat.cpp:37, lens l0 + l1 =
    # charLen) with A = pi r^2. Rolling without slip, (7/5) m r^2 domega/dt = -dt E pi r^2 omega,
    # so v decays at the rate dt E pi / (7/5 m). A large sphere and a small step keep that rate
    # times dt small (3e-4), where the discrete steps follow the exponential (the fit over one
    # second averages the tangential spring's oscillation; it comes within 1.4 % of the rate, 4 %
    # at four times the step)
    r, E = 1., 1e7
    v, dt, mass, kn = _rolling_on_floor(relRollStiff=1., rollTanPhi=1.5, v0=.1, r=r, E=E, steps=2830, dtf=.025)
    t = np.arange(len(v)) * dt
    rate = -np.polyfit(t, np.log(v), 1)[0]
    expect = dt * E * math.pi / (1.4 * mass)
    assert rate * dt < 1e-3
    assert rate == pytest.approx(expect, rel=2e-2)


def test_sliding_sphere_rolls_at_five_sevenths():
    # a sphere sliding on a wall ends up rolling at 5/7 of its initial speed (solid sphere),
    # independent of the friction coefficient; checks IdealElPl slip, torque branch and rotation
    for tanPhi in (.3, .6):
        r, rho, E = .01, 2500., 1e7
        o = Oracle(damping=0., gravity=(0, 0, -9.81))
        m = o.material(density=rho, young=E, ktDivKn=.2, tanPhi=tanPhi)
        o.wall((0, 0, 0), 2, 1, m)
        mass = (4 / 3.) * math.pi * r ** 3 * rho
        kn = E * math.pi * r / 2
        o.sphere((0, 0, r - mass * 9.81 / kn), r, m)
        v0 = 0.2
        o.set("vel", np.array([[0, 0, 0], [v0, 0, 0]]))
        o.set_params(dt=.2 * o.pwave_dt())
        t_slip = 2 / 7. * v0 / (tanPhi * 9.81)
        nsteps = int(3 * t_slip / o.dt)
        o.run(ENG, nsteps)
        v = o.get("vel")[1]
        w = o.get("angVel")[1]
        assert v[0] == pytest.approx(5 / 7. * v0, rel=0.01)
        # rolling without slip on a floor: contact-point velocity v_x - w_y r = 0
        assert w[1] * r == pytest.approx(v[0], rel=0.01)


def test_coulomb_cap_is_exact_while_sliding():
    r, rho, E, tanPhi = .01, 2500., 1e7, .4
    o = Oracle(damping=0., gravity=(0, 0, -9.81))
    m = o.material(density=rho, young=E, ktDivKn=.2, tanPhi=tanPhi)
    o.wall((0, 0, 0), 2, 1, m)
    mass = (4 / 3.) * math.pi * r ** 3 * rho
    o.sphere((0, 0, r - mass * 9.81 / (E * math.pi * r / 2)), r, m)
    o.set("vel", np.array([[0, 0, 0], [0.5, 0, 0]]))
    o.set_params(dt=.2 * o.pwave_dt())
    o.run(ENG, 50)
    ids, cd, d = o.contacts(real_only=True)
    assert ids.tolist() == [[0, 1]]
    Fn, Ft = d[0, 9], np.hypot(d[0, 10], d[0, 11])
    assert Ft == pytest.approx(abs(Fn) * tanPhi, rel=1e-14)


def test_free_fall_leapfrog_closed_form():
    # gravity enters at the force reset, so step 0 integrates with zero force
    g, dt, n, v0 = -9.81, 1e-4, 137, 0.3
    for damping in (0., .2):
        o = Oracle(damping=damping, gravity=(0, 0, g), dt=dt)
        m = o.material()
        o.sphere((0, 0, 1.), .01, m)
        o.set("vel", np.array([[0, 0, v0]]))
        o.run(["Leapfrog"], 1)
        o.set("vel", np.array([[0, 0, 0.]]))  # from rest, force = m g now
        z0 = o.get("pos")[0, 2]
        o.run(["Leapfrog"], n)
        # non-viscous damping opposes the motion: a *= (1 - damping) while falling
        geff = g * (1 - damping)
        vz = o.get("vel")[0, 2]
        z = o.get("pos")[0, 2]
        assert vz == pytest.approx(n * geff * dt, rel=1e-12)
        assert z - z0 == pytest.approx(geff * dt * dt * n * (n + 1) / 2, rel=1e-10)


def test_spherical_rotation_constant_angvel():
    dt, n = 1e-3, 250
    w = np.array([0.3, -1.2, 2.0])
    o = Oracle(damping=0., gravity=(0, 0, 0), dt=dt)
    m = o.material()
    o.sphere((0, 0, 0), .01, m)
    o.set("angVel", w[None, :])
    o.run(["Leapfrog"], n)
    q = o.get("ori")[0]
    th = np.linalg.norm(w) * dt * n
    ax = w / np.linalg.norm(w)
    qe = np.concatenate([[math.cos(th / 2)], math.sin(th / 2) * ax])
    assert np.max(np.abs(q - qe)) < 1e-12


def test_dyndt_single_wall_contact():
    # one vertical wall contact: kt=kn/5 so the critical step is sqrt(2 m / kn) (DynDt.cpp:40-58)
    r, rho, E = .01, 2000., 1e7
    o = Oracle(damping=0., gravity=(0, 0, 0))
    mt = o.material(density=rho, young=E, ktDivKn=.2, tanPhi=.5)
    o.wall((0, 0, 0), 2, 1, mt)
    o.sphere((0, 0, r * (1 - 1e-4)), r, mt)
    o.set_params(dt=.1 * o.pwave_dt())
    o.run(["Collider", "ContactLoop"], 1)
    mass = o.get("mass")[1]
    kn = E * math.pi * r / 2
    assert o.crit_dt() == pytest.approx(math.sqrt(2 * mass / kn), rel=1e-13)


def test_lattice_observables():
    # simple cubic lattice of touching spheres in a periodic cell: Z = 6, solid fraction -> pi/6
    m, r = 6, .01
    a = 2 * r * (1 - 1e-7)
    L = m * a
    o = Oracle(periodic=1, cell=(L, L, L), physKind=0, lawKind=0, damping=0., gravity=(0, 0, 0), verletDist=0.)
    mt = o.material(density=1000., young=1e7)
    for i in range(m):
        for j in range(m):
            for k in range(m):
                o.sphere(((i + .5) * a, (j + .5) * a, (k + .5) * a), r, mt)
    o.set_params(dt=.1 * o.pwave_dt())
    o.run(["Collider", "ContactLoop"], 1)
    assert o.coord_number(False) == 6.
    assert o.coord_number(True) == 6.
    assert len(o.contacts(real_only=True, data=False)[0]) == 3 * m ** 3
    assert o.solid_fraction((0, 0, 0), (L, L, L)) == pytest.approx(math.pi / 6, rel=1e-6)


def test_stress_compressed_lattice_closed_form():
    # simple cubic lattice, pitch a < 2r, one linear contact per sphere and axis:
    # sigma = diag(kn uN / a^2) with kn = E pi r / 2 and uN = a - 2r (compression negative)
    m, r, E = 5, .01, 1e7
    a = 2 * r * (1 - 1e-4)
    L = m * a
    o = Oracle(periodic=1, cell=(L, L, L), physKind=0, lawKind=0, damping=0., gravity=(0, 0, 0), verletDist=0.)
    mt = o.material(density=1000., young=E)
    for i in range(m):
        for j in range(m):
            for k in range(m):
                o.sphere(((i + .5) * a, (j + .5) * a, (k + .5) * a), r, mt)
    o.set_params(dt=.1 * o.pwave_dt())
    o.run(["Collider", "ContactLoop"], 1)
    kn = E * math.pi * r / 2
    sig = o.stress()
    expect = kn * (a - 2 * r) / a ** 2
    np.testing.assert_allclose(np.diag(sig), [expect] * 3, rtol=1e-9)
    assert np.max(np.abs(sig - np.diag(np.diag(sig)))) < 1e-9 * abs(expect)
    with pytest.raises(RuntimeError):
        Oracle().stress()  # aperiodic needs a volume


def test_stress_tangential_term_symmetric():
    # two spheres pressed along x and sheared along y: the tangential force enters as
    # (fT_i n_j + fT_j n_i)/2; compare with the tensor rebuilt from the contact's frame
    r, L = .01, .1
    o = Oracle(periodic=1, cell=(L, L, L), physKind=0, lawKind=0, damping=0., gravity=(0, 0, 0), verletDist=0.)
    mt = o.material(density=1000., young=1e7, tanPhi=.8)
    o.sphere((.05, .05, .05), r, mt)
    o.sphere((.05 + 2 * r * (1 - 1e-3), .05, .05), r, mt)
    o.set("vel", np.array([[0, -.01, 0], [0, .01, 0]]))
    o.set_params(dt=.05 * o.pwave_dt())
    o.run(ENG, 20)
    ids, _, d = o.contacts(real_only=True)
    assert len(ids) == 1
    t0, t1 = d[0, 0:3], d[0, 3:6]
    t2 = np.cross(t0, t1)
    F = d[0, 9] * t0 + d[0, 10] * t1 + d[0, 11] * t2
    n = t0
    fN = F @ n
    fT = F - n * fN
    p = o.get("pos")
    d0 = np.linalg.norm(p[1] - p[0])
    expect = d0 * (fN * np.outer(n, n) + .5 * (np.outer(fT, n) + np.outer(n, fT))) / L ** 3
    sig = o.stress()
    assert abs(fT[1]) > 1e-3 * abs(fN)
    np.testing.assert_allclose(sig, expect, rtol=1e-9, atol=1e-12 * np.max(np.abs(expect)))
    np.testing.assert_allclose(sig, sig.T, rtol=0, atol=1e-15 * np.max(np.abs(sig)))


def test_contact_loop_stress_parts():
    # ContactLoop.stress (evalStress, ContactLoop.cpp:216-229,243-244) is sum F branch^T / V at the
    # contact loop's positions. Its symmetric part is DemFuncs' stress (branch = d0 n for spheres);
    # its antisymmetric part is the torque the contacts put on the particles: sum F x b =
    # -sum b x F, and b x F = d0 n x F = (l0 + l1) n x F is the torque on pA and pB together
    # (contact point at l0 from A, l1 from B). After a reset and a contact loop alone, the
    # particles' torques are those contact torques and their positions the loop's. DemFuncs takes
    # the normal from node.ori, the quaternion of the contact's frame, whose axes are kept
    # orthonormal only every trsfRenorm steps: the two agree to about 1e-9, not to roundoff.
    sc = periodic_hertz(n=600, spacing=0.92, seed=34, vel_sigma=1.)
    o = build_oracle(sc, evalStress=1)
    o.run(ENG, 30)
    o.run(["ForceResetter", "Collider", "ContactLoop"], 1)
    S = o.cl_stress()
    V = float(np.prod(sc["cell"]))
    sym = .5 * (S + S.T)
    sig = o.stress()
    scale = float(np.max(np.abs(sig)))
    assert scale > 0
    np.testing.assert_allclose(sym, sig, rtol=0, atol=1e-7 * scale)
    w = V * np.array([S[1, 2] - S[2, 1], S[2, 0] - S[0, 2], S[0, 1] - S[1, 0]])
    tsum = o.get("torque").sum(axis=0)
    assert np.max(np.abs(tsum)) > 1e-6 * scale * V / sc["cell"][0]
    np.testing.assert_allclose(w, -tsum, rtol=0, atol=1e-10 * np.max(np.abs(tsum)))
    # off (the default) or in an aperiodic scene it is zero
    o2 = build_oracle(sc)
    o2.run(ENG, 2)
    assert not o2.cl_stress().any()


@pytest.mark.parametrize("seed", [1, 2])
def test_sap_incremental_equals_full_rebuild_and_bruteforce(seed):
    sc = deposition(n=1200, dense=True, seed=seed, vel_sigma=0.3)
    a = build_oracle(sc, colliderMode=0)
    b = build_oracle(sc, colliderMode=1)
    for k in range(6):
        a.run(ENG, 15)
        b.run(ENG, 15)
        ia, _, _ = a.contacts(data=False)
        ib, _, _ = b.contacts(data=False)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(a.get("pos"), b.get("pos"))
    assert a.nRefresh > 1
    # right after a refresh the potential set is the brute-force enlarged-AABB set plus real contacts
    real = set(map(tuple, a.contacts(real_only=True, data=False)[0].tolist()))
    brute = set(map(tuple, a.brute_aabb_pairs().tolist()))
    pot = set(map(tuple, a.contacts(data=False)[0].tolist()))
    assert pot == brute | real


def test_periodic_celldist_normal_contact():
    # TestPBCCollisions.testNormalContact (test_pbc.py:162-172)
    for shift2 in [(0, 0, 0), (1, 2, 3), (6, 8, -4)]:
        o = Oracle(periodic=1, cell=(1, 1, 1), verletDist=0., gravity=(0, 0, 0))
        mt = o.material()
        o.sphere((.5, .5, .5), .2, mt)
        o.sphere(np.array([.8, .6, .5]) + shift2, .2, mt)
        o.run(["Collider"], 1)
        ids, cd, _ = o.contacts(data=False)
        assert ids.tolist() == [[0, 1]]
        assert cd.tolist() == [[-s for s in shift2]]


def test_periodic_over_half_contact():
    # TestPBCCollisions.testOverHalfContact (test_pbc.py:135-149)
    for move in [(0, 0, 0), (13, 14, 15)]:
        o = Oracle(periodic=1, cell=(1, 1, 1), verletDist=0., gravity=(0, 0, 0))
        mt = o.material()
        o.sphere(np.array([.3, .5, .5]) + move, .4, mt)
        o.sphere((.7, .9, .5), .1, mt)
        o.run(["Collider"], 1)
        ids, cd, _ = o.contacts(data=False)
        assert ids.tolist() == [[0, 1]]
        assert cd.tolist() == [list(move)]


def test_periodic_double_contact_raises():
    # TestPBCCollisions.testDoubleContact (test_pbc.py:151-160)
    o = Oracle(periodic=1, cell=(1, 1, 1), verletDist=0., gravity=(0, 0, 0))
    mt = o.material()
    o.sphere((.3, .5, .5), .4, mt)
    o.sphere((.8, .9, .5), .2, mt)
    with pytest.raises(RuntimeError):
        o.run(["Collider"], 1)


def test_elastic_energy_balance_hertz_periodic():
    # no damping, en outside (0,1): kinetic energy is recovered after collisions end (1e-3)
    sc = periodic_hertz(n=64, spacing=1.3, seed=3, vel_sigma=0.2, en=float("nan"))
    o = build_oracle(sc)
    e0 = o.kinetic_energy()
    for _ in range(400):
        o.run(ENG, 50)
        if len(o.contacts(real_only=True, data=False)[0]) == 0:
            break
    assert len(o.contacts(real_only=True, data=False)[0]) == 0
    assert o.kinetic_energy() == pytest.approx(e0, rel=1e-3)


def test_deterministic_and_openmp_modes_agree():
    sc = deposition(n=900, dense=True, seed=4, vel_sigma=0.05)
    a = build_oracle(sc, deterministic=1)
    b = build_oracle(sc, deterministic=0, nThreads=4)
    a.run(ENG, 20)
    b.run(ENG, 20)
    np.testing.assert_allclose(b.get("pos"), a.get("pos"), rtol=0, atol=1e-12)
    ia, _, _ = a.contacts(real_only=True, data=False)
    ib, _, _ = b.contacts(real_only=True, data=False)
    np.testing.assert_array_equal(ia, ib)


def test_column_contact_force_sign():
    # TestContact.testForceSign (test_core.py:436-453): a resting column touches at (4,5),
    # and the contact pushes 4 down and 5 up
    r = .01
    o = Oracle(damping=.2, gravity=(0, 0, -9.81))
    mt = o.material(density=1000., young=1e6)
    for i in range(10):
        o.sphere((0, 0, r + i * 2 * r * (1 - 1e-3)), r, mt)
    o.wall((0, 0, 0), 2, 1, mt)
    o.set_params(dt=.2 * o.pwave_dt())
    o.run(ENG, 5)
    o.run(["Collider", "ContactLoop"], 1)
    ids, _, d = o.contacts(real_only=True)
    pairs = ids.tolist()
    assert [4, 5] in pairs
    k = pairs.index([4, 5])
    # global normal points from pA to pB (+z); local normal force is compressive (negative)
    assert d[k, 2] > 0.999 and d[k, 9] < 0


# ---------------------------------------------------------------- facets
def _floor_square(o, m, half=1., fakeVel=None, z=0.):
    # two facets tiling [-half, half]^2 in the plane z
    a, b, c, d = (-half, -half, z), (half, -half, z), (half, half, z), (-half, half, z)
    o.facet([a, b, c], fakeVel, 0., m)
    o.facet([a, c, d], fakeVel, 0., m)


@pytest.mark.parametrize("law", ["idealelpl", "hertz"])
def test_facet_face_contact_equals_wall(law):
    # a sphere in the face region of a facet sees the same contact as on a wall in that plane:
    # normal and uN from the plane, lens (r, r), contA pi r^2, force on the sphere only
    runs = []
    for kind in ("wall", "facet"):
        if law == "hertz":
            o = Oracle(physKind=1, lawKind=1, poisson=.25, en=.6, damping=0., gravity=(0, 0, -9.81))
            m = o.material(kind=1, density=2500., young=1e7, ktDivKn=.2, tanPhi=.4)
        else:
            o = Oracle(damping=0., gravity=(0, 0, -9.81))
            m = o.material(density=2500., young=1e7, ktDivKn=.2, tanPhi=.4)
        if kind == "wall":
            o.wall((0, 0, 0), 2, 1, m)
            o.wall((0, 0, 5), 2, -1, m, mask=0)  # keeps the particle count and ids equal
        else:
            _floor_square(o, m)
        r = .01
        o.sphere((0.01, -0.02, r * (1 - 1e-4)), r, m)
        o.set("vel", np.array([[0, 0, 0], [0, 0, 0], [0.2, 0.05, 0.]]))
        o.set_params(dt=.2 * o.pwave_dt())
        o.run(ENG, 3000)
        runs.append((o.get("pos")[2], o.get("vel")[2], o.get("angVel")[2], len(o.contacts(real_only=True)[0])))
    (pw, vw, ww, nw), (pf, vf, wf, nf) = runs
    assert nw == 1 and nf == 1
    assert np.max(np.abs(pf - pw)) <= 1e-12 * .01
    assert np.max(np.abs(vf - vw)) <= 1e-9 * np.max(np.abs(vw))
    assert np.max(np.abs(wf - ww)) <= 1e-9 * np.max(np.abs(ww))


def test_facet_fakevel_drags_resting_sphere_to_two_sevenths():
    # Facet.fakeVel u moves the surface under a sphere at rest: friction accelerates it until the
    # contact point moves with the surface (v - w_y r = u), after which it rolls at 2/7 u
    u, r, rho, E, tanPhi = .1, .01, 2500., 1e7, .5
    o = Oracle(damping=0., gravity=(0, 0, -9.81))
    m = o.material(density=rho, young=E, ktDivKn=.2, tanPhi=tanPhi)
    _floor_square(o, m, fakeVel=(u, 0, 0))
    mass = (4 / 3.) * math.pi * r ** 3 * rho
    o.sphere((0, 0, r - mass * 9.81 / (E * math.pi * r / 2)), r, m)
    o.set_params(dt=.2 * o.pwave_dt())
    t_slip = 2 / 7. * u / (tanPhi * 9.81)
    o.run(ENG, int(3 * t_slip / o.dt))
    v, w = o.get("vel")[2], o.get("angVel")[2]
    assert v[0] == pytest.approx(2 / 7. * u, rel=0.01)
    assert v[0] - w[1] * r == pytest.approx(u, rel=0.01)
    # the facets are fixed and get no force from the contact loop (multinodal particles)
    assert np.all(o.get("force")[:2] == 0.)


@pytest.mark.parametrize("centre,closest", [
    ((0.3, 0.2, 0.006), (0.3, 0.2, 0.)),        # face
    ((0.5, -0.003, 0.002), (0.5, 0., 0.)),       # edge v0-v1
    ((0.5, 0.502, -0.001), (0.499, 0.501, 0.)),  # hypotenuse v1-v2, below the plane
    ((-0.003, 0.4, 0.001), (0., 0.4, 0.)),       # edge v2-v0
    ((-0.002, -0.002, 0.001), (0., 0., 0.)),     # vertex v0
    ((1.002, -0.001, 0.002), (1., 0., 0.)),      # vertex v1
    ((-0.001, 1.003, -0.002), (0., 1., 0.)),     # vertex v2
])
def test_facet_voronoi_regions(centre, closest):
    # closest point of the triangle by region, normal from it to the centre, uN = |d| - r
    r = .008
    o = Oracle(damping=0., gravity=(0, 0, 0))
    m = o.material(density=2500., young=1e7, ktDivKn=.2, tanPhi=.4)
    o.facet([(0, 0, 0), (1, 0, 0), (0, 1, 0)], None, 0., m)
    o.sphere(centre, r, m)
    o.run(["Collider", "ContactLoop"], 1)
    ids, _, d = o.contacts(real_only=True)
    assert ids.tolist() == [[0, 1]]
    dvec = np.array(centre) - np.array(closest)
    dist = np.linalg.norm(dvec)
    np.testing.assert_allclose(d[0, 0:3], dvec / dist, rtol=0, atol=1e-14)
    np.testing.assert_allclose(d[0, 6:9], closest, rtol=0, atol=1e-15)
    assert d[0, 12] == pytest.approx(dist - r, rel=1e-12)


def test_facet_halfthick_shifts_contact():
    # halfThick thickens the facet on both sides: uN = |d| - r - h, contact point moved by h
    r, h = .008, .001
    o = Oracle(damping=0., gravity=(0, 0, 0))
    m = o.material(density=2500., young=1e7, ktDivKn=.2, tanPhi=.4)
    o.facet([(0, 0, 0), (1, 0, 0), (0, 1, 0)], None, h, m)
    o.sphere((0.3, 0.2, -0.0085), r, m)
    o.run(["Collider", "ContactLoop"], 1)
    ids, _, d = o.contacts(real_only=True)
    assert ids.tolist() == [[0, 1]]
    np.testing.assert_allclose(d[0, 0:3], (0, 0, -1), atol=1e-15)
    np.testing.assert_allclose(d[0, 6:9], (0.3, 0.2, -h), atol=1e-15)
    assert d[0, 12] == pytest.approx(0.0085 - r - h, rel=1e-10)


def _global_force(d):
    """global force on pA of each contact row of contacts(): local force in the frame
    (normal, y axis, normal x y)"""
    n, y = d[:, 0:3], d[:, 3:6]
    z = np.cross(n, y)
    return d[:, 9:10] * n + d[:, 10:11] * y + d[:, 11:12] * z


def _two_facets(fold_deg, flip=False):
    """two unit right triangles sharing the edge (0,0,0)-(0,1,0): A in the plane z=0 on x<0 with
    normal +z, B on x>0 folded down about the edge by fold_deg (0: coplanar), listed so that its
    normal is A's turned by the fold (sin, 0, cos); flip lists B's vertices the other way round"""
    t = math.radians(fold_deg)
    a = [(0., 0., 0.), (0., 1., 0.), (-1., 0., 0.)]
    b = [(0., 0., 0.), (math.cos(t), 0., -math.sin(t)), (0., 1., 0.)]
    if flip:
        b = [b[0], b[2], b[1]]
    return a, b


@pytest.mark.parametrize("fold", [0., 30., 90.])
@pytest.mark.parametrize("flip", [False, True])
def test_facet_neighbor_angles(fold, flip):
    # Facet::computeNeighborAngles (Facet.cpp:180-216): the shared edge gets n1 . n2 with the
    # neighbour's normal taken in our orientation: cos(fold) whichever way the neighbour lists its
    # vertices; the unshared edges get NaN
    o = Oracle(gravity=(0, 0, 0))
    m = o.material()
    a, b = _two_facets(fold, flip)
    ia = o.facet(np.array(a).ravel(), None, 0., m)
    ib = o.facet(np.array(b).ravel(), None, 0., m)
    o.facet_neighbor_angles()
    na, nb = o.facet_n21lim(ia), o.facet_n21lim(ib)
    # A's edge 0 (v0, v1) is the shared one
    assert np.isnan(na[1]) and np.isnan(na[2])
    assert na[0] == pytest.approx(math.cos(math.radians(fold)), abs=1e-15)
    shared_b = [k for k in range(3) if not np.isnan(nb[k])]
    assert len(shared_b) == 1 and nb[shared_b[0]] == pytest.approx(math.cos(math.radians(fold)), abs=1e-15)


def _edge_contact(fold, smooth, centre, r=.2):
    o = Oracle(gravity=(0, 0, 0), damping=0.)
    m = o.material(young=1e7)
    a, b = _two_facets(fold)
    ia = o.facet(np.array(a).ravel(), None, 0., m)
    o.facet(np.array(b).ravel(), None, 0., m)
    o.sphere(centre, r, m)
    if smooth:
        o.facet_neighbor_angles()
    o.set_params(dt=1e-7)
    o.run(["Collider", "ContactLoop"], 1)
    ids, _, d = o.contacts(real_only=True)
    k = [q for q in range(len(ids)) if ia in ids[q]]
    assert len(k) == 1
    return d[k[0], 0:3], d[k[0], 6:9]


def test_n21lim_limits_the_edge_normal_to_the_neighbours():
    # Facet::adjustBoundaryContactGeom_impl (Facet.cpp:239-265): a sphere past A's edge, above a
    # neighbour folded by 30 degrees, touches A's edge with a normal tilted towards it; n21lim =
    # cos 30 bounds the tilt: the normal becomes cos30 fN + sin30 out (A's unit outer edge vector,
    # +x), i.e. the neighbour's normal, and the contact point the centre's projection along it
    # onto the plane through the edge
    c = np.array([.15, .4, .1])
    n0, p0 = _edge_contact(30., False, c)
    np.testing.assert_allclose(n0, np.array([.15, 0, .1]) / math.hypot(.15, .1), atol=1e-14)
    n1, p1 = _edge_contact(30., True, c)
    t = math.radians(30.)
    nb = np.array([math.sin(t), 0., math.cos(t)])
    np.testing.assert_allclose(n1, nb, atol=1e-14)
    np.testing.assert_allclose(p1, c - nb * (c @ nb), atol=1e-14)
    # inside the bound (a normal steeper than the neighbour's) nothing changes
    c2 = np.array([.02, .4, .15])
    n2, _ = _edge_contact(30., True, c2)
    np.testing.assert_allclose(n2, np.array([.02, 0, .15]) / math.hypot(.02, .15), atol=1e-14)


def test_n21lim_folds_past_a_right_angle():
    # n21lim = cos(fold): up to a right angle (n21lim >= 0; cos 90 is 6e-17 here) a normal between
    # the two facets' normals is left as it is; past a right angle (n21lim < 0) the same normal is
    # set to nnLim fN + sqrt(1 - nnLim^2) out, which is the neighbour's normal (sin, 0, cos)
    c = np.array([.1, .4, .12])
    n0, _ = _edge_contact(90., False, c)
    n1, _ = _edge_contact(90., True, c)
    np.testing.assert_array_equal(n1, n0)
    t = math.radians(120.)
    nb = np.array([math.sin(t), 0., math.cos(t)])
    n2, p2 = _edge_contact(120., True, c)
    np.testing.assert_allclose(n2, nb, atol=1e-14)
    np.testing.assert_allclose(p2, c - nb * (c @ nb), atol=1e-14)


@pytest.mark.parametrize("centre", [
    (0.2, 0.3, 0.0079),     # face: weights (0.5, 0.2, 0.3)
    (0.5, -0.002, 0.0075),  # edge v0-v1: node 2 gets nothing
    (1.002, -0.001, 0.0075),  # vertex v1: node 1 gets everything, with no moment
])
def test_in2_facet_single_contact_weights(centre):
    # In2_Facet on the unit right triangle (0,0,0), (1,0,0), (0,1,0): the barycentric coordinates
    # of a point (x, y, 0) are (1 - x - y, x, y); node i gets w_i F and (c - v_i) x (w_i F),
    # F being the force on the facet, which is minus the force on the only sphere
    r = .008
    o = Oracle(damping=0., gravity=(0, 0, 0))
    m = o.material(density=2500., young=1e7, ktDivKn=.2, tanPhi=.4)
    o.facet([(0, 0, 0), (1, 0, 0), (0, 1, 0)], None, 0., m)
    o.sphere(centre, r, m)
    o.set("vel", np.array([[0, 0, 0], [0.05, -0.03, 0.]]))
    o.set_params(dt=.2 * o.pwave_dt())
    o.run(["Collider", "ContactLoop", "IntraForce", "Leapfrog"], 4)
    o.run(["ForceResetter", "Collider", "ContactLoop", "IntraForce"], 1)
    ids, _, d = o.contacts(real_only=True)
    assert ids.tolist() == [[0, 1]]
    c = d[0, 6:9]
    assert c[2] == 0.
    F = -o.get("force")[1]
    assert abs(F[0]) > 0 and abs(F[1]) > 0 and F[2] < 0  # friction acts and the sphere presses down
    np.testing.assert_allclose(F, _global_force(d)[0], rtol=0, atol=1e-14 * np.abs(F).max())
    w = np.array([1 - c[0] - c[1], c[0], c[1]])
    nf, nt = o.facet_nodes(0)
    V = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0)], float)
    for i in range(3):
        np.testing.assert_allclose(nf[i], w[i] * F, rtol=0, atol=1e-15 * np.abs(F).max())
        np.testing.assert_allclose(nt[i], np.cross(c - V[i], w[i] * F), rtol=0, atol=1e-15 * np.abs(F).max())
    if centre[0] > 1:
        assert np.all(nf[[0, 2]] == 0) and np.all(nt == 0)
    if centre[1] < 0:
        assert np.all(nf[2] == 0) and np.all(nt[2] == 0)
    with pytest.raises(RuntimeError):
        o.facet_nodes(1)  # a sphere


def test_in2_facet_weights_reproduce_contact_point_on_tilted_triangle():
    # on a scalene triangle out of the coordinate planes, the weights read back from the node
    # forces (w_i = F_i . F / F . F) are the barycentric coordinates of the contact point: they
    # sum to one and interpolate the vertices to it
    V = np.array([(0.1, -0.2, 0.05), (0.9, 0.1, 0.3), (0.2, 0.7, -0.1)])
    n = np.cross(V[1] - V[0], V[2] - V[0])
    n /= np.linalg.norm(n)
    r = .008
    o = Oracle(damping=0., gravity=(0, 0, 0))
    m = o.material(density=2500., young=1e7, ktDivKn=.2, tanPhi=.4)
    o.facet(V, None, 0., m)
    target = .2 * V[0] + .5 * V[1] + .3 * V[2]
    o.sphere(target + n * r * (1 - 1e-3), r, m)
    o.set("vel", np.array([[0, 0, 0], [0.05, -0.03, 0.02]]))
    o.set_params(dt=.2 * o.pwave_dt())
    o.run(["Collider", "ContactLoop", "IntraForce", "Leapfrog"], 4)
    ids, _, d = o.contacts(real_only=True)
    assert ids.tolist() == [[0, 1]]
    c = d[0, 6:9]
    nf, nt = o.facet_nodes(0)
    F = nf.sum(0)
    w = nf @ F / (F @ F)
    assert w.sum() == pytest.approx(1., abs=1e-14)
    np.testing.assert_allclose(w @ V, c, rtol=0, atol=1e-14)
    np.testing.assert_allclose(w, (.2, .5, .3), atol=1e-3)  # the sphere has barely moved
    for i in range(3):
        np.testing.assert_allclose(nt[i], np.cross(c - V[i], nf[i]), rtol=0, atol=1e-15 * np.abs(F).max())


def test_in2_facet_drum_force_and_moment_balance():
    # In the rotating drum every facet's node forces add up to the force of its contacts, and the
    # node forces and torques have the same moment about any point as the contact forces acting
    # at their contact points (weights summing to one)
    from tests.scenes import drum, n_boundary
    sc = drum(n=500, seed=3, omega=4.)
    o = build_oracle(sc)
    eng = ["Collider", "ContactLoop", "IntraForce", "Leapfrog"]
    o.run(eng, 300)
    nW = n_boundary(sc)
    facets = [(k + len(sc["walls"]), f) for k, f in enumerate(sc["facets"])]
    ids, _, d = o.contacts(real_only=True)
    Fg = _global_force(d)
    pt = np.array([0.01, -0.02, 0.005])
    touched = 0
    for pid, f in facets:
        nf, nt = o.facet_nodes(pid)
        rows = np.nonzero(ids[:, 0] == pid)[0]
        assert np.all(ids[rows, 1] >= nW)
        Fc, cc = Fg[rows], d[rows, 6:9]
        scale = max(np.abs(Fc).sum(), 1e-300)
        np.testing.assert_allclose(nf.sum(0), Fc.sum(0), rtol=0, atol=1e-13 * scale)
        M_nodes = sum(np.cross(f["verts"][i] - pt, nf[i]) + nt[i] for i in range(3))
        M_cont = np.cross(cc - pt, Fc).sum(0) if len(rows) else np.zeros(3)
        np.testing.assert_allclose(M_nodes, M_cont, rtol=0, atol=1e-13 * scale * .1)
        if len(rows) == 0:
            assert np.all(nf == 0) and np.all(nt == 0)
        touched += len(rows) > 0
    assert touched > 10


def _rot(axis, ang):
    """Rotation matrix about a unit axis (Rodrigues)."""
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _corotating_pair(omega, steps, dt=1e-4, r=.01):
    """Two touching spheres (normal along x) moved as one rigid body rotating at omega about
    their midpoint, positions and velocities set each step (no integrator: the contact loop
    sees exactly the rigid motion, with leapfrog's convention of velocities at mid-step).
    Returns the contact's (normal, local y axis, local force) at creation and after `steps`."""
    # no renormalisation (trsfRenorm 0): it would project the frame back onto the new normal
    # and hide how the frame was carried
    o = Oracle(physKind=0, lawKind=0, damping=0., gravity=(0, 0, 0), dt=dt, trsfRenorm=0)
    mt = o.material(density=2000., young=1e7, ktDivKn=.2, tanPhi=.5)
    half = r * (1 - 1e-7)
    x0 = np.array([[-half, 0., 0.], [half, 0., 0.]])
    o.sphere(x0[0], r, mt)
    o.sphere(x0[1], r, mt)
    w = np.asarray(omega, float)
    wn = np.linalg.norm(w)

    def place(t):
        x = x0 @ _rot(w, wn * t).T
        xm = x0 @ _rot(w, wn * (t - dt / 2)).T
        o.set("pos", x)
        o.set("vel", np.cross(w, xm))
        o.set("angVel", np.tile(w, (2, 1)))

    place(0.)
    o.run(["Collider", "ContactLoop"], 1)
    d0 = o.contacts(real_only=True)[2][0]
    for k in range(1, steps + 1):
        place(k * dt)
        o.run(["ContactLoop"], 1)
    ids, _, d = o.contacts(real_only=True)
    assert len(ids) == 1
    return d0, d[0], _rot(w, wn * steps * dt)


def test_l6geom_corotation_perpendicular_to_normal():
    # a touching pair turning rigidly about an axis perpendicular to its normal (x), and oblique
    # to the frame's initial y axis (-y): the contact frame turns with it (normRotVec,
    # L6Geom.cpp:74-117) and the material points at the contact do not slide, so the shear force
    # stays at the chord error of the midpoint rule. A wrong sign of normRotVec turns the frame's
    # y axis the other way about the rotation axis; a wrong branch vector gives velT ~ omega * 2r.
    steps, dt, r, w = 400, 1e-4, .01, np.array([0., 2.5, 2.5]) / math.sqrt(2)
    d0, d, R = _corotating_pair(w, steps, dt, r)
    assert np.max(np.abs(d[0:3] - R @ d0[0:3])) < 1e-12
    assert np.max(np.abs(d[3:6] - R @ d0[3:6])) < 1e-6
    kt = d[14]
    wrong = kt * np.linalg.norm(w) * 2 * r * dt * steps
    assert math.hypot(d[10], d[11]) < 1e-6 * wrong


def test_l6geom_twist_about_normal():
    # both spheres spinning at the same rate about the normal: the frame's y axis turns about the
    # normal by the spin angle (normTwistVec = n dt (n.(w1+w2))/2), no relative motion, no shear
    steps, dt, w = 400, 1e-4, np.array([3.0, 0., 0.])
    d0, d, R = _corotating_pair(w, steps, dt)
    assert np.max(np.abs(d[0:3] - d0[0:3])) < 1e-15
    assert np.max(np.abs(d[3:6] - R @ d0[3:6])) < 1e-6
    # half the rate (a factor 1/2 too many) or the opposite sense would be off by >= 0.03 rad
    assert np.max(np.abs(d[3:6] - d0[3:6])) > 0.05
    assert math.hypot(d[10], d[11]) == 0.


def test_trsf_orthonormal_after_renormalisation():
    # trsfRenorm (L6Geom.cpp:124-136): at steps that are multiples of 100 the frame is
    # re-orthonormalised; in between it drifts. Right after the renormalising step every real
    # contact's columns 0 and 1 are unit and orthogonal to rounding.
    sc = periodic_hertz(n=400, seed=5, spacing=.8, vel_sigma=.5)
    o = build_oracle(sc)
    o.run(ENG, 100)  # steps 0..99; the frame drifts
    d99 = o.contacts(real_only=True)[2]
    o.run(ENG, 1)  # step 100 renormalises
    d = o.contacts(real_only=True)[2]
    assert len(d) > 50

    def dev(x):
        n, y = x[:, 0:3], x[:, 3:6]
        return max(np.max(np.abs(np.linalg.norm(n, axis=1) - 1)), np.max(np.abs(np.linalg.norm(y, axis=1) - 1)),
                   np.max(np.abs(np.sum(n * y, axis=1))))

    print("deviation before / after", dev(d99), dev(d))
    assert dev(d) < 4e-16
    assert dev(d99) > dev(d)


def test_dyndt_rotational_stiffness_decides():
    # one vertical wall contact with kt = 0.8 kn: the rotational term of DynDt.cpp:16-58,
    # I / (l^2 kt) = (2/5) m / (0.8 kn) = 0.5 m / kn, is below m / kn, so the critical step is
    # sqrt(2 * 0.5 m / kn). A rotational stiffness without l^2, with kn in place of kt, or left
    # out, gives another value.
    r, rho, E = .01, 2000., 1e7
    o = Oracle(damping=0., gravity=(0, 0, 0))
    mt = o.material(density=rho, young=E, ktDivKn=.8, tanPhi=.5)
    o.wall((0, 0, 0), 2, 1, mt)
    o.sphere((0, 0, r * (1 - 1e-4)), r, mt)
    o.set_params(dt=.1 * o.pwave_dt())
    o.run(["Collider", "ContactLoop"], 1)
    mass = o.get("mass")[1]
    kn = E * math.pi * r / 2
    assert o.crit_dt() == pytest.approx(math.sqrt(2 * 0.5 * mass / kn), rel=1e-12)
