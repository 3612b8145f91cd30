"""ContactLoop.updatePhys on the HIP engine against the oracle (ContactLoop.hpp:119,
ContactLoop.cpp:179-180 and 265-266).

The material is edited in place between runs (young x4, then back). With 'never' each contact
keeps the phys of the materials it was created with, so the old contacts keep the old
stiffness and the contacts created after the edit take the new one; 'always' recomputes every
contact's phys at every step; 'once' recomputes them in the next contact loop only and then
reads 'never'. The runs go step by step (wdg_run) and through the K-step graph replay
(wdg_run_graph), whose captured contact loop carries the contacts' material epochs. Bars as
in the other parity tests: identical real-contact sets, positions within 1e-9 r, and each
contact's stiffness kn within 1e-12 relative.
"""
import numpy as np
import pytest

from tests.scenes import _particle_order, build_gpu, build_oracle, contact_kw, deposition, periodic_hertz

pytestmark = pytest.mark.gpu

ENG = ["Collider", "ContactLoop", "Leapfrog"]


def _scene(kind):
    if kind == "frict":
        # spheres on a lattice with slight overlaps, kicked: contacts break and form again
        return deposition(n=1500, seed=5, dense=True, vel_sigma=.05)
    return periodic_hertz(n=1000, spacing=0.92, seed=33, vel_sigma=1.)


def _radii(sc):
    return np.array([sc["radius"][i] if k == "s" else 1. for k, i in _particle_order(sc)])


def _compare(o, d, sc, tag):
    io, _, do = o.contacts(real_only=True)
    ig, _, dg = d.contacts(real_only=True)
    np.testing.assert_array_equal(io, ig, err_msg=tag)
    err = float(np.max(np.linalg.norm(d.get("pos") - o.get("pos"), axis=1) / _radii(sc)))
    assert err <= 1e-9, (tag, err)
    np.testing.assert_allclose(dg[:, 13], do[:, 13], rtol=1e-12, err_msg=tag)
    return {tuple(p) for p in io}


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("up", ["never", "always", "once"])
@pytest.mark.parametrize("kind", ["frict", "hertz"])
def test_update_phys_matches_oracle(kind, up, graph):
    sc = _scene(kind)
    o = build_oracle(sc)
    d = build_gpu(sc)
    run_kw = dict(contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0, graph=graph)
    step = 0
    o.run(ENG, 30)
    step, _ = d.run(ENG, 30, step, sc["dt"], **run_kw)
    before = _compare(o, d, sc, f"{kind} {up}: before the edit")
    m0 = dict(sc["materials"][0])
    m1 = dict(m0, young=4 * m0["young"])
    o.set_material(0, **m1)
    d.set_materials([m1])
    o.update_phys = up
    d.update_phys = up
    o.run(ENG, 40)
    step, _ = d.run(ENG, 40, step, sc["dt"], **run_kw)
    after = _compare(o, d, sc, f"{kind} {up}: after the edit")
    expect = "always" if up == "always" else "never"
    assert o.update_phys == expect and d.update_phys == expect
    # contacts created after the edit take the new materials: the case holds some
    assert len(after - before) > 0
    # back to the first table: with 'never', three tables are referred to (the third equal to
    # the first), with 'always' one
    d.set_materials([m0])
    o.set_material(0, **m0)
    o.run(ENG, 30)
    step, _ = d.run(ENG, 30, step, sc["dt"], **run_kw)
    _compare(o, d, sc, f"{kind} {up}: after the second edit")
    d.close()


def test_update_phys_argument_checks():
    sc = _scene("frict")
    d = build_gpu(sc)
    assert d.update_phys == "never"
    with pytest.raises(Exception):
        d.update_phys = 3
    d.update_phys = "once"
    assert d.update_phys == "once"
    d.close()


def test_snapshot_keeps_each_contacts_material_epoch():
    # with 'never' after a material change, a snapshot keeps which contacts have the old table:
    # stepping on from the restored snapshot repeats the first continuation bit for bit
    sc = _scene("frict")
    d = build_gpu(sc)
    kw = dict(contact=contact_kw(sc), leapfrog=dict(damping=sc["damping"]), dynDtPeriod=0)
    step, _ = d.run(ENG, 30, 0, sc["dt"], **kw)
    m0 = dict(sc["materials"][0])
    d.set_materials([dict(m0, young=4 * m0["young"])])
    step, _ = d.run(ENG, 20, step, sc["dt"], **kw)
    d.snapshot(1)
    s1, _ = d.run(ENG, 40, step, sc["dt"], **kw)
    p1, k1 = d.get("pos"), d.contacts(real_only=True)[2][:, 13]
    d.restore(1)
    s2, _ = d.run(ENG, 40, step, sc["dt"], **kw)
    np.testing.assert_array_equal(d.get("pos"), p1)
    np.testing.assert_array_equal(d.contacts(real_only=True)[2][:, 13], k1)
    assert len(set(np.round(k1 / k1.min(), 6).tolist())) > 1  # both tables are in use
    d.close()
