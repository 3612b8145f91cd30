"""Checkpoint and resume of the device state with its contact history.

The save half is the real contacts of wdg_contact_list with the node fields; the load half is
wdg_restore_contacts after a fresh upload. The bar is bit-exact: stepping on from a restored
context gives the positions, velocities and contact states of the uninterrupted run, with
difference 0, and both equal the CPU oracle's uninterrupted run.
"""
import numpy as np
import pytest

from tests.scenes import GpuRunner, build_oracle, deposition, drum, periodic_hertz

pytestmark = pytest.mark.gpu

ENG = ["Collider", "ContactLoop", "Leapfrog"]


def scene(kind):
    if kind == "frict_deposition":
        return deposition(n=800, dense=True, seed=61, vel_sigma=0.05)
    if kind == "hertz_periodic":
        return periodic_hertz(n=1500, spacing=0.9, seed=62, en=.6)
    return drum(n=1500, seed=63)


@pytest.mark.parametrize("kind", ["frict_deposition", "hertz_periodic", "drum_facets"])
def test_resume_is_bit_exact(kind, tmp_path):
    from woo_amd._native import Device

    sc = scene(kind)
    g = GpuRunner(sc)
    g.run(ENG, 120)
    saved = g.d.checkpoint()
    np.savez(tmp_path / "ck.npz", **saved)
    ck = dict(np.load(tmp_path / "ck.npz"))
    n_real = len(ck["c_ids"])
    print(kind, "real contacts in the checkpoint", n_real)
    assert n_real > 50
    # the uninterrupted run
    g.run(ENG, 80)
    # a new context from the file: the same real contacts with the same state
    r = GpuRunner(sc)
    r.d.close()
    r.d = Device.from_checkpoint(ck, verletDist=r.vd)
    r.step = 120
    ids, cd, data = r.d.contacts(real_only=True)
    np.testing.assert_array_equal(ids, ck["c_ids"])
    np.testing.assert_array_equal(cd, ck["c_cellDist"])
    np.testing.assert_array_equal(data[:, :13], ck["c_data"][:, :13])
    r.run(ENG, 80)
    for f in ("pos", "vel", "angVel", "ori"):
        np.testing.assert_array_equal(r.d.get(f), g.d.get(f), err_msg=f)
    a, b = g.d.contacts(real_only=True), r.d.contacts(real_only=True)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[2][:, :13], b[2][:, :13])
    # and the oracle's uninterrupted run
    o = build_oracle(sc)
    o.run(ENG, 200)
    np.testing.assert_array_equal(r.d.get("pos"), o.get("pos"))


def test_restore_rejects_inconsistent_contacts():
    from woo_amd._native import Device, WdgError

    sc = scene("frict_deposition")
    g = GpuRunner(sc)
    g.run(ENG, 60)
    ck = g.d.checkpoint()
    # particles out of role order
    bad = dict(ck)
    bad["c_ids"] = ck["c_ids"][:, ::-1].copy()
    sph = np.nonzero(np.all(ck["p_shape"][ck["c_ids"]] == 0, axis=1))[0]
    bad["c_ids"] = bad["c_ids"][sph]
    bad["c_cellDist"] = ck["c_cellDist"][sph]
    bad["c_data"] = ck["c_data"][sph]
    with pytest.raises(WdgError, match="role order"):
        Device.from_checkpoint(bad)
    # a contact between particles whose boxes do not overlap
    far = dict(ck)
    pos = ck["p_pos"].copy()
    i = int(ck["c_ids"][sph[0], 1])
    pos[i] += 0.05
    far["p_pos"] = pos
    with pytest.raises(WdgError, match="do not overlap"):
        Device.from_checkpoint(far)
    # listed twice
    dup = dict(ck)
    dup["c_ids"] = np.concatenate([ck["c_ids"], ck["c_ids"][:1]])
    dup["c_cellDist"] = np.concatenate([ck["c_cellDist"], ck["c_cellDist"][:1]])
    dup["c_data"] = np.concatenate([ck["c_data"], ck["c_data"][:1]])
    with pytest.raises(WdgError, match="twice"):
        Device.from_checkpoint(dup)


def test_resume_keeps_energy_terms(tmp_path):
    # Scene.energy travels with the checkpoint; the resumed terms equal the uninterrupted run's
    # up to summation order (the sums run over the workgroups of a different storage order)
    from woo_amd._native import Device

    sc = scene("frict_deposition")
    g = GpuRunner(sc)
    g.d.set_energy_tracking(True)
    g.run(ENG, 100)
    saved = g.d.checkpoint()
    np.savez(tmp_path / "ck.npz", **saved)
    ck = dict(np.load(tmp_path / "ck.npz"))
    assert bool(ck["track_energy"]) and ck["energy"][4] > 0  # plast
    g.run(ENG, 60)
    r = GpuRunner(sc)
    r.d.close()
    r.d = Device.from_checkpoint(ck, verletDist=r.vd)
    r.step = 100
    e0 = r.d.energy()
    assert [e0[k] for k in e0] == list(ck["energy"])
    r.run(ENG, 60)
    np.testing.assert_array_equal(r.d.get("pos"), g.d.get("pos"))
    a, b = g.d.energy(), r.d.energy()
    tot = sum(abs(v) for v in a.values())
    for k in a:
        assert abs(a[k] - b[k]) <= 1e-11 * max(abs(a[k]), 1e-3 * tot), (k, a[k], b[k])
