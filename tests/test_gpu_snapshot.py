"""In-memory snapshots of the device state (the role of Scene.saveTmp / loadTmp): a restored
state runs on exactly as the state did when it was kept, and the distance to a snapshot is the
largest displacement since."""
import numpy as np
import pytest

from tests.scenes import build_gpu, contact_kw, deposition, gpu_arrays

pytestmark = pytest.mark.gpu


def test_restore_replays_the_same_trajectory():
    sc = deposition(n=600, seed=7, dense=True, vel_sigma=0.05)
    a = gpu_arrays(sc)
    d = build_gpu(sc)
    eng = ["Collider", "ContactLoop", "Leapfrog"]
    kw = dict(collide=dict(verletDist=-.05, forceInit=0), contact=contact_kw(sc), leapfrog=dict(damping=.2),
              dynDtPeriod=0)
    step, _ = d.run(eng, 50, 0, sc["dt"], **kw)
    d.snapshot(1)
    assert d.snapshot_distance(1) == 0.
    p0 = d.get("pos")
    # the uninterrupted run
    d.run(eng, 80, step, sc["dt"], **kw)
    pu, vu = d.get("pos"), d.get("vel")
    # a continuation from the snapshot: the contact history comes back with it, so it is the
    # uninterrupted run bit for bit
    d.restore(1)
    assert np.array_equal(d.get("pos"), p0)
    d.run(eng, 80, step, sc["dt"], **kw)
    p1, v1 = d.get("pos"), d.get("vel")
    assert np.array_equal(p1, pu) and np.array_equal(v1, vu)
    moved = d.snapshot_distance(1)
    sph = a["shape"] == 0
    assert moved == pytest.approx(np.max(np.linalg.norm((p1 - p0)[sph], axis=1)), rel=1e-12)
    assert moved > 0
    # the same again: identical
    d.restore(1)
    assert np.array_equal(d.get("pos"), p0)
    d.run(eng, 80, step, sc["dt"], **kw)
    assert np.array_equal(d.get("pos"), p1) and np.array_equal(d.get("vel"), v1)
    with pytest.raises(RuntimeError):
        d.restore(3)  # nothing kept there
    d.close()
