"""The learned coarse step on the GPU: gradients of the BF16 MFMA kernels against the CPU
restatement (oracle/mpnn.py), the device-built contact graph against the host-built one, and
what training on the engine's own trajectories gains over the method without it."""
import numpy as np
import pytest

from oracle.mpnn import Net, layout
from tests.test_mpnn_oracle import random_graph, random_params

pytestmark = pytest.mark.gpu


def as_f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


@pytest.mark.parametrize("seed", [11, 12])
def test_gradients_match_cpu_oracle(seed):
    from woo_amd.surrogate import A0, Graph, Network

    rng = np.random.default_rng(seed)
    rowOff, send, rev, x, ef, t = random_graph(rng, N=700, deg=6)
    p = as_f32(random_params(rng))
    x, ef, t = as_f32(x), as_f32(ef), as_f32(t)
    net = Network(seed=1)
    net.set_params(p)
    g = Graph(rowOff, send, rev, x, ef)
    net.set_graph(g, t)
    loss_gpu = net.loss_grad(g, t * A0)
    grad_gpu = net.grads().astype(np.float64)
    loss_cpu, grad_cpu = Net(p, emulate=True).loss_grad(x, rowOff, send, rev, ef, t)
    print("loss gpu", loss_gpu, "cpu", loss_cpu)
    assert abs(loss_gpu - loss_cpu) <= 2e-3 * loss_cpu
    off, _ = layout()
    for name, (o, s) in off.items():
        a, b = grad_gpu[o:o + s], grad_cpu[o:o + s]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))
        print(f"{name:6s} rel {rel:.2e} cos {cos:.6f}")
        assert rel < 2e-2 and cos > 0.9995, name
    # forward outputs against the oracle
    out = net.predict(g) / A0
    a_cpu, _ = Net(p, emulate=True).forward(x, rowOff, send, ef)
    assert np.max(np.abs(out - a_cpu)) <= 1e-2 * np.max(np.abs(a_cpu))


def test_device_graph_equals_host_graph():
    # several columns as independent scenes of one context, far enough into the collapse that the
    # forced refresh of the graph builder re-sorts the particles
    from woo_amd._native import Device
    from woo_amd.surrogate import (Network, column, contact_graph, contact_params, dem_dt, dem_run, engine_arrays,
                                   material)

    parts = []
    for c in range(4):
        p, r = column(0.03, 0.3, seed=5000 + c, offset=((c % 2) * 0.08, (c // 2) * 0.08, 0.))
        parts.append(engine_arrays(p, r))
    arr = {k: np.concatenate([q[k] for q in parts]) for k in parts[0]}
    arr["scene"] = np.concatenate([np.full(len(q["shape"]), i, np.int32) for i, q in enumerate(parts)])
    sph = arr["shape"] == 0
    rad = arr["radius"][sph]
    d = Device(0)
    d.set_field(gravity=(0., 0., -9.81), periodic=False)
    d.set_materials([material()])
    d.upload_particles(**arr)
    d.set_scenes(np.full(4, dem_dt()))
    step = dem_run(d, 12000, 0, dem_dt())
    net = Network(seed=3)
    # random weights everywhere (the decoder of a new network starts at zero)
    net.set_params(as_f32(random_params(np.random.default_rng(4))))
    n, e, rb = net.graph_from(d, floorZ=0., force=True)
    a_dev = net.forward()
    gd, row = net.graph()
    P, V = d.get("pos")[sph], d.get("vel")[sph]
    g = contact_graph(P, V, rad)
    print("nodes", n, "edges", e, "host edges", g.E)
    assert n == len(rad) and e == g.E and e > 1000
    # the same edges, with the same inputs (as stored in BF16)
    recv_d = np.repeat(np.arange(gd.N), np.diff(gd.rowOff))
    recv_h = np.repeat(np.arange(g.N), np.diff(g.rowOff))
    kd = row[recv_d] * n + row[gd.send]
    kh = recv_h.astype(np.int64) * n + g.send
    od, oh = np.argsort(kd), np.argsort(kh)
    assert np.array_equal(kd[od], kh[oh])
    bf = as_bf16
    np.testing.assert_array_equal(gd.ef[od], bf(g.ef[oh]))
    xd = np.empty_like(gd.x)
    xd[row] = gd.x
    np.testing.assert_array_equal(xd, bf(g.x))
    a_host = net.predict(g)
    np.testing.assert_allclose(a_dev, a_host, rtol=1e-5, atol=1e-6 * np.max(np.abs(a_host)))
    # a refresh re-sorts the context's particles; the outputs of the graph built before stay in
    # the user-id order of its spheres
    net.graph_from(d, floorZ=0., force=True)
    d.run(["Collider", "ContactLoop", "Leapfrog"], 1, step, dem_dt(), collide=dict(verletDist=-.16, forceInit=1),
          contact=contact_params(), leapfrog=dict(damping=0.), dynDtPeriod=0)
    np.testing.assert_array_equal(net.forward(), a_dev)


def as_bf16(a):
    """float32 values rounded to BF16 (nearest even), as the device stores inputs"""
    from oracle.mpnn import bf16
    return bf16(np.asarray(a, dtype=np.float32)).astype(np.float32)


def test_training_gains_over_untrained_and_ballistic():
    from woo_amd.surrogate import Network, column_samples

    train, _ = column_samples(n_columns=48, snapshots=10, every=1200, seed=200)
    test, _ = column_samples(n_columns=16, snapshots=10, every=1200, seed=900)
    net = Network(seed=7)

    def rmse(pred_fn):
        e = [np.mean(np.sum((pred_fn(g) - a) ** 2, axis=1)) for g, a, _, _ in test]
        return float(np.sqrt(np.mean(e)))

    # the learned step: the network's accelerations for spheres in contact, free flight for the
    # others (the untrained network is the ballistic step)
    before = rmse(net.step_acc)
    ballistic = rmse(lambda g: np.tile([0., 0., -9.81], (g.N, 1)))
    hist = net.train([(g, a) for g, a, _, _ in train], epochs=100, lr=1e-3)
    after = rmse(net.step_acc)
    print("rmse untrained", before, "ballistic", ballistic, "trained", after, "loss", hist[0], hist[-1])
    # held-out columns: the trained network predicts the coarse-step accelerations better than
    # free fall (the coarse step without it) and than the untrained network
    assert hist[-1] < 0.8 * hist[0]
    assert abs(before - ballistic) < 1e-3 * ballistic
    assert after < 0.85 * ballistic


def test_kept_graphs_train_like_host_graphs():
    # a graph built on the device and kept there gives the same loss as the current graph, and
    # training over kept graphs lowers the loss
    import math

    from woo_amd.surrogate import A0, Network, column_training_set, held_out_rmse

    net = Network(seed=5)
    slots, ball, n = column_training_set(net, 0, n_columns=16, width=0.03, height=0.1, snapshots=4, every=1200,
                                         seed=300)
    assert n > 100 and len(slots) == 4
    l0 = [net.kept_loss(k) for k in slots]
    assert l0 == [net.kept_loss(k) for k in slots]  # deterministic
    hist = net.train_kept(slots, epochs=60, lr=1e-3, lr_end=1e-4)
    l1 = [net.kept_loss(k) for k in slots]
    learned, ballistic = held_out_rmse(net, slots)
    print("loss before", np.mean(l0), "after", np.mean(l1), "rmse learned", learned, "ballistic", ballistic)
    # the untrained network is the ballistic step; training lowers the error of the spheres in contact
    assert abs(math.sqrt(np.mean(l0)) * A0 - ballistic) < 1e-3 * ballistic
    assert np.mean(l1) < 0.7 * np.mean(l0)
