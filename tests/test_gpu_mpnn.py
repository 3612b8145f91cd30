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
    from woo_amd.surrogate import Graph, Network

    rng = np.random.default_rng(seed)
    rowOff, send, rev, x, ef, t = random_graph(rng, N=700, deg=6)
    p = as_f32(random_params(rng))
    x, ef, t = as_f32(x), as_f32(ef), as_f32(t)
    net = Network(seed=1)
    net.set_params(p)
    g = Graph(rowOff, send, rev, x, ef)
    net.set_graph(g, t)
    loss_gpu = net.loss_grad(g, t * 9.81)
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
    out = net.predict(g) / 9.81
    a_cpu, _ = Net(p, emulate=True).forward(x, rowOff, send, ef)
    assert np.max(np.abs(out - a_cpu)) <= 1e-2 * np.max(np.abs(a_cpu))


def test_device_graph_equals_host_graph():
    from woo_amd._native import Device
    from woo_amd.surrogate import (Network, column, contact_graph, dem_dt, dem_run, engine_arrays, material)

    pos, rad = column(0.04, 0.08, seed=5)
    d = Device(0)
    d.set_field(gravity=(0., 0., -9.81), periodic=False)
    d.set_materials([material()])
    d.upload_particles(**engine_arrays(pos, rad))
    step = dem_run(d, 4000, 0, dem_dt())
    net = Network(seed=3)
    n, e, rb = net.graph_from(d, floorZ=0., force=True)
    a_dev = net.forward()
    P, V = d.get("pos")[1:], d.get("vel")[1:]
    g = contact_graph(P, V, rad)
    print("nodes", n, "edges", e, "host edges", g.E)
    assert n == len(rad) and e == g.E and e > 100
    a_host = net.predict(g)
    np.testing.assert_allclose(a_dev, a_host, rtol=1e-5, atol=1e-6 * np.max(np.abs(a_host)))


def test_training_gains_over_untrained_and_ballistic():
    from woo_amd.surrogate import Network, column_samples

    train, _ = column_samples(n_columns=48, snapshots=10, every=1200, seed=200)
    test, _ = column_samples(n_columns=16, snapshots=10, every=1200, seed=900)
    net = Network(seed=7)

    def rmse(pred_fn):
        e = [np.mean(np.sum((pred_fn(g) - a) ** 2, axis=1)) for g, a, _, _ in test]
        return float(np.sqrt(np.mean(e)))

    before = rmse(net.predict)
    ballistic = rmse(lambda g: np.tile([0., 0., -9.81], (g.N, 1)))
    hist = net.train([(g, a) for g, a, _, _ in train], epochs=100, lr=1e-3)
    after = rmse(net.predict)
    print("rmse untrained", before, "ballistic", ballistic, "trained", after, "loss", hist[0], hist[-1])
    # held-out columns: the trained network predicts the coarse-step accelerations better than
    # free fall (the coarse step without it) and than the untrained network
    assert hist[-1] < 0.8 * hist[0]
    assert after < 0.9 * ballistic and after < 0.85 * before
