"""The CPU restatement of the learned step's network (oracle/mpnn.py): its hand-derived
backward pass is pinned by central finite differences of its forward pass (plain float64),
and the BF16-emulating mode stays close to it."""
import numpy as np
import pytest

from oracle.mpnn import NEI, NIN, Net, bf16, layout


def random_graph(rng, N=40, deg=4):
    pairs = set()
    while len(pairs) < N * deg // 2:
        i, j = rng.integers(0, N, 2)
        if i != j:
            pairs.add((min(i, j), max(i, j)))
    p = np.array(sorted(pairs))
    recv = np.concatenate([p[:, 0], p[:, 1]])
    send = np.concatenate([p[:, 1], p[:, 0]])
    o = np.lexsort((send, recv))
    recv, send = recv[o], send[o]
    rowOff = np.concatenate([[0], np.cumsum(np.bincount(recv, minlength=N))])
    key = recv * N + send
    rev = np.searchsorted(key, send * N + recv)
    x = rng.normal(size=(N, NIN))
    ef = rng.normal(size=(len(send), NEI))
    t = rng.normal(size=(N, 3))
    return rowOff, send, rev, x, ef, t


def random_params(rng, scale=.15):
    off, n = layout()
    p = rng.normal(0, scale, size=n)
    # the padding rows of the inputs stay zero, as on the device
    o, s = off["encW"]
    p[o:o + s].reshape(32, 64)[NIN:] = 0.
    for l in range(2):
        o, s = off[f"mW{l}"]
        p[o:o + s].reshape(160, 64)[128 + NEI:] = 0.
    for name in ("encB", "mB0", "mB1", "uB0", "uB1"):
        o, s = off[name]
        p[o:o + s] = np.abs(p[o:o + s]) + .05  # keep units active
    return p


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    rowOff, send, rev, x, ef, t = random_graph(rng)
    p = random_params(rng)
    net = Net(p, emulate=False)
    loss, g = net.loss_grad(x, rowOff, send, rev, ef, t)
    off, n = layout()
    # a sample of parameters from every tensor
    idx = []
    for name, (o, s) in off.items():
        nz = [o + k for k in range(s) if p[o + k] != 0. or name.endswith("B")]
        idx += list(rng.choice(nz, size=min(6, len(nz)), replace=False))
    h = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += h
        lp, _ = Net(q, emulate=False).loss_grad(x, rowOff, send, rev, ef, t)
        q[k] -= 2 * h
        lm, _ = Net(q, emulate=False).loss_grad(x, rowOff, send, rev, ef, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1., abs(fd)), (k, fd, g[k])


def test_bf16_rounding_and_emulation():
    # ties to even, and a few exact values
    assert bf16(1.0) == 1.0 and bf16(1 + 2 ** -8) == 1.0 and bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6
    # against torch's float32 -> bfloat16 conversion (round to nearest even) on random values
    import torch
    v = np.random.default_rng(5).standard_normal(100000).astype(np.float32) * np.float32(1e3)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(bf16(v), ref)
    rng = np.random.default_rng(4)
    rowOff, send, rev, x, ef, t = random_graph(rng)
    p = random_params(rng)
    l64, g64 = Net(p, emulate=False).loss_grad(x, rowOff, send, rev, ef, t)
    l16, g16 = Net(p, emulate=True).loss_grad(x, rowOff, send, rev, ef, t)
    assert abs(l16 - l64) < 2e-2 * l64
    assert np.linalg.norm(g16 - g64) < 5e-2 * np.linalg.norm(g64)
