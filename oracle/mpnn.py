"""CPU oracle of the learned coarse step's network — TEST INFRASTRUCTURE ONLY.

A NumPy restatement of the message-passing network that the HIP kernels in
woo_amd/csrc/wdg_mpnn.hip train and run, with its forward pass and the hand-derived
backward pass. Used only by tests/ to check the GPU gradients; the product never calls it.

Two modes:
  * emulate=True rounds every tensor the GPU stores in BF16 at the same point (inputs,
    activations, output gradients, weight images) and computes in float64 otherwise, so that
    it differs from the GPU only by FP32 accumulation and the roundings that follows;
  * emulate=False is plain float64, used to pin the backward pass by finite differences.
The network has no counterpart in the reference (it is the build's learned form), so this
restatement is pinned by finite differences of its own forward pass, not by the reference.
"""
import numpy as np

H, KX, KE, NOUT = 64, 32, 32, 3
KM, KU = 2 * H + KE, 2 * H
NIN, NEI = 7, 15  # node and edge input columns in use (the rest of KX, KE are zero padding)


def layout():
    """parameter offsets, in the order of the device layout"""
    off, o = {}, 0
    for name, size in (("encW", KX * H), ("encB", H)):
        off[name] = (o, size)
        o += size
    for l in range(2):
        for name, size in ((f"mW{l}", KM * H), (f"mB{l}", H), (f"uW{l}", KU * H), (f"uB{l}", H)):
            off[name] = (o, size)
            o += size
    for name, size in (("decW", H * NOUT), ("decB", NOUT)):
        off[name] = (o, size)
        o += size
    return off, o


def bf16(x):
    """round to the nearest BF16 (ties to even), as the device's f2bf"""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


class Net:
    def __init__(self, params, emulate=True):
        self.off, self.n = layout()
        self.p = np.asarray(params, dtype=np.float64).copy()
        self.emulate = emulate

    def _r(self, x):
        return bf16(x) if self.emulate else x

    def W(self, name, shape):
        o, s = self.off[name]
        return self.p[o:o + s].reshape(shape)

    def Wimg(self, name, shape):
        # GEMM weights enter the MFMA as BF16 images of the FP32 masters
        return self._r(self.W(name, shape))

    def forward(self, x, rowOff, send, ef):
        N, E = len(rowOff) - 1, len(send)
        recv = np.repeat(np.arange(N), np.diff(rowOff))
        X = np.zeros((N, KX))
        X[:, :x.shape[1]] = self._r(x)
        F = np.zeros((E, KE))
        F[:, :ef.shape[1]] = self._r(ef)
        c = dict(X=X, F=F, recv=recv, send=send, N=N, E=E)
        c["h"] = [self._r(np.maximum(X @ self.Wimg("encW", (KX, H)) + self.W("encB", (H,)), 0.))]
        c["m"], c["agg"], c["u"], c["Z"] = [], [], [], []
        for l in range(2):
            h = c["h"][l]
            Z = np.concatenate([h[recv], h[send], F], axis=1)
            m = self._r(np.maximum(Z @ self.Wimg(f"mW{l}", (KM, H)) + self.W(f"mB{l}", (H,)), 0.))
            agg = np.zeros((N, H))
            np.add.at(agg, recv, m)
            agg = self._r(agg)
            U = np.concatenate([h, agg], axis=1)
            u = self._r(np.maximum(U @ self.Wimg(f"uW{l}", (KU, H)) + self.W(f"uB{l}", (H,)), 0.))
            c["Z"].append(Z)
            c["m"].append(m)
            c["agg"].append(agg)
            c["u"].append(u)
            c["h"].append(self._r(h + u))
        a = c["h"][2] @ self.W("decW", (H, NOUT)) + self.W("decB", (NOUT,))
        c["a"] = a
        return a, c

    def contact(self, c):
        """nodes in contact: a positive floor overlap input (column 4) or a positive overlap input
        (column 3) on one of their edges; the others move in free flight and carry no loss"""
        flag = c["X"][:, 4] > 0
        hit = np.zeros(c["N"], dtype=bool)
        np.logical_or.at(hit, c["recv"], c["F"][:, 3] > 0)
        return flag | hit

    def loss_grad(self, x, rowOff, send, rev, ef, t):
        a, c = self.forward(x, rowOff, send, ef)
        N, recv = c["N"], c["recv"]
        d = (a - t) * self.contact(c)[:, None]
        loss = float(np.sum(d * d) / N)
        dA = 2 * d / N
        g = np.zeros(self.n)

        def put(name, val):
            o, s = self.off[name]
            g[o:o + s] = val.reshape(-1)

        h2 = c["h"][2]
        put("decW", h2.T @ dA)
        put("decB", dA.sum(0))
        dh = dA @ self.W("decW", (H, NOUT)).T
        for l in (1, 0):
            h, agg, u, m, Z = c["h"][l], c["agg"][l], c["u"][l], c["m"][l], c["Z"][l]
            dYu = self._r(dh * (u > 0))
            U = np.concatenate([h, agg], axis=1)
            put(f"uW{l}", U.T @ dYu)
            put(f"uB{l}", dYu.sum(0))
            dU = dYu @ self.Wimg(f"uW{l}", (KU, H)).T
            dYm = self._r(dU[recv, H:] * (m > 0))
            put(f"mW{l}", Z.T @ dYm)
            put(f"mB{l}", dYm.sum(0))
            dZ = dYm @ self.Wimg(f"mW{l}", (KM, H))[:KU].T
            dhn = dh + dU[:, :H]
            np.add.at(dhn, recv, dZ[:, :H])
            np.add.at(dhn, c["send"], dZ[:, H:])
            dh = dhn
        dY0 = self._r(dh * (c["h"][0] > 0))
        put("encW", c["X"].T @ dY0)
        put("encB", dY0.sum(0))
        return loss, g
