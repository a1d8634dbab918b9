"""CPU oracle of the learned surrogate (include/fdtd_surrogate.h), numpy.

Test infrastructure only: tests/ compare the device network against it.
It restates the network of fdtd_surrogate.h (no reference counterpart:
the reference has no learned part) in float64 with an optional emulation
of the device's BF16 operand rounding (every matrix-product operand, the
input voxels and the coordinates rounded to BF16, round to nearest even),
and its backward pass by hand.  Pinned by central finite differences of
its own float64 loss (tests/test_surrogate_oracle.py).
"""
import numpy as np
from math import erf, sqrt

NL = 4


def bf16(x):
    """Round to BF16 (nearest even) and back, as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


_erf = np.vectorize(erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / sqrt(2.0))) + z * np.exp(-0.5 * z * z) / sqrt(2.0 * np.pi)


class Config:
    def __init__(self, grid, in_channels, channels, coords, width, outputs, queries):
        self.grid, self.in_channels, self.channels = grid, in_channels, list(channels)
        self.coords, self.width, self.outputs, self.queries = coords, width, outputs, queries
        self.C = [in_channels] + list(channels)
        self.S = [grid // (2 ** l) for l in range(NL + 1)]
        self.F = self.S[NL] ** 3 * self.C[NL]

    def shapes(self):
        """Parameter shapes in the flat order of fdtd_surrogate.h."""
        out = []
        for l in range(NL):
            out += [(27 * self.C[l], self.C[l + 1]), (self.C[l + 1],)]
        w = self.width
        out += [(self.F, w), (self.coords, w), (w,), (w, w), (w,), (w, self.outputs), (self.outputs,)]
        return out

    def unflatten(self, flat):
        ps, off = [], 0
        for sh in self.shapes():
            n = int(np.prod(sh))
            ps.append(np.asarray(flat[off:off + n], dtype=np.float64).reshape(sh))
            off += n
        assert off == len(flat)
        return ps


def voxels(maps, S):
    """[n, S, S, C] (x, z) maps -> [n, S(x), S(y), S(z), C]."""
    return np.repeat(maps[:, :, None, :, :], S, axis=2)


def conv_fwd(x, W, b, r):
    """3x3x3, stride 2, padding 1 over [n, S, S, S, Cin]; W [27 Cin, Cout]."""
    n, S, _, _, cin = x.shape
    so = S // 2
    xp = np.zeros((n, S + 2, S + 2, S + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    out = np.zeros((n, so, so, so, W.shape[1]))
    Wr = r(W)
    for a in range(3):
        for bb in range(3):
            for c in range(3):
                q = 9 * a + 3 * bb + c
                patch = xp[:, a:a + 2 * so:2, bb:bb + 2 * so:2, c:c + 2 * so:2, :]
                out += patch @ Wr[q * cin:(q + 1) * cin]
    return out + b


def conv_bwd(x, W, dz, r):
    """Gradients of conv_fwd wrt W (with the rounded operands x and dz) and
    wrt x (dz and W rounded, as the device's products)."""
    n, S, _, _, cin = x.shape
    so = S // 2
    xp = np.zeros((n, S + 2, S + 2, S + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp)
    dzr, Wr = r(dz), r(W)
    for a in range(3):
        for bb in range(3):
            for c in range(3):
                q = 9 * a + 3 * bb + c
                patch = xp[:, a:a + 2 * so:2, bb:bb + 2 * so:2, c:c + 2 * so:2, :]
                dW[q * cin:(q + 1) * cin] = np.einsum("nxyzi,nxyzo->io", patch, dzr)
                dxp[:, a:a + 2 * so:2, bb:bb + 2 * so:2, c:c + 2 * so:2, :] += dzr @ Wr[q * cin:(q + 1) * cin].T
    db = r(dz).sum(axis=(0, 1, 2, 3))
    return dW, db, dxp[:, 1:-1, 1:-1, 1:-1]


def forward(cfg, params, maps, coords, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    ps = params
    n = maps.shape[0]
    x = r(voxels(np.asarray(maps, dtype=np.float64), cfg.grid))
    cache = {"x": [], "z": []}
    for l in range(NL):
        cache["x"].append(x)
        z = conv_fwd(x, ps[2 * l], ps[2 * l + 1], r)
        cache["z"].append(z)
        x = r(gelu(z))
    feat = x.reshape(n, cfg.F)
    W1f, W1c, b1, W2, b2, W3, b3 = ps[2 * NL:]
    P = cfg.queries
    c = r(np.asarray(coords, dtype=np.float64).reshape(n * P, cfg.coords))
    U = feat @ r(W1f)
    V = c @ r(W1c)
    h1z = np.repeat(U, P, axis=0) + V + b1
    h1 = r(gelu(h1z))
    h2z = h1 @ r(W2) + b2
    h2 = r(gelu(h2z))
    y = h2 @ r(W3) + b3
    cache.update(feat=feat, c=c, h1z=h1z, h1=h1, h2z=h2z, h2=h2)
    return y.reshape(n, P, cfg.outputs), cache


def loss_and_grad(cfg, params, maps, coords, labels, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    y, k = forward(cfg, params, maps, coords, emulate_bf16)
    n, P = y.shape[0], cfg.queries
    d = (y - labels).reshape(n * P, cfg.outputs)
    loss = np.mean(d * d)
    dy = r(2.0 * d / d.size)
    W1f, W1c, b1, W2, b2, W3, b3 = params[2 * NL:]
    g = [None] * len(params)
    g[2 * NL + 5] = k["h2"].T @ dy
    g[2 * NL + 6] = dy.sum(axis=0)
    dz2 = r(dy @ r(W3).T * gelu_grad(k["h2z"]))
    g[2 * NL + 3] = k["h1"].T @ dz2
    g[2 * NL + 4] = dz2.sum(axis=0)
    dz1f = (dz2 @ r(W2).T) * gelu_grad(k["h1z"])
    dz1 = r(dz1f)
    g[2 * NL + 1] = k["c"].T @ dz1
    g[2 * NL + 2] = dz1.sum(axis=0)
    dU = r(dz1f.reshape(n, P, -1).sum(axis=1))
    g[2 * NL] = k["feat"].T @ dU
    dact = (dU @ r(W1f).T).reshape(k["z"][NL - 1].shape)
    for l in range(NL - 1, -1, -1):
        dz = dact * gelu_grad(k["z"][l])
        dW, db, dx = conv_bwd(k["x"][l], params[2 * l], dz, r)
        g[2 * l], g[2 * l + 1] = dW, db
        dact = dx
    return loss, g


def flatten(arrs):
    return np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in arrs])
