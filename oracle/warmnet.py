"""CPU restatement of the warm-start network (test infrastructure only).

The network has no counterpart in the reference: it proposes the initial guess
of the compliance solves that the reference warm-starts from the previous
solution (src/topopt/Sensitivity.cpp:169-176). This module restates the
device network of sparlab_amd/csrc/warmnet.cpp so the tests can check the GPU
forward pass and every parameter gradient.

Two arithmetic modes:
  * ``bf16``: the device's rounding points -- convolution operands (inputs,
    weights, activations, output gradients) rounded to bfloat16 (round to
    nearest even), products accumulated exactly (float64) and the epilogues in
    float32; the device accumulates in FP32, so results agree to a tolerance
    set by FP32 accumulation and the occasional bfloat16 rounding flip;
  * ``exact``: float64 throughout, no rounding; used to pin the analytic
    backward pass against central finite differences.

Pinning (tests/test_warmnet_oracle.py): the convolution against
scipy.ndimage.correlate, pooling and upsampling against brute-force loops and
their adjoints by the dot-product identity, and the whole backward pass
against finite differences in exact mode.

Layout: voxel v = (z * Y + y) * X + x, tensors [V][C]; weights [cout][27][cin]
with tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); layers (cin, cout, level) =
(16,16,1) (16,32,2) (32,64,3) (64,32,2) (32,16,1) (16,16,1).
"""
from __future__ import annotations

import numpy as np

LAYERS = ((16, 16, 0), (16, 32, 1), (32, 64, 2), (64, 32, 1), (32, 16, 0), (16, 16, 0))
C_IN = 16


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def parameter_count():
    return sum(co * 27 * ci + co for ci, co, _ in LAYERS)


def unpack(params):
    """Flat FP32 vector -> [(W (cout, 27, cin), b (cout,))] per layer."""
    out, off = [], 0
    for ci, co, _ in LAYERS:
        w = np.asarray(params[off:off + co * 27 * ci]).reshape(co, 27, ci)
        off += co * 27 * ci
        b = np.asarray(params[off:off + co])
        off += co
        out.append((w, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def level_shapes(X, Y, Z):
    s = [(X, Y, Z)]
    for _ in range(2):
        x, y, z = s[-1]
        s.append(((x + 1) // 2, (y + 1) // 2, (z + 1) // 2))
    return s


def _taps():
    for t in range(27):
        yield t, t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def _shifted(x, shape, dx, dy, dz):
    """x[nbr(v)] with zero outside the grid, as [V][C]."""
    X, Y, Z = shape
    C = x.shape[1]
    xp = np.zeros((Z + 2, Y + 2, X + 2, C), dtype=x.dtype)
    xp[1:-1, 1:-1, 1:-1] = x.reshape(Z, Y, X, C)
    return xp[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X].reshape(-1, C)


def conv(x, w, shape):
    """out[v][co] = sum_{tap, ci} x[v + tap][ci] w[co][tap][ci], float64."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    out = np.zeros((x.shape[0], w.shape[0]))
    for t, dx, dy, dz in _taps():
        out += _shifted(x, shape, dx, dy, dz) @ w[:, t, :].T
    return out


def conv_adjoint(g, w, shape):
    """The adjoint of conv with respect to its input: flipped taps, transposed channels."""
    g = np.asarray(g, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    out = np.zeros((g.shape[0], w.shape[2]))
    for t, dx, dy, dz in _taps():
        out += _shifted(g, shape, -dx, -dy, -dz) @ w[:, t, :]
    return out


def conv_wgrad(g, x, shape):
    """dW[co][tap][ci] = sum_v g[v][co] x[v + tap][ci]."""
    g = np.asarray(g, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    dw = np.zeros((g.shape[1], 27, x.shape[1]))
    for t, dx, dy, dz in _taps():
        dw[:, t, :] = g.T @ _shifted(x, shape, dx, dy, dz)
    return dw


def _children(shape):
    """For 2x2x2 pooling: per (dz, dy, dx) the fine indices of each coarse voxel (or -1)."""
    X, Y, Z = shape
    X2, Y2, Z2 = (X + 1) // 2, (Y + 1) // 2, (Z + 1) // 2
    cz, cy, cx = np.meshgrid(np.arange(Z2), np.arange(Y2), np.arange(X2), indexing="ij")
    out = []
    for dz in range(2):
        for dy in range(2):
            for dx in range(2):
                x, y, z = 2 * cx + dx, 2 * cy + dy, 2 * cz + dz
                ok = (x < X) & (y < Y) & (z < Z)
                out.append(np.where(ok, (z * Y + y) * X + x, -1).ravel())
    return out


def pool(x, shape, dtype=np.float32):
    ch = _children(shape)
    s = np.zeros((ch[0].size, x.shape[1]), dtype=dtype)
    cnt = np.zeros(ch[0].size, dtype=dtype)
    for idx in ch:
        ok = idx >= 0
        s[ok] += np.asarray(x, dtype=dtype)[idx[ok]]
        cnt[ok] += 1
    return s / cnt[:, None]


def pool_adjoint(gc, shape, dtype=np.float32):
    ch = _children(shape)
    cnt = np.zeros(ch[0].size, dtype=dtype)
    for idx in ch:
        cnt[idx >= 0] += 1
    X, Y, Z = shape
    gf = np.zeros((X * Y * Z, gc.shape[1]), dtype=dtype)
    q = np.asarray(gc, dtype=dtype) / cnt[:, None]
    for idx in ch:
        ok = idx >= 0
        gf[idx[ok]] = q[ok]
    return gf


def parent(shape):
    X, Y, Z = shape
    X2, Y2 = (X + 1) // 2, (Y + 1) // 2
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    return ((z // 2) * Y2 + y // 2) * X2 + x // 2


def upsample(xc, shape):
    return np.asarray(xc)[parent(shape).ravel()]


def upsample_adjoint(gf, shape, dtype=np.float32):
    ch = _children(shape)
    s = np.zeros((ch[0].size, gf.shape[1]), dtype=dtype)
    for idx in ch:
        ok = idx >= 0
        s[ok] += np.asarray(gf, dtype=dtype)[idx[ok]]
    return s


class Net:
    """Forward and backward passes of the device network on one grid."""

    def __init__(self, X, Y, Z, nin, nout, mode="bf16"):
        assert mode in ("bf16", "exact")
        self.shapes = level_shapes(X, Y, Z)
        self.V = X * Y * Z
        self.nin, self.nout, self.mode = nin, nout, mode
        self.ft = np.float32 if mode == "bf16" else np.float64

    def _r(self, a):
        """Round a convolution operand the way the device stores it."""
        return bf16(a) if self.mode == "bf16" else np.asarray(a, dtype=np.float64)

    def _layer(self, li, x, w, b, add=None, relu=True):
        shape = self.shapes[LAYERS[li][2]]
        z = conv(x, self._r(w), shape).astype(self.ft) + np.asarray(b, dtype=self.ft)
        if add is not None:
            z = z + np.asarray(add, dtype=self.ft)
        if relu:
            z = np.maximum(z, 0)
        return z

    def forward(self, params, x):
        L = unpack(params)
        x16 = np.zeros((self.V, C_IN), dtype=self.ft)
        x16[:, :self.nin] = np.asarray(x, dtype=self.ft).reshape(self.V, self.nin)
        a = {"x": self._r(x16)}
        s1, s2, _ = self.shapes
        a["e1"] = self._r(self._layer(0, a["x"], *L[0]))
        a["p1"] = self._r(pool(a["e1"], s1, self.ft))
        a["e2"] = self._r(self._layer(1, a["p1"], *L[1]))
        a["p2"] = self._r(pool(a["e2"], s2, self.ft))
        a["e3"] = self._r(self._layer(2, a["p2"], *L[2]))
        a["u3"] = upsample(a["e3"], s2)
        a["d2"] = self._r(self._layer(3, a["u3"], *L[3], add=a["e2"]))
        a["u2"] = upsample(a["d2"], s1)
        a["d1"] = self._r(self._layer(4, a["u2"], *L[4], add=a["e1"]))
        a["y"] = self._layer(5, a["d1"], *L[5], relu=False)
        self.acts = a
        return a["y"][:, :self.nout]

    def loss(self, params, x, t, w=None):
        """sum_v w_v sum_k (y - t)^2 / (nout sum_v w_v), w = 1 by default."""
        y = self.forward(params, x)
        d = np.asarray(y, dtype=np.float64) - np.asarray(t, dtype=np.float64).reshape(self.V, self.nout)
        if w is None:
            return float((d * d).sum() / (self.V * self.nout))
        w = np.asarray(w, dtype=np.float64).reshape(self.V)
        return float((w[:, None] * d * d).sum() / (w.sum() * self.nout))

    def gradients(self, params, x, t, w=None):
        """Loss and d loss / d params (flat, parameter order)."""
        L = unpack(params)
        loss = self.loss(params, x, t, w)
        a, ft = self.acts, self.ft
        s1, s2, s3 = self.shapes
        t = np.asarray(t, dtype=ft).reshape(self.V, self.nout)
        gy = np.zeros((self.V, 16), dtype=ft)
        if w is None:
            gy[:, :self.nout] = ft(2) * (a["y"][:, :self.nout] - t) / (ft(self.V) * ft(self.nout))
        else:
            wv = np.asarray(w, dtype=ft).reshape(self.V, 1)
            W = ft(np.asarray(w, dtype=np.float64).sum())
            gy[:, :self.nout] = ft(2) * wv * (a["y"][:, :self.nout] - t) / (W * ft(self.nout))
        grads = [None] * 6

        def back(li, gz, inp):
            """weight / bias gradients of layer li; returns the input gradient (float)."""
            w = L[li][0]
            shape = self.shapes[LAYERS[li][2]]
            gz_r = self._r(gz)
            grads[li] = (conv_wgrad(gz_r, inp, shape), np.asarray(gz, dtype=np.float64).sum(axis=0))
            return conv_adjoint(gz_r, self._r(w), shape).astype(ft)

        def relu_adj(g, act):
            return np.where(act > 0, g, 0).astype(ft)

        g_d1 = back(5, gy, a["d1"])
        g_z5 = relu_adj(g_d1, a["d1"])
        g_u2 = back(4, g_z5, a["u2"])
        g_d2 = upsample_adjoint(g_u2, s1, ft)
        g_z4 = relu_adj(g_d2, a["d2"])
        g_u3 = back(3, g_z4, a["u3"])
        g_e3 = upsample_adjoint(g_u3, s2, ft)
        g_z3 = relu_adj(g_e3, a["e3"])
        g_p2 = back(2, g_z3, a["p2"])
        g_e2 = pool_adjoint(g_p2, s2, ft) + g_z4
        g_z2 = relu_adj(g_e2, a["e2"])
        g_p1 = back(1, g_z2, a["p1"])
        g_e1 = pool_adjoint(g_p1, s1, ft) + g_z5
        g_z1 = relu_adj(g_e1, a["e1"])
        back(0, g_z1, a["x"])
        return loss, pack(grads)


def adam(params, grads, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected), float32 like the device."""
    f = np.float32
    m = f(b1) * m + f(1 - b1) * grads
    v = f(b2) * v + f(1 - b2) * grads * grads
    c1, c2 = f(1 - b1 ** step), f(1 - b2 ** step)
    return params - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps)), m, v
