"""Pins the CPU restatement of the warm-start network (oracle/warmnet.py).

The convolution is checked against scipy.ndimage.correlate, pooling and
upsampling against brute-force loops and their adjoints by the dot-product
identity, bfloat16 rounding against torch's conversion, and the analytic
backward pass against central finite differences (float64, no rounding).
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import torch

from oracle import warmnet as W


def test_bf16_rounding_matches_torch_and_ties_to_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 4000),
                        np.float32([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0, -0.0, 3.0e38])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(W.bf16(x), ref)
    assert W.bf16(np.float32([1 + 2 ** -8]))[0] == 1.0           # tie -> even
    assert W.bf16(np.float32([1 + 3 * 2 ** -8]))[0] == 1 + 2 ** -6


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    X, Y, Z, ci, co = 5, 4, 3, 3, 2
    x = rng.standard_normal((X * Y * Z, ci))
    w = rng.standard_normal((co, 27, ci))
    out = W.conv(x, w, (X, Y, Z))
    x3 = x.reshape(Z, Y, X, ci)
    for o in range(co):
        ref = sum(ndi.correlate(x3[..., c], w[o, :, c].reshape(3, 3, 3), mode="constant", cval=0.0)
                  for c in range(ci))
        np.testing.assert_allclose(out[:, o], ref.ravel(), rtol=1e-12, atol=1e-12)


def test_conv_adjoints_by_dot_product():
    rng = np.random.default_rng(2)
    shape, ci, co = (4, 3, 5), 3, 4
    V = 60
    x, g = rng.standard_normal((V, ci)), rng.standard_normal((V, co))
    w = rng.standard_normal((co, 27, ci))
    lhs = (W.conv(x, w, shape) * g).sum()
    assert lhs == pytest.approx((x * W.conv_adjoint(g, w, shape)).sum(), rel=1e-12)
    assert lhs == pytest.approx((w * W.conv_wgrad(g, x, shape)).sum(), rel=1e-12)


def test_pool_and_upsample_brute_force_and_adjoints():
    rng = np.random.default_rng(3)
    shape = (5, 3, 4)
    X, Y, Z = shape
    X2, Y2, Z2 = 3, 2, 2
    x = rng.standard_normal((X * Y * Z, 2))
    p = W.pool(x, shape, np.float64)
    for cz in range(Z2):
        for cy in range(Y2):
            for cx in range(X2):
                kids = [((z * Y + y) * X + xx) for z in (2 * cz, 2 * cz + 1) for y in (2 * cy, 2 * cy + 1)
                        for xx in (2 * cx, 2 * cx + 1) if xx < X and y < Y and z < Z]
                np.testing.assert_allclose(p[(cz * Y2 + cy) * X2 + cx], x[kids].mean(axis=0), rtol=1e-14)
    u = W.upsample(p, shape)
    for z in range(Z):
        for y in range(Y):
            for xx in range(X):
                assert np.array_equal(u[(z * Y + y) * X + xx], p[((z // 2) * Y2 + y // 2) * X2 + xx // 2])
    gc = rng.standard_normal(p.shape)
    gf = rng.standard_normal(x.shape)
    assert (p * gc).sum() == pytest.approx((x * W.pool_adjoint(gc, shape, np.float64)).sum(), rel=1e-12)
    assert (u * gf).sum() == pytest.approx((p * W.upsample_adjoint(gf, shape, np.float64)).sum(), rel=1e-12)


def _params(rng, scale=0.3):
    return rng.standard_normal(W.parameter_count()) * scale


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    X, Y, Z, nin, nout = 5, 4, 3, 5, 3
    net = W.Net(X, Y, Z, nin, nout, mode="exact")
    p = _params(rng)
    x = rng.standard_normal((X * Y * Z, nin))
    t = rng.standard_normal((X * Y * Z, nout))
    w = rng.uniform(0.1, 2.0, X * Y * Z)
    loss, g = net.gradients(p, x, t, w)
    assert loss == pytest.approx(net.loss(p, x, t, w), rel=1e-14)
    # every layer's weights and biases, with a parameter that carries signal
    offs, off = [], 0
    for ci, co, _ in W.LAYERS:
        offs.append((off, co * 27 * ci, co))
        off += co * 27 * ci + co
    picks = []
    for wo, nw, nb in offs:
        picks += list(wo + rng.choice(nw, 6, replace=False)) + list(wo + nw + rng.choice(nb, 2, replace=False))
    h = 1e-6
    for i in picks:
        e = np.zeros_like(p)
        e[i] = h
        fd = (net.loss(p + e, x, t, w) - net.loss(p - e, x, t, w)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=2e-5, abs=1e-9 * max(1.0, abs(loss))), i


def test_bf16_mode_tracks_exact_mode():
    rng = np.random.default_rng(5)
    X, Y, Z, nin, nout = 6, 5, 4, 8, 3
    p = _params(rng, 0.2).astype(np.float32)
    x = rng.standard_normal((X * Y * Z, nin)).astype(np.float32)
    t = rng.standard_normal((X * Y * Z, nout)).astype(np.float32)
    le, ge = W.Net(X, Y, Z, nin, nout, "exact").gradients(p, x, t)
    lb, gb = W.Net(X, Y, Z, nin, nout, "bf16").gradients(p, x, t)
    assert lb == pytest.approx(le, rel=2e-2)
    assert np.linalg.norm(gb - ge) <= 5e-2 * np.linalg.norm(ge)


def test_adam_first_step_moves_by_learning_rate():
    # bias correction makes the first step lr * sign(g) (up to eps)
    g = np.float32([0.5, -2.0, 1e-3])
    p, m, v = W.adam(np.zeros(3, np.float32), g, np.zeros(3, np.float32), np.zeros(3, np.float32), 1, 1e-2)
    np.testing.assert_allclose(p, -1e-2 * np.sign(g), rtol=1e-4)
