"""GPU parity of the warm-start network against its CPU restatement, and what
the learned warm start gains inside an OC run.

The oracle rounds the same operands to bfloat16 but accumulates exactly, so
outputs and gradients agree to a tolerance (stated per test) set by the
device's FP32 accumulation and the occasional bfloat16 rounding flip that
follows from it.
"""
import numpy as np
import pytest

from oracle import warmnet as W
from sparlab_amd import model as M
from sparlab_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    print("gpu warmnet: creating context", flush=True)
    c = N.Context(0)
    yield c
    c.close()


def _case(shape, nin, nout, seed):
    rng = np.random.default_rng(seed)
    V = shape[0] * shape[1] * shape[2]
    x = rng.standard_normal((V, nin)).astype(np.float32)
    t = (0.1 * rng.standard_normal((V, nout))).astype(np.float32)
    return x, t


def _layer_slices():
    out, off = [], 0
    for ci, co, _ in W.LAYERS:
        out.append(slice(off, off + co * 27 * ci + co))
        off += co * 27 * ci + co
    return out


@pytest.mark.parametrize("shape", [(9, 7, 5), (17, 9, 1), (16, 8, 8)])
def test_forward_matches_oracle(ctx, shape):
    print(f"test_forward_matches_oracle[{shape}]", flush=True)
    nin, nout = 8, 3
    net = N.WarmNet(*shape, nin, nout, seed=11, ctx=ctx)
    p = net.parameters()
    p[:] += 0.05 * np.random.default_rng(1).standard_normal(p.size).astype(np.float32)  # nonzero biases
    net.set_parameters(p)
    x, _ = _case(shape, nin, nout, 2)
    yg = net.forward(x)
    yo = W.Net(*shape, nin, nout, "bf16").forward(p, x)
    # bf16 activations: a rounding flip moves a value by 2^-8 relative
    assert np.linalg.norm(yg - yo) <= 1e-2 * np.linalg.norm(yo)
    np.testing.assert_allclose(yg, yo, rtol=0, atol=3e-2 * np.abs(yo).max())


@pytest.mark.parametrize("shape,weighted", [((9, 7, 5), False), ((16, 8, 8), True)])
def test_gradients_match_oracle(ctx, shape, weighted):
    print(f"test_gradients_match_oracle[{shape}-{weighted}]", flush=True)
    nin, nout = 8, 3
    net = N.WarmNet(*shape, nin, nout, seed=12, ctx=ctx)
    p = net.parameters()
    p[:] += 0.05 * np.random.default_rng(3).standard_normal(p.size).astype(np.float32)
    net.set_parameters(p)
    x, t = _case(shape, nin, nout, 4)
    w = np.random.default_rng(5).uniform(1e-3, 1.0, x.shape[0]).astype(np.float32) if weighted else None
    lg, gg = net.gradients(x, t, w)
    lo, go = W.Net(*shape, nin, nout, "bf16").gradients(p, x, t, w)
    assert lg == pytest.approx(lo, rel=5e-3)
    for li, sl in enumerate(_layer_slices()):
        err = np.linalg.norm(gg[sl] - go[sl])
        assert err <= 3e-2 * np.linalg.norm(go[sl]), f"layer {li + 1}: {err / np.linalg.norm(go[sl]):.3e}"


def test_adam_step_and_training(ctx):
    print("test_adam_step_and_training", flush=True)
    shape, nin, nout = (12, 8, 6), 8, 3
    net = N.WarmNet(*shape, nin, nout, seed=13, ctx=ctx)
    x, t = _case(shape, nin, nout, 5)
    p0 = net.parameters()
    _, g = net.gradients(x, t)
    net.train(x, t, 1, lr=1e-3)
    p1 = net.parameters()
    z = np.zeros_like(p0)
    pe, _, _ = W.adam(p0, g, z, z, 1, 1e-3)
    np.testing.assert_allclose(p1, pe, rtol=1e-5, atol=1e-8)
    first, last = net.train(x, t, 60, lr=3e-3)
    assert last < 0.5 * first


def test_learned_warm_start_in_an_oc_run(ctx):
    """The same OC run from the plain warm start (u_prev, as the reference) and
    from the learned one. What the trained network gains is measured in the
    K-norm distance of the starting point from the converged solution, relative
    to u_prev's: for the network's direction alone, for the previous
    displacement change alone, and for the Galerkin combination CG starts from.
    The compliances of both runs agree (CG runs to the same tolerance)."""
    print("test_learned_warm_start_in_an_oc_run", flush=True)
    m = M.cantilever_3d(24, 12, 12, volume_fraction=0.3, radius_cells=1.5)
    iters, train = 30, 12
    plain = N.HipProblem(m, ctx=ctx)
    xp, hp, _, fp = plain.run_oc(m.initial_design(), iters, change_tol=0.0, tol=1e-8)
    ctx2 = N.Context(0)
    try:
        learned = N.HipProblem(m, ctx=ctx2)
        learned.warm_start(True, train_solves=train, min_samples=3, steps_per_solve=24, buffer=8,
                           learning_rate=2e-3)
        xl, hl, _, fl = learned.run_oc(m.initial_design(), iters, change_tol=0.0, tol=1e-8)
        st = learned.warm_start_stats()
    finally:
        ctx2.close()
    k = st["predictions"]
    ratios = {name: st[f"{name}_ratio_sum"] / k for name in ("energy", "net", "secant")}
    print("plain iterations", hp[:, 11].astype(int).tolist(), flush=True)
    print("learned iterations", hl[:, 11].astype(int).tolist(), flush=True)
    print("warm-start stats", st, "mean K-norm ratios", ratios, flush=True)
    assert st["samples"] == train and st["train_steps"] == train * 24
    assert st["last_loss"] < st["first_loss"]
    assert k >= iters - 3 and st["rejected"] == 0
    assert ratios["net"] < 1.0                             # the trained direction alone beats u_prev
    assert ratios["energy"] <= ratios["secant"] + 1e-9     # and never hurts the combination
    assert ratios["energy"] < 0.8
    assert hl[train + 1:, 11].sum() <= hp[train + 1:, 11].sum()
    np.testing.assert_allclose(hl[:, 2], hp[:, 2], rtol=1e-4)
    assert fl == pytest.approx(fp, rel=1e-4)


def test_learned_warm_start_at_full_size():
    """The learned warm start's workload at its size: Hex8 96 x 48 x 48, 200 OC
    iterations at rtol 1e-8 on the mixed hierarchy, the network trained on the
    first 40 solves. Three runs: the plain warm start (the previous solution, as
    the reference), the previous solution plus the K-optimal multiple of the
    previous change (the method without the network), and the learned start.
    All three reach the same designs (CG runs to the same tolerance); the
    starting point CG gets is never worse in the K-norm than the method without
    the network; what the network adds over it is stated."""
    print("test_learned_warm_start_at_full_size", flush=True)
    m = M.cantilever_3d(96, 48, 48, volume_fraction=0.3, radius_cells=2.0)
    runs = {}
    for label in ("plain", "previous_change", "learned"):
        c = N.Context(0)
        g = N.HipProblem(m, ctx=c)
        x0 = m.initial_design()
        g.assemble(1e-9 + (1 - 1e-9) * np.clip(g.filter(x0), 0, 1) ** 3)
        g.amg_setup(precision=1)
        if label == "learned":
            g.warm_start(True, train_solves=40, min_samples=4, steps_per_solve=16, buffer=8, learning_rate=1e-3)
        elif label == "previous_change":
            g.warm_start(True, train_solves=0, min_samples=1 << 30)
        x, h, done, comp = g.run_oc(x0, 200, change_tol=0.0, tol=1e-8)
        st = g.warm_start_stats() if label != "plain" else None
        runs[label] = (x, h, comp, st)
        print(f"{label}: {int(h[40:, 11].sum())} CG iterations after iteration 40, c = {comp!r}", flush=True)
        c.close()
    xp, hp, cp, _ = runs["plain"]
    for label in ("previous_change", "learned"):
        x, h, comp, st = runs[label]
        assert comp == pytest.approx(cp, rel=1e-7)
        np.testing.assert_allclose(h[:, 2], hp[:, 2], rtol=1e-6)
        assert st["predictions"] >= 195 and st["rejected"] == 0
    _, hs, _, ss = runs["previous_change"]
    _, hl, _, sl = runs["learned"]
    k = sl["predictions"]
    comb, net, sec = (sl[f"{n}_ratio_sum"] / k for n in ("energy", "net", "secant"))
    print(f"K-norm error of the start / the previous solution's: combination {comb:.4f}, network alone {net:.4f}, "
          f"previous change alone {sec:.4f}", flush=True)
    assert sl["samples"] == 40 and sl["train_steps"] == 40 * 16
    assert comb <= sec + 1e-9 and net < 1.0 and sec < 0.6
    assert hs[40:, 11].sum() < 0.95 * hp[40:, 11].sum()
    assert hl[40:, 11].sum() <= 1.02 * hs[40:, 11].sum()
