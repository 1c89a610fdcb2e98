"""slx_kff_apply_noalloc (no device allocation) against slx_kff_apply, through the
C ABI: the same product bit for bit, x.Kx against the host dot of the product, on
two calls in a row (the second on a smaller system), on the matrix-free Hex8
operator and on an assembled Q4 CSR, and the size limit's error."""
import numpy as np
import pytest

from sparlab_amd import model as M
from sparlab_amd import native as N

pytestmark = pytest.mark.gpu


def _check(g, seed):
    x = np.random.default_rng(seed).standard_normal(g.nfree)
    y_ref, mf_ref = g.kff_apply(x)
    y, xkx, mf = g.kff_apply_noalloc(x)
    assert mf == mf_ref
    np.testing.assert_array_equal(y, y_ref)
    assert xkx == pytest.approx(float(x @ y_ref), rel=1e-12)
    return mf


def test_noalloc_product_matches_the_allocating_entry_point():
    print("test_noalloc_product_matches_the_allocating_entry_point", flush=True)
    ctx = N.Context(0)
    # two calls in a row on one context: a larger Hex8 system, then a smaller one
    for k, (nx, ny, nz) in enumerate([(24, 12, 12), (12, 6, 5)]):
        m = M.cantilever_3d(nx, ny, nz)
        g = N.HipProblem(m, ctx=ctx)
        g.assemble(1e-9 + (1 - 1e-9) * M.seeded_density(m.nelem, seed=40 + k) ** 3)
        assert _check(g, 50 + k)
    # the assembled CSR path (Q4), again larger then smaller
    for k, (nx, ny) in enumerate([(120, 40), (30, 15)]):
        m = M.mbb_beam(nx, ny)
        g = N.HipProblem(m, ctx=ctx)
        g.assemble(M.seeded_density(m.nelem, seed=60 + k) ** 3)
        assert not _check(g, 70 + k)
    ctx.close()


def test_noalloc_refuses_systems_above_its_limit():
    print("test_noalloc_refuses_systems_above_its_limit", flush=True)
    ctx = N.Context(0)
    m = M.mbb_beam(1100, 1000)  # 2 204 101 free DOFs > 2 097 152
    g = N.HipProblem(m, ctx=ctx)
    g.assemble()
    assert g.nfree > 2 ** 21
    with pytest.raises(N.ConfigError):
        g.kff_apply_noalloc(np.ones(g.nfree))
    # a smaller system on the same context still works after the refusal
    m2 = M.mbb_beam(40, 20)
    g2 = N.HipProblem(m2, ctx=ctx)
    g2.assemble()
    _check(g2, 80)
    ctx.close()
