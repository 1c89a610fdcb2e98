"""The symmetric Gauss-Seidel smoother (src/fem/Multigrid.cpp:788-806) on the
device against the oracle's sequential sweeps.

The device sweeps a level's rows by dependency levels (rows of one level never
read each other's unknowns), which reproduces the sequential sweep's values;
what differs from the oracle is the rounding of the matrix-free level-0
residual and of the coarse solve, so the comparisons are to tolerance (stated
per test): one V-cycle, PCG solves, an OC run and the mixed hierarchies.
"""
import numpy as np
import pytest

from oracle.refcpu import Oracle
from sparlab_amd import model as M
from sparlab_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    print("gauss-seidel: creating context", flush=True)
    c = N.Context(0)
    yield c
    c.close()


def _model(case):
    # coarse sizes that give a three-level hierarchy on both
    if case == "hex":
        return M.cantilever_3d(16, 8, 8), 300
    return M.mbb_beam(60, 20), 100


@pytest.mark.parametrize("case", ["hex", "quad"])
def test_gauss_seidel_vcycle_and_pcg_against_oracle(ctx, case):
    print(f"test_gauss_seidel_vcycle_and_pcg_against_oracle[{case}]", flush=True)
    m, cs = _model(case)
    g = N.HipProblem(m, ctx=ctx)
    o = Oracle(m, tol=1e-12, amg=dict(coarse_size=cs, smoother=1))
    cheb = Oracle(m, tol=1e-12, amg=dict(coarse_size=cs, smoother=0))
    # the Hex8 box takes a 1000:1 contrast; on the Q4 beam that contrast leaves
    # PCG so ill-conditioned (over 120 iterations to 1e-12) that its iteration
    # count follows last-bit rounding, so the beam gets a 5:1 field
    u = np.random.default_rng(2).uniform(size=m.nelem)
    scale = np.where(u < 0.5, 1.0, 1e-3) if case == "hex" else 0.2 + 0.8 * u
    g.assemble(scale)
    o.assemble_kff(scale)
    cheb.assemble_kff(scale)
    g.amg_setup(coarse_size=cs, smoother=1)
    o.amg_setup()
    cheb.amg_setup()
    assert g.amg_stats().levels == o.amg_levels() >= 3
    r = np.random.default_rng(4).standard_normal(g.nfree)
    zg, zo, zc = g.amg_apply(r), o.amg_apply(r), cheb.amg_apply(r)
    assert np.linalg.norm(zg - zo) <= 1e-9 * np.linalg.norm(zo)
    # the cycle is the Gauss-Seidel one, not the Chebyshev one (they differ by
    # about 1e-3 relative on the beam, far above the agreement bound)
    assert np.linalg.norm(zc - zo) > 1e-6 * np.linalg.norm(zo)
    b = m.free_loads()[0]
    xg, itg, rrg, cvg = g.solve(b, tol=1e-12)
    xo, ito, rro, cvo = o.pcg(b, tol=1e-12)
    print(f"{case}: CG iterations device {itg}, oracle {ito}", flush=True)
    assert cvg and cvo and abs(itg - ito) <= 1
    assert np.linalg.norm(xg - xo) <= 1e-9 * np.linalg.norm(xo)
    assert b @ xg == pytest.approx(b @ xo, rel=1e-12)


def test_gauss_seidel_oc_run_against_oracle(ctx):
    print("test_gauss_seidel_oc_run_against_oracle", flush=True)
    m, cs = _model("quad")
    g = N.HipProblem(m, ctx=ctx)
    o = Oracle(m, tol=1e-10, amg=dict(coarse_size=cs, smoother=1))
    x0 = m.initial_design()
    g.assemble()
    g.amg_setup(coarse_size=cs, smoother=1)
    xg, hg, ng, fg = g.run_oc(x0, 10, change_tol=0.0, tol=1e-10)
    xo, ho, no, fo = o.run_oc(x0, 10, change_tol=0.0)
    assert ng == no == 10
    np.testing.assert_allclose(hg[:, 2], ho[:, 2], rtol=1e-8)
    np.testing.assert_allclose(hg[:, 4], ho[:, 4], rtol=1e-8)
    assert fg == pytest.approx(fo, rel=1e-7)
    assert np.abs(xg - xo).max() <= 1e-6


@pytest.mark.parametrize("precision", [1, 2])
def test_gauss_seidel_mixed_hierarchies_against_oracle(ctx, precision):
    """FP32 sweeps (bfloat16 coarse operators with precision 2) under the FP64
    flexible CG: the solution to the same tolerance as the oracle's FP64 solve."""
    print(f"test_gauss_seidel_mixed_hierarchies_against_oracle[{precision}]", flush=True)
    m, cs = _model("hex")
    g = N.HipProblem(m, ctx=ctx)
    o = Oracle(m, tol=1e-10, amg=dict(coarse_size=cs, smoother=1))
    scale = 0.1 + 0.9 * M.seeded_density(m.nelem, seed=9)
    g.assemble(scale)
    o.assemble_kff(scale)
    g.amg_setup(coarse_size=cs, smoother=1, precision=precision)
    o.amg_setup()
    b = m.free_loads()[0]
    xg, itg, rrg, cvg = g.solve(b, tol=1e-10)
    xo, ito, rro, cvo = o.pcg(b, tol=1e-10)
    print(f"precision {precision}: CG iterations device {itg}, oracle {ito}", flush=True)
    assert cvg and rrg <= 1e-10
    assert abs(itg - ito) <= 3
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)
    assert b @ xg == pytest.approx(b @ xo, rel=1e-9)


def test_gauss_seidel_option_errors(ctx):
    print("test_gauss_seidel_option_errors", flush=True)
    m, cs = _model("hex")
    g = N.HipProblem(m, ctx=ctx)
    g.assemble()
    with pytest.raises(N.ConfigError):
        g.amg_setup(smoother=2)
    with pytest.raises(N.ConfigError):
        g.amg_setup(smoother=1, coarse_reuse=2)
    g.amg_setup(coarse_size=cs, smoother=1)
    assert g.amg_stats().levels >= 3
