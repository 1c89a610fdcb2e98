"""GPU parity of the HIP path against the refcpu oracle, through the C ABI.

Bit-exact: Ke0, the DOF map, the K_ff pattern and values, the density filter,
the aggregates of every level. To tolerance (stated per test): the multigrid
operators, PCG solutions and compliances, sensitivities, OC updates and the
OC loop. Sizes are ones the oracle finishes in seconds.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.refcpu import Oracle
from sparlab_amd import model as M
from sparlab_amd import native as N

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))


def _csr(rows, cols, ptr, col, val):
    return sp.csr_matrix((val, col, ptr), shape=(rows, cols))


@pytest.fixture(scope="module")
def ctx():
    print("gpu parity: creating context", flush=True)
    c = N.Context(0)
    yield c
    c.close()


def _pair(m, ctx, tol=1e-8, **amg):
    return N.HipProblem(m, ctx=ctx), Oracle(m, tol=tol, amg=amg or None)


def _partly_fixed_box():
    """Hex8 13 x 7 x 5 with nodes of every free-component mask: the clamped
    face, rollers (z fixed) on the bottom face, y fixed on a side strip and x
    and z on a few nodes inside."""
    m = M.cantilever_3d(13, 7, 5)
    m.fix(m.select_box(zmax=0.0), [2])
    m.fix(m.select_box(ymax=0.0, xmin=4.0, xmax=9.0), [1])
    m.fix(m.select_box(xmin=6.0, xmax=7.0, ymin=3.0, ymax=4.0, zmin=2.0, zmax=3.0), [0, 2])
    return m


@pytest.mark.parametrize("case", ["hex", "hex_partly_fixed", "quad"])
def test_ke0_dofs_pattern_and_assembly_bitwise(ctx, case, monkeypatch):
    """K_ff against the oracle bit for bit; Hex8 through both device kernels (node
    blocks, and per entry with SPARLAB_ROW_ASSEMBLY=1)."""
    print(f"test_ke0_dofs_pattern_and_assembly_bitwise[{case}]", flush=True)
    m = {"hex": lambda: M.cantilever_3d(12, 6, 5), "hex_partly_fixed": _partly_fixed_box,
         "quad": lambda: M.mbb_beam(30, 15)}[case]()
    g, o = _pair(m, ctx)
    assert np.array_equal(g.ke0(), o.ke0())
    assert np.array_equal(g.free(), o.free())
    scale = M.seeded_density(m.nelem, seed=3) ** 3
    g.assemble(scale)
    gp, gc, gv = g.kff()
    op_, oc, ov = o.assemble_kff(scale)
    assert np.array_equal(gp, op_) and np.array_equal(gc, oc)
    assert np.array_equal(gv, ov)
    if case != "quad":
        monkeypatch.setenv("SPARLAB_ROW_ASSEMBLY", "1")
        g.assemble(scale)
        assert np.array_equal(g.kff()[2], ov)


def test_aggregates_and_hierarchy(ctx):
    print("test_aggregates_and_hierarchy", flush=True)
    m = M.cantilever_3d(16, 8, 8)
    g, o = _pair(m, ctx, coarse_size=300)
    scale = M.seeded_density(m.nelem, seed=5) ** 3
    g.assemble(scale)
    o.assemble_kff(scale)
    g.amg_setup(coarse_size=300)
    o.amg_setup()
    st = g.amg_stats()
    assert st.levels == o.amg_levels() >= 3
    for l in range(st.levels - 1):
        nn = len(o.amg_aggregates(l))
        assert np.array_equal(g.amg_aggregates(l, nn), o.amg_aggregates(l)), f"aggregates differ at level {l}"
        for w in ("Ptent", "P", "A"):
            G = _csr(*g.amg_matrix(l, w))
            R = _csr(*o.amg_matrix(l, w))
            assert G.shape == R.shape
            assert np.array_equal(G.indptr, R.indptr) and np.array_equal(G.indices, R.indices)
            if w == "Ptent":
                # the same column-pivoted QR on the same near-null-space rows: bit for bit
                assert np.array_equal(G.data, R.data), l
            # values: Lanczos dots and product orders differ at rounding level
            assert abs(G - R).max() <= 1e-10 * abs(R).max(), (l, w)
        assert g.amg_stats().lambda_max[l] == pytest.approx(o.amg_lambda(l), rel=1e-10)
    assert st.operator_complexity == pytest.approx(o.amg_operator_complexity(), rel=1e-15)


@pytest.mark.parametrize("row", [r for r in GOLDEN["hex8_bench_amg"] if 16 <= r["nx"] <= 48], ids=lambda r: f"nx{r['nx']}")
def test_bench_box_matches_published(ctx, row):
    print(f"test_bench_box_matches_published[{row['nx']}]", flush=True)
    m = M.bench_box(row["nx"])
    g = N.HipProblem(m, ctx=ctx)
    g.assemble()
    g.amg_setup()
    st = g.amg_stats()
    assert st.levels == row["levels"]
    assert st.operator_complexity == pytest.approx(row["operator_complexity"], rel=1e-15)
    b = m.free_loads()[0]
    x, it, rr, cv = g.solve(b, tol=1e-10)
    assert cv and rr <= 1e-10
    c = float(b @ x)
    assert abs(c - row["compliance"]) <= 1e-11 * row["compliance"]
    assert abs(it - row["iterations"]) <= 1


def test_vcycle_and_pcg_against_oracle(ctx):
    print("test_vcycle_and_pcg_against_oracle", flush=True)
    m = M.cantilever_3d(16, 8, 8)
    g, o = _pair(m, ctx, coarse_size=300)
    scale = np.where(np.random.default_rng(2).uniform(size=m.nelem) < 0.5, 1.0, 1e-3)
    g.assemble(scale)
    o.assemble_kff(scale)
    g.amg_setup(coarse_size=300)
    o.amg_setup()
    r = np.random.default_rng(4).standard_normal(g.nfree)
    zg, zo = g.amg_apply(r), o.amg_apply(r)
    assert np.linalg.norm(zg - zo) <= 1e-9 * np.linalg.norm(zo)
    b = m.free_loads()[0]
    xg, itg, rrg, cvg = g.solve(b, tol=1e-12)
    xo, ito, rro, cvo = o.pcg(b, tol=1e-12)
    assert cvg and cvo and abs(itg - ito) <= 1
    assert np.linalg.norm(xg - xo) <= 1e-9 * np.linalg.norm(xo)
    assert b @ xg == pytest.approx(b @ xo, rel=1e-12)


@pytest.mark.parametrize("case", ["hex", "quad"])
def test_filter_bitwise(ctx, case):
    print(f"test_filter_bitwise[{case}]", flush=True)
    m = M.cantilever_3d(10, 6, 5, radius_cells=2.0) if case == "hex" else M.mbb_beam(40, 20)
    g, o = _pair(m, ctx)
    x = np.random.default_rng(1).uniform(size=m.nelem)
    assert np.array_equal(g.filter(x), o.filter(x))
    assert np.array_equal(g.filter(x, adjoint=True), o.filter(x, adjoint=True))


@pytest.mark.parametrize("case", ["hex", "quad"])
def test_evaluate_and_oc_against_oracle(ctx, case):
    print(f"test_evaluate_and_oc_against_oracle[{case}]", flush=True)
    m = M.cantilever_3d(12, 6, 6) if case == "hex" else M.mbb_beam(60, 20)
    g, o = _pair(m, ctx, tol=1e-12)
    # a design whose volume target lies within the move limit
    x = m.volume_fraction * np.random.default_rng(7).uniform(0.5, 1.5, m.nelem)
    eg = g.evaluate(x, tol=1e-12)
    eo = o.evaluate(x)
    assert eg["compliance"] == pytest.approx(eo["compliance"], rel=1e-10)
    assert eg["volume"] == pytest.approx(eo["volume"], rel=1e-13)
    np.testing.assert_allclose(eg["dc_dx"], eo["dc_dx"], rtol=0, atol=1e-9 * np.abs(eo["dc_dx"]).max())
    np.testing.assert_allclose(eg["dv_dx"], eo["dv_dx"], rtol=1e-14)
    np.testing.assert_allclose(eg["energy"], eo["energy"], rtol=0, atol=1e-9 * np.abs(eo["energy"]).max())
    xg, sg = g.oc(x, eo["dc_dx"], eo["dv_dx"])
    xo, so = o.oc(x, eo["dc_dx"], eo["dv_dx"])
    assert sg["converged"] and so["converged"]
    np.testing.assert_allclose(xg, xo, rtol=0, atol=1e-9)
    assert sg["lambda_"] == pytest.approx(so["lambda_"], rel=1e-8)


def test_short_oc_run_against_oracle(ctx):
    print("test_short_oc_run_against_oracle", flush=True)
    m = M.mbb_beam(30, 15)
    g, o = _pair(m, ctx, tol=1e-10)
    x0 = m.initial_design()
    xg, hg, ng, fg = g.run_oc(x0, 15, change_tol=0.0, tol=1e-10)
    xo, ho, no, fo = o.run_oc(x0, 15, change_tol=0.0)
    assert ng == no == 15
    np.testing.assert_allclose(hg[:, 2], ho[:, 2], rtol=1e-8)
    np.testing.assert_allclose(hg[:, 4], ho[:, 4], rtol=1e-8)
    assert fg == pytest.approx(fo, rel=1e-7)


def test_projected_evaluate_and_oc_run_against_oracle(ctx):
    """Heaviside projection: the projected objective and its gradients, then a
    short run through three beta stages, against the oracle."""
    print("test_projected_evaluate_and_oc_run_against_oracle", flush=True)
    m = M.mbb_beam(30, 15)
    g, o = _pair(m, ctx, tol=1e-10)
    x = np.random.default_rng(3).uniform(0.2, 0.8, m.nelem)
    o.set_projection(4.0, 0.5)
    eg = g.evaluate(x, tol=1e-10, beta=4.0, eta=0.5)
    eo = o.evaluate(x)
    o.set_projection(0.0)
    assert eg["compliance"] == pytest.approx(eo["compliance"], rel=1e-9)
    assert eg["volume"] == pytest.approx(eo["volume"], rel=1e-13)
    np.testing.assert_allclose(eg["dc_dx"], eo["dc_dx"], rtol=1e-7, atol=1e-9 * np.abs(eo["dc_dx"]).max())
    np.testing.assert_allclose(eg["dv_dx"], eo["dv_dx"], rtol=1e-12)
    x0 = m.initial_design()
    pr = dict(beta_start=1.0, beta_max=8.0, beta_interval=5)
    xg, hg, ng, fg = g.run_oc(x0, 20, change_tol=0.0, tol=1e-10, projection=pr)
    xo, ho, no, fo = o.run_oc(x0, 20, change_tol=0.0, projection=pr)
    assert ng == no == 20
    np.testing.assert_array_equal(hg[:, 12], ho[:, 12])
    assert hg[-1, 12] == 8.0 and g.last_run["final_beta"] == 8.0
    np.testing.assert_allclose(hg[:, 2], ho[:, 2], rtol=1e-7)
    np.testing.assert_allclose(hg[:, 4], ho[:, 4], rtol=1e-8)
    assert fg == pytest.approx(fo, rel=1e-7)


def test_robust_projected_run_against_oracle(ctx):
    """The robust formulation: objective on the eroded design, volume on the dilated
    design against its rescaled target, blueprint reported; GPU loop vs oracle loop."""
    print("test_robust_projected_run_against_oracle", flush=True)
    m = M.mbb_beam(30, 15)
    g, o = _pair(m, ctx, tol=1e-10)
    x0 = m.initial_design()
    pr = dict(beta_start=1.0, beta_max=8.0, beta_interval=6, robust=True, robust_delta=0.1, robust_volume_interval=4)
    xg, hg, ng, fg = g.run_oc(x0, 20, change_tol=0.0, tol=1e-10, projection=pr, move=0.1)
    xo, ho, no, fo = o.run_oc(x0, 20, change_tol=0.0, projection=pr, move=0.1)
    assert ng == no == 20
    np.testing.assert_array_equal(hg[:, 12], ho[:, 12])
    for col in (2, 3, 4, 13, 14, 15):
        np.testing.assert_allclose(hg[:, col], ho[:, col], rtol=1e-7, err_msg=f"history column {col}")
    assert fg == pytest.approx(fo, rel=1e-7)
    lr, orr = g.last_run, o.last_robust
    for k in ("compliance_eroded", "compliance_dilated", "volume_fraction_eroded", "volume_fraction_dilated",
              "dilated_target_fraction"):
        assert lr[k] == pytest.approx(orr[k], rel=1e-7), k
    assert lr["compliance_eroded"] > fg > lr["compliance_dilated"]


def test_adjoint_solves_with_warm_slots(ctx):
    """ComplianceObjective::solve_adjoint: on the compliance operator the adjoint of the
    load is the displacement; a warm slot starts from its last solution."""
    print("test_adjoint_solves_with_warm_slots", flush=True)
    m = M.mbb_beam(60, 20)
    g, o = _pair(m, ctx, tol=1e-11)
    with pytest.raises(N.ModelError):
        g.solve_adjoint(m.loads[0])
    x = np.random.default_rng(11).uniform(0.2, 0.9, m.nelem)
    g.evaluate(x, tol=1e-11)
    u = o.evaluate(x, want_u=True)["u"][0]
    f = m.loads[0]
    lam, it0 = g.solve_adjoint(f, slot=0, tol=1e-10)
    np.testing.assert_allclose(lam, u, rtol=1e-7, atol=1e-9 * np.abs(u).max())
    assert np.all(lam[m.fixed.astype(bool)] == 0.0)
    lam2, it1 = g.solve_adjoint(1.001 * f, slot=0, tol=1e-10)
    np.testing.assert_allclose(lam2, 1.001 * lam, rtol=1e-7, atol=1e-9 * np.abs(u).max())
    lam3, it3 = g.solve_adjoint(1.001 * f, slot=-1, tol=1e-10)
    np.testing.assert_allclose(lam3, lam2, rtol=1e-7, atol=1e-9 * np.abs(u).max())
    print(f"adjoint iterations: cold {it0}, warm slot {it1}, no slot {it3}", flush=True)
    assert it1 < it3
    # a modal analysis or a plain assembly replaces K_ff: the adjoint of the last
    # evaluation is no longer available until the design is evaluated again
    g.modal(density=1.0, num_modes=2)
    with pytest.raises(N.ModelError):
        g.solve_adjoint(f, slot=0)
    g.assemble()
    with pytest.raises(N.ModelError):
        g.solve_adjoint(f, slot=0)
    g.evaluate(x, tol=1e-11)
    lam4, _ = g.solve_adjoint(f, slot=0, tol=1e-10)
    np.testing.assert_allclose(lam4, u, rtol=1e-7, atol=1e-9 * np.abs(u).max())


def test_fp64_runs_are_bitwise_reproducible():
    """Two FP64 runs in separate contexts give the same history bit for bit (the
    CSR kernels' lane counts follow a fixed rule on the FP64 path)."""
    print("test_fp64_runs_are_bitwise_reproducible", flush=True)
    m = M.mbb_beam(60, 20)
    runs = []
    for _ in range(2):
        c = N.Context(0)
        g = N.HipProblem(m, ctx=c)
        x, h, n, f = g.run_oc(m.initial_design(), 25, change_tol=0.0, tol=1e-8)
        runs.append((x, h[:, [2, 4, 5, 6, 7, 11]], f))
        c.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


def test_c1_mbb_120x40_hundred_oc_iterations(ctx):
    """The MBB half-beam at 120 x 40 (r 1.5, vf 0.5, p 3, rtol 1e-8), exactly 100
    OC iterations with no stopping rule: GPU loop against the oracle's loop."""
    print("test_c1_mbb_120x40_hundred_oc_iterations", flush=True)
    m = M.mbb_beam(120, 40)
    g, o = _pair(m, ctx, tol=1e-8)
    x0 = m.initial_design()
    xg, hg, ng, fg = g.run_oc(x0, 100, change_tol=0.0, tol=1e-8)
    xo, ho, no, fo = o.run_oc(x0, 100, change_tol=0.0)
    assert ng == no == 100
    np.testing.assert_allclose(hg[:, 2], ho[:, 2], rtol=1e-6)
    np.testing.assert_allclose(hg[:, 4], ho[:, 4], rtol=1e-8)
    assert fg == pytest.approx(fo, rel=1e-6)
    assert np.abs(xg - xo).max() < 1e-4


def test_c2_hex8_64x32x32_seeded_fp64_against_oracle(ctx):
    """Hex8 64 x 32 x 32 cantilever, densities seeded in [0.001, 1], SIMP p 3:
    FP64 assembly + AMG-PCG at rtol 1e-12 against the oracle's solve."""
    print("test_c2_hex8_64x32x32_seeded_fp64_against_oracle", flush=True)
    m = M.cantilever_3d(64, 32, 32, volume_fraction=0.3, radius_cells=2.0)
    g, o = _pair(m, ctx, tol=1e-12)
    x = M.seeded_density(m.nelem)
    eg = g.evaluate(x, tol=1e-12)
    eo = o.evaluate(x, want_u=True)
    assert eg["compliance"] == pytest.approx(eo["compliance"], rel=1e-9)
    ug = g.solution(0)
    uo = eo["u"][0][g.free()]
    assert np.linalg.norm(ug - uo) <= 1e-9 * np.linalg.norm(uo)
    np.testing.assert_allclose(eg["dc_dx"], eo["dc_dx"], rtol=0, atol=1e-8 * np.abs(eo["dc_dx"]).max())


def test_mixed_precision_solve_meets_tolerance(ctx):
    print("test_mixed_precision_solve_meets_tolerance", flush=True)
    m = M.cantilever_3d(24, 12, 12)
    g = N.HipProblem(m, ctx=ctx)
    x = M.seeded_density(m.nelem)
    scale = 1e-9 + (1 - 1e-9) * x ** 3
    g.assemble(scale)
    b = m.free_loads()[0]
    g.amg_setup()
    x64, it64, rr64, cv64 = g.solve(b, tol=1e-10)
    g.amg_setup(precision=1)
    x32, it32, rr32, cv32 = g.solve(b, tol=1e-10)
    assert cv64 and cv32 and rr32 <= 1e-10
    assert b @ x32 == pytest.approx(b @ x64, rel=1e-9)
    assert it32 <= it64 + 3
    # coarse operators in bfloat16: the same tolerance, the same answer
    g.amg_setup(precision=2)
    x16, it16, rr16, cv16 = g.solve(b, tol=1e-10)
    print(f"iterations fp64 {it64}, fp32 cycle {it32}, bf16 coarse {it16}", flush=True)
    assert cv16 and rr16 <= 1e-10
    assert b @ x16 == pytest.approx(b @ x64, rel=1e-9)
    assert it16 <= it64 + 8


def test_matrix_free_operator_matches_assembled_kff(ctx):
    """The matrix-free Hex8 product equals the assembled K_ff's (partial tiles:
    node counts that are not multiples of the 8 x 8 x 4 tile)."""
    print("test_matrix_free_operator_matches_assembled_kff", flush=True)
    m = M.cantilever_3d(20, 9, 6)
    g = N.HipProblem(m, ctx=ctx)
    g.assemble(1e-9 + (1 - 1e-9) * M.seeded_density(m.nelem, seed=21) ** 3)
    K = _csr(g.nfree, g.nfree, *g.kff())
    rng = np.random.default_rng(8)
    for _ in range(2):
        x = rng.standard_normal(g.nfree)
        y, mf = g.kff_apply(x)
        assert mf
        ref = K @ x
        np.testing.assert_allclose(y, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_matrix_free_and_assembled_solves_agree(ctx, monkeypatch):
    print("test_matrix_free_and_assembled_solves_agree", flush=True)
    m = M.cantilever_3d(24, 12, 12)
    scale = 1e-9 + (1 - 1e-9) * M.seeded_density(m.nelem, seed=22) ** 3
    b = m.free_loads()[0]
    r = np.random.default_rng(23).standard_normal(len(b))
    out, cyc = {}, {}
    for mode in ("mf", "csr"):
        if mode == "csr":
            monkeypatch.setenv("SPARLAB_NO_MATRIX_FREE", "1")
        g = N.HipProblem(m, ctx=ctx)
        g.assemble(scale)
        assert g.kff_apply(np.ones(g.nfree))[1] == (mode == "mf")
        g.amg_setup()
        # one V-cycle: level-0 smoothing and the transfers (matrix-free: P = (I - w D^-1 K) P-hat
        # and R = P^T applied through the operator; assembled: the stored P and R)
        cyc[mode] = g.amg_apply(r)
        out[mode] = g.solve(b, tol=1e-10)
    np.testing.assert_allclose(cyc["mf"], cyc["csr"], rtol=0, atol=1e-11 * np.abs(cyc["csr"]).max())
    (xm, im, rm, cm), (xc, ic, rc, cc) = out["mf"], out["csr"]
    assert cm and cc and abs(im - ic) <= 1
    assert np.linalg.norm(xm - xc) <= 1e-9 * np.linalg.norm(xc)


@pytest.mark.parametrize("precision", [0, 1])
def test_captured_cg_graphs_match_eager_launches(ctx, precision, monkeypatch):
    """The CG iteration replayed from HIP graphs (SPARLAB_GRAPHS=1) gives the
    eager launches' bits, on every replay, with eager work (a V-cycle, a fresh
    setup) in between."""
    print(f"test_captured_cg_graphs_match_eager_launches[{precision}]", flush=True)
    m = M.cantilever_3d(16, 8, 8)
    g = N.HipProblem(m, ctx=ctx)
    g.assemble(1e-9 + (1 - 1e-9) * M.seeded_density(m.nelem, seed=11) ** 3)
    b = m.free_loads()[0]
    r = np.random.default_rng(6).standard_normal(g.nfree)
    monkeypatch.delenv("SPARLAB_GRAPHS", raising=False)
    g.amg_setup(coarse_size=300, precision=precision)
    x_eager, it_eager, rr_eager, _ = g.solve(b, tol=1e-10)
    monkeypatch.setenv("SPARLAB_GRAPHS", "1")
    for replay in range(3):
        if replay == 2:
            g.amg_setup(coarse_size=300, precision=precision)
        x, it, rr, cv = g.solve(b, tol=1e-10)
        assert cv and it == it_eager and rr == rr_eager
        assert np.array_equal(x, x_eager), f"replay {replay}"
        g.amg_apply(r)


def test_error_paths(ctx):
    print("test_error_paths", flush=True)
    # under-constrained: no supports -> the coarse operator is singular
    m = M.cantilever_3d(6, 3, 3)
    m.fixed[:] = 0
    g = N.HipProblem(m, ctx=ctx)
    g.assemble()
    with pytest.raises(N.SolverError):
        g.amg_setup()
    # iteration cap -> not converged
    m2 = M.cantilever_3d(12, 6, 6)
    g2 = N.HipProblem(m2, ctx=ctx)
    g2.assemble()
    g2.amg_setup()
    x, it, rr, cv = g2.solve(m2.free_loads()[0], tol=1e-14, maxit=2)
    assert not cv and it == 2
    with pytest.raises(N.ConvergenceError):
        g2.evaluate(np.full(m2.nelem, 0.5), tol=1e-14, max_iterations=2)
    with pytest.raises(N.ModelError):
        g2.assemble(-np.ones(m2.nelem))


def test_smoother_timing_leaves_the_solver_usable(ctx):
    """The smoother timing hook overwrites a level's work vectors (level 0: the
    matrix-free step); the next V-cycle and solve must be unaffected."""
    print("test_smoother_timing_leaves_the_solver_usable", flush=True)
    m = M.cantilever_3d(24, 12, 12)
    g = N.HipProblem(m, ctx=ctx)
    g.assemble(1e-9 + (1 - 1e-9) * M.seeded_density(m.nelem, seed=31) ** 3)
    g.amg_setup(precision=1, coarse_size=300)
    assert g.amg_stats().levels >= 3
    b = m.free_loads()[0]
    r = np.random.default_rng(32).standard_normal(g.nfree)
    z0 = g.amg_apply(r)
    x0, it0, _, cv0 = g.solve(b, tol=1e-10)
    ms = g.time_smoother(1, 5)
    ms0 = g.time_smoother(0, 5)
    assert ms > 0.0 and ms0 > 0.0
    z1 = g.amg_apply(r)
    x1, it1, _, cv1 = g.solve(b, tol=1e-10)
    np.testing.assert_array_equal(z0, z1)
    assert cv0 and cv1 and it0 == it1
    np.testing.assert_array_equal(x0, x1)
    with pytest.raises(N.ConfigError):
        g.time_smoother(g.amg_stats().levels - 1, 5)  # the coarsest level has no smoother


def test_coarse_reuse_keeps_the_answers_and_refreshes_level0():
    """coarse_reuse (mixed path): evaluations that keep A_1 and below refresh level 0
    only (no Galerkin time), an OC run follows the fresh-hierarchy run to the solve
    tolerance, and K_ff read back after a lagged evaluation is the current operator's."""
    print("test_coarse_reuse_keeps_the_answers_and_refreshes_level0", flush=True)
    m = M.cantilever_3d(24, 12, 12)
    x0 = m.initial_design()
    runs = {}
    for reuse in (0, 3):
        c = N.Context(0)
        g = N.HipProblem(m, ctx=c)
        g.assemble(1e-9 + (1 - 1e-9) * np.clip(g.filter(x0), 0, 1) ** 3)
        g.amg_setup(precision=1, coarse_reuse=reuse)
        x, h, n, f = g.run_oc(x0, 12, change_tol=0.0, tol=1e-10)
        runs[reuse] = (x, h, f)
        if reuse:
            # a full setup restarts the reuse: one full evaluation, then a lagged one (level 0 only)
            g.amg_setup(precision=1, coarse_reuse=reuse)
            x1 = np.clip(x * 1.01, 0.0, 1.0)
            e1 = g.evaluate(x1, tol=1e-10)
            st = g.amg_stats()
            x2 = np.clip(x1 * 0.99, 0.0, 1.0)
            e2 = g.evaluate(x2, tol=1e-10)
            st2 = g.amg_stats()
            print(f"coarse reuse: setups {st.setup_seconds * 1e3:.2f} / {st2.setup_seconds * 1e3:.2f} ms, "
                  f"galerkin {st.galerkin_seconds * 1e3:.2f} / {st2.galerkin_seconds * 1e3:.2f} ms, "
                  f"its {e1['iterations']} / {e2['iterations']}", flush=True)
            assert st.galerkin_seconds > 0.0
            assert st2.galerkin_seconds == 0.0 and st2.reused_aggregates
            # the lagged solution solves the current operator: K u = f to the tolerance
            u = g.solution(0)
            y, _ = g.kff_apply(u)
            b = m.free_loads()[0]
            assert np.linalg.norm(y - b) <= 1e-9 * np.linalg.norm(b)
            # K_ff read back after the lagged evaluation (assembled on demand) is the one a
            # fresh context assembles for x2
            pg, cg, vg = g.kff()
            cf = N.Context(0)
            gf = N.HipProblem(m, ctx=cf)
            gf.evaluate(x2, tol=1e-10)
            assert np.array_equal(vg, gf.kff()[2])
            cf.close()
        c.close()
    (xa, ha, fa), (xb, hb, fb) = runs[0], runs[3]
    np.testing.assert_allclose(hb[:, 2], ha[:, 2], rtol=1e-7)
    assert fb == pytest.approx(fa, rel=1e-7)
    print(f"coarse reuse: CG its fresh {ha[:, 11].astype(int).tolist()} reused {hb[:, 11].astype(int).tolist()}",
          flush=True)


def test_matrix_core_level0_products_match_the_vector_path(ctx, monkeypatch):
    """The FP32 V-cycle's level-0 products on the matrix cores (split-bfloat16
    MFMA, FP32 accumulation) against the FP32 vector-FMA path: one V-cycle to
    1e-4, a solve to rtol 1e-10 in the same iterations (+-1) and to the same
    compliance; the Chebyshev step's time for both."""
    print("test_matrix_core_level0_products_match_the_vector_path", flush=True)
    m = M.cantilever_3d(24, 12, 12)
    g = N.HipProblem(m, ctx=ctx)
    g.assemble(1e-9 + (1 - 1e-9) * M.seeded_density(m.nelem, seed=91) ** 3)
    b = m.free_loads()[0]
    r = np.random.default_rng(92).standard_normal(g.nfree)
    out = {}
    for mode in ("vector", "matrix"):
        monkeypatch.setenv("SPARLAB_MF_MFMA", "0" if mode == "vector" else "1")
        g.amg_setup(precision=1, coarse_size=300)
        z = g.amg_apply(r)
        x, it, rr, cv = g.solve(b, tol=1e-10)
        ms = g.time_smoother(0, 20)
        out[mode] = (z, x, it, cv, ms)
    (zv, xv, itv, cvv, msv), (zm, xm, itm, cvm, msm) = out["vector"], out["matrix"]
    dz = np.linalg.norm(zm - zv) / np.linalg.norm(zv)
    print(f"matrix-core level 0: V-cycle rel diff {dz:.2e}, iterations {itv} / {itm}, "
          f"Chebyshev step {1e3 * msv:.1f} / {1e3 * msm:.1f} us", flush=True)
    assert cvv and cvm and dz <= 1e-4 and abs(itm - itv) <= 1
    assert b @ xm == pytest.approx(b @ xv, rel=1e-10)


@pytest.mark.parametrize("precision", [0, 1])
def test_matrix_free_tile_shapes_give_the_same_bits(ctx, precision, monkeypatch):
    """Every tile shape of the matrix-free kernels (fixed with SPARLAB_MF_TILE)
    gives the V-cycle and the solve the same bits, on a grid that none of the
    shapes divides (21 x 10 x 8 nodes): a node's sums run in one order whatever
    the tile. Without the variable the setup times the shapes and keeps one."""
    print(f"test_matrix_free_tile_shapes_give_the_same_bits[{precision}]", flush=True)
    m = M.cantilever_3d(20, 9, 7)
    scale = 1e-9 + (1 - 1e-9) * M.seeded_density(m.nelem, seed=31) ** 3
    b = m.free_loads()[0]
    r = np.random.default_rng(32).standard_normal(len(b))
    shapes = {0: (8, 8, 4), 1: (16, 4, 4), 2: (32, 4, 2), 3: (16, 8, 2)}

    def run():
        # a new problem per shape: a setup that reuses the structure keeps its tile
        g = N.HipProblem(m, ctx=ctx)
        g.assemble(scale)
        g.amg_setup(precision=precision, coarse_size=300)
        assert g.amg_stats().levels >= 2
        z = g.amg_apply(r)
        x, it, rr, cv = g.solve(b, tol=1e-10)
        assert cv
        return g.mf_tile(), (z, x, it, rr)

    ref = None
    for t, dims in shapes.items():
        monkeypatch.setenv("SPARLAB_MF_TILE", str(t))
        tile, out = run()
        assert tile == dims
        if ref is None:
            ref = out
        else:
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), f"tile {dims}"
            assert out[2:] == ref[2:]
    monkeypatch.delenv("SPARLAB_MF_TILE")
    tile, out = run()
    print(f"precision {precision}: the setup picked tile {tile}", flush=True)
    assert tile in shapes.values()
    assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])


def _three_case_model(kind):
    """Three load cases: the model's own, a second load elsewhere (another
    iteration count) and an unloaded case (zero right-hand side)."""
    if kind == "q4":
        m = M.mbb_beam(120, 40)
        c = m.add_load_case(0.5)
        m.add_point_load(c, m.select_nearest([60.0, 40.0]), [0.3, -1.0])
    else:
        m = M.cantilever_3d(24, 12, 12)
        c = m.add_load_case(0.5)
        m.add_point_load(c, m.select_box(xmin=24.0, ymin=12.0), [0.0, 0.0, 1.0])
    m.add_load_case(0.25)
    return m


@pytest.mark.parametrize("kind", ["hex8_matrix_free", "hex8_assembled", "q4", "hex8_matrix_free_fp32_cycle",
                                  "q4_bf16_coarse", "hex8_matrix_free_gauss_seidel", "q4_gauss_seidel_fp32_cycle"])
def test_block_load_cases_equal_case_by_case_solves(ctx, kind, monkeypatch):
    """The load cases of an evaluation solved as one block PCG (one operator
    product and one V-cycle pass per iteration for every case still iterating)
    give each case's case-by-case result bit for bit: displacements, iteration
    counts, compliance and sensitivities, over three evaluations that warm-start
    from the previous solutions, with one case unloaded. FP64 V-cycle, the
    mixed hierarchies (FP32 V-cycle under flexible CG; bfloat16 coarse operators)
    and the Gauss-Seidel smoother."""
    print(f"test_block_load_cases_equal_case_by_case_solves[{kind}]", flush=True)
    precision = 1 if "fp32" in kind else 2 if "bf16" in kind else 0
    smoother = 1 if "gauss_seidel" in kind else 0
    if precision:
        # the mixed path times the CSR kernels' lane counts per context: fix them
        monkeypatch.setenv("SPARLAB_NO_AUTOTUNE", "1")
    if kind == "hex8_assembled":
        monkeypatch.setenv("SPARLAB_NO_MATRIX_FREE", "1")
    m = _three_case_model("q4" if kind.startswith("q4") else "hex8")
    designs = [M.seeded_density(m.nelem, seed=31 + k) for k in range(3)]
    out = {}
    for mode in ("block", "case_by_case"):
        if mode == "case_by_case":
            monkeypatch.setenv("SPARLAB_NO_BLOCK_CASES", "1")
        c = N.Context(0)
        c.timing(True)
        g = N.HipProblem(m, ctx=c)
        if precision or smoother:
            g.assemble()
            g.amg_setup(precision=precision, smoother=smoother)
        runs = []
        for x in designs:
            ev = g.evaluate(x, tol=1e-10)
            runs.append((ev, [g.solution(l) for l in range(3)]))
        out[mode] = runs
        launches = c.timer("block_spmv")[1]
        print(f"{kind} {mode}: iterations {[r[0]['iterations'] for r in runs]}, block products {launches}",
              flush=True)
        assert (launches > 0) == (mode == "block")
        c.close()
        monkeypatch.delenv("SPARLAB_NO_BLOCK_CASES", raising=False)
    for (eb, ub), (es, us) in zip(out["block"], out["case_by_case"]):
        assert eb["iterations"] == es["iterations"]
        assert eb["compliance"] == es["compliance"]
        assert np.array_equal(eb["dc_dx"], es["dc_dx"])
        for a, b in zip(ub, us):
            assert np.array_equal(a, b)
        assert not ub[2].any()  # the unloaded case


@pytest.mark.parametrize("precision", [0, 1])
def test_block_cases_replayed_from_graphs_match_eager_launches(precision, monkeypatch):
    """With SPARLAB_GRAPHS=1 the block PCG replays its captured iteration; the
    evaluations give the eager launches' bits (iterations, compliance,
    sensitivities, every case's displacements), across a new setup per design."""
    print(f"test_block_cases_replayed_from_graphs_match_eager_launches[{precision}]", flush=True)
    if precision:
        monkeypatch.setenv("SPARLAB_NO_AUTOTUNE", "1")
    m = _three_case_model("hex8")
    designs = [M.seeded_density(m.nelem, seed=51 + k) for k in range(2)]
    out = {}
    for mode in ("eager", "graphs"):
        if mode == "graphs":
            monkeypatch.setenv("SPARLAB_GRAPHS", "1")
        else:
            monkeypatch.delenv("SPARLAB_GRAPHS", raising=False)
        c = N.Context(0)
        g = N.HipProblem(m, ctx=c)
        if precision:
            g.assemble()
            g.amg_setup(precision=precision)
        out[mode] = [(g.evaluate(x, tol=1e-10), [g.solution(l) for l in range(3)]) for x in designs]
        c.close()
    for (ee, ue), (eg, ug) in zip(out["eager"], out["graphs"]):
        assert ee["iterations"] == eg["iterations"] and ee["compliance"] == eg["compliance"]
        assert np.array_equal(ee["dc_dx"], eg["dc_dx"])
        for a, b in zip(ue, ug):
            assert np.array_equal(a, b)


def test_block_load_cases_oc_run_equals_case_by_case(monkeypatch):
    """A short multi-case OC run (the block PCG inside every evaluation) gives
    the case-by-case run's history and design bit for bit."""
    print("test_block_load_cases_oc_run_equals_case_by_case", flush=True)
    m = _three_case_model("q4")
    res = {}
    for mode in ("block", "case_by_case"):
        if mode == "case_by_case":
            monkeypatch.setenv("SPARLAB_NO_BLOCK_CASES", "1")
        c = N.Context(0)
        g = N.HipProblem(m, ctx=c)
        res[mode] = g.run_oc(m.initial_design(), 12, tol=1e-10)
        c.close()
        monkeypatch.delenv("SPARLAB_NO_BLOCK_CASES", raising=False)
    (xb, hb, ib, cb), (xs, hs, is_, cs) = res["block"], res["case_by_case"]
    assert ib == is_ and cb == cs
    keep = [k for k in range(hb.shape[1]) if k != N.H_SECONDS]  # wall-clock seconds differ
    assert np.array_equal(xb, xs) and np.array_equal(hb[:, keep], hs[:, keep])


def test_caller_csr_is_checked_on_the_host_and_a_new_pattern_gets_a_new_setup():
    """slx_factorize_csr checks every index the device and the aggregation will
    follow (row pointers, column range, layout unknowns) and refuses bad input
    with a ConfigError; a second matrix of the same size, non-zero count and
    layout but another pattern gets a setup of its own (its solution equals a
    fresh context's bit for bit)."""
    print("test_caller_csr_is_checked_on_the_host_and_a_new_pattern_gets_a_new_setup", flush=True)
    m = M.bench_box(8)
    ptr, col, val = Oracle(m, tol=1e-10).assemble_kff()
    n = len(ptr) - 1
    coords, free = m.node_coords(), m.free_dofs()
    s = N.HipAmgCgSolver(tolerance=1e-10, ctx=N.Context(0))
    s.set_layout(3, coords, free, dofs_per_node=3)
    bad = col.copy()
    bad[-1] = n + 5
    with pytest.raises(N.ConfigError):
        s.factorize(ptr, bad, val)
    bad = ptr.copy()
    bad[5] = bad[6] + 1
    with pytest.raises(N.ConfigError):
        s.factorize(bad, col, val)
    s.set_layout(3, coords, np.where(np.arange(n) == 3, 3 * m.nnodes, free), dofs_per_node=3)
    with pytest.raises(N.ConfigError):
        s.factorize(ptr, col, val)
    s.set_layout(3, coords, free, dofs_per_node=3)
    s.factorize(ptr, col, val)
    b = m.free_loads()[0]
    x0 = s.solve(b)
    # a permuted ordering: the same n, nnz and layout, another pattern
    K = sp.csr_matrix((val, col, ptr), shape=(n, n))
    rev = np.random.default_rng(5).permutation(n)
    Kr = K[rev][:, rev].tocsr()
    Kr.sort_indices()
    assert Kr.nnz == K.nnz and not np.array_equal(Kr.indices, col)
    s.factorize(Kr.indptr, Kr.indices, Kr.data)
    x1 = s.solve(b[rev])
    fresh = N.HipAmgCgSolver(tolerance=1e-10, ctx=N.Context(0))
    fresh.set_layout(3, coords, free, dofs_per_node=3)
    fresh.factorize(Kr.indptr, Kr.indices, Kr.data)
    x2 = fresh.solve(b[rev])
    assert np.array_equal(x1, x2) and s.last_iterations() == fresh.last_iterations()
    assert np.linalg.norm(x1 - x0[rev]) <= 1e-8 * np.linalg.norm(x0)  # (K[rev][:, rev]) x[rev] = b[rev]
    s.ctx.close()
    fresh.ctx.close()


def test_coarsest_level_above_eight_thousand_unknowns(ctx):
    """A one-level hierarchy (max_levels 1) makes the whole system the coarsest
    level: 8 712 unknowns, above the reference's dense limit (Multigrid.cpp:749-764
    takes its sparse LDL^T there) and the former 8 000 of this path. The dense
    inverse then solves it directly: one or two CG iterations, the solution of a sparse
    direct solve. Above 16 384 unknowns the setup refuses with the options to change."""
    import scipy.sparse.linalg as sla
    print("test_coarsest_level_above_eight_thousand_unknowns", flush=True)
    m = M.cantilever_3d(24, 10, 10)
    g = N.HipProblem(m, ctx=ctx)
    scale = 0.2 + 0.8 * M.seeded_density(m.nelem, seed=13)
    g.assemble(scale)
    g.amg_setup(max_levels=1)
    st = g.amg_stats()
    assert st.levels == 1 and st.unknowns[0] == g.nfree == 8712
    assert 0.0 < st.coarse_pivot_ratio <= 1.0
    b = m.free_loads()[0]
    x, it, rr, cv = g.solve(b, tol=1e-10)
    K = _csr(g.nfree, g.nfree, *g.kff())
    xd = sla.spsolve(K.tocsc(), b)
    assert cv and it <= 2
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    big = M.cantilever_3d(32, 16, 16)
    gb = N.HipProblem(big, ctx=ctx)
    gb.assemble()
    with pytest.raises(N.SolverError, match="lower amg.coarse_size or raise amg.max_levels"):
        gb.amg_setup(max_levels=1)
