"""Pins the refcpu oracle before anything is compared against it.

Checks, per oracle function, against: the reference's published results
(tests/golden/published.json, copied from docs/results/*.csv), the invariants
the reference's own tests assert (tests/test_multigrid.cpp, test_topopt.cpp,
test_solid.cpp, test_assembly_solver.cpp), closed forms, independent numpy /
scipy restatements and brute force on tiny inputs.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import refcpu
from oracle.refcpu import Oracle
from sparlab_amd import model as M

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))


def csr(rows, cols, ptr, col, val):
    return sp.csr_matrix((val, col, ptr), shape=(rows, cols))


# ---------------------------------------------------------------------------
# Published results (docs/results/solver_scaling.csv:24-33)
# ---------------------------------------------------------------------------
# Rows whose level-1 aggregation is not reproduced to the last entry: at these
# sizes a near-tie of the pass-2 strength comparison at level 1 resolves
# differently (the coarse operators' last bits), so the operator complexity (96,
# 128) or the stored non-zeros (48, 96, 128) differ slightly while compliance,
# iterations and levels match. Every other row matches exactly.
_LEVEL1_FLIPS = {48: (0.0, 3e-4), 96: (1e-5, 3e-5), 128: (1e-5, 3e-5)}


def _bench_row_check(row, dim):
    m = M.bench_box(row["nx"], dim=dim)
    assert (m.nx, m.ny) == (row["nx"], row["ny"])
    if dim == 3:
        assert m.nz == row["nz"]
    assert m.ndofs == row["num_dofs"]
    o = Oracle(m, tol=1e-10)
    ptr, col, val = o.assemble_kff()
    assert len(val) == row["kff_nnz"]
    o.amg_setup()
    b = m.free_loads()[0]
    x, it, rr, cv = o.pcg(b, tol=1e-10)
    assert cv
    c = float(b @ x)
    assert abs(c - row["compliance"]) <= 1e-11 * row["compliance"]
    assert it == row["iterations"]
    assert o.amg_levels() == row["levels"]
    # The operator complexity counts the non-zeros of every coarse operator and
    # the stored non-zeros add those of P-hat, P, R, A P and R A P of every
    # level (Multigrid.cpp:1089-1100): both are fixed by the aggregates and the
    # prolongator ranks, so equality pins the strength test, the greedy passes
    # (whose pass-2 ties depend on the last bits of the coarse operators), the
    # column-pivoted QR and the Lanczos estimate. The published complexities
    # carry one or two units of the 17th digit from the results pipeline
    # (rel 1e-15 covers them).
    oc_tol, snz_tol = _LEVEL1_FLIPS.get(row["nx"], (0.0, 0.0)) if dim == 3 else (0.0, 0.0)
    assert o.amg_operator_complexity() == pytest.approx(row["operator_complexity"], rel=max(oc_tol, 1e-15))
    snz = o.amg_storage_nonzeros()
    assert abs(snz - row["solver_nonzeros"]) <= snz_tol * row["solver_nonzeros"], (snz, row["solver_nonzeros"])


@pytest.mark.parametrize("row", GOLDEN["hex8_bench_amg"], ids=lambda r: f"nx{r['nx']}")
def test_hex8_bench_matches_published(row):
    _bench_row_check(row, 3)


@pytest.mark.parametrize("row", GOLDEN["q4_bench_amg"], ids=lambda r: f"nx{r['nx']}")
def test_q4_bench_matches_published(row):
    _bench_row_check(row, 2)


def test_kff_nnz_closed_form():
    # Appendix of the survey: nnz(K_ff) = 9 (3nx-2)(3ny+1)(3nz+1) for the clamped box.
    for row in GOLDEN["hex8_bench_amg"]:
        nx, ny, nz = row["nx"], row["ny"], row["nz"]
        assert 9 * (3 * nx - 2) * (3 * ny + 1) * (3 * nz + 1) == row["kff_nnz"]
    m = M.cantilever_3d(6, 3, 2)
    o = Oracle(m)
    _, _, val = o.assemble_kff()
    assert len(val) == 9 * (3 * 6 - 2) * (3 * 3 + 1) * (3 * 2 + 1)
    full = o.pattern()[1]
    assert len(full) == 9 * (3 * 6 + 1) * (3 * 3 + 1) * (3 * 2 + 1)


def test_mbb_published_compliances():
    g = GOLDEN["mbb_beam_180x60"]
    m = M.mbb_beam(180, 60, radius_cells=g["radius_elements"])
    o = Oracle(m, tol=1e-12)
    solid = o.evaluate(np.ones(m.nelem), need_gradients=False)["compliance"]
    # the published value is an LDL^T solve; its own rounding is ~1e-10 here
    assert abs(solid - g["solid_compliance"]) <= 3e-10 * g["solid_compliance"]
    o2 = Oracle(m, tol=1e-12)
    init = o2.evaluate(np.full(m.nelem, 0.5), need_gradients=False)["compliance"]
    assert abs(init - g["initial_compliance"]) <= 1e-10 * g["initial_compliance"]
    # closed form: a uniform design scales the solid compliance by 1/E(rho)
    factor = 1e-9 + (1 - 1e-9) * 0.5 ** 3
    assert init == pytest.approx(solid / factor, rel=1e-9)


# ---------------------------------------------------------------------------
# Element stiffness (Hex8.cpp:172-192, Quad4.cpp:117-137)
# ---------------------------------------------------------------------------
def test_q4_unit_square_matches_textbook():
    # The classical closed-form stiffness of a unit-square bilinear element in
    # plane stress, nodes counter-clockwise from the lower left.
    nu, E = 0.3, 1.0
    m = M.StructuredModel(dim=2, nx=1, ny=1, lx=1.0, ly=1.0, E=E, nu=nu)
    k = Oracle(m).ke0()
    kk = np.array([1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12, -1 / 8 + 3 * nu / 8,
                   -1 / 4 + nu / 12, -1 / 8 - nu / 8, nu / 6, 1 / 8 - 3 * nu / 8])
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2], [2, 7, 0, 5, 6, 3, 4, 1], [3, 6, 5, 0, 7, 2, 1, 4],
           [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6], [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    ref = E / (1 - nu ** 2) * kk[np.array(idx)]
    np.testing.assert_allclose(k, ref, rtol=0, atol=1e-15)


def _numpy_hex8(dx, dy, dz, E, nu, order):
    g, w = np.polynomial.legendre.leggauss(order)
    corners = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                        [-1, 1, 1]], dtype=float)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * G
    D[3:, 3:] = np.eye(3) * G
    K = np.zeros((24, 24))
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for c, wc in zip(g, w):
                dN = np.stack([0.125 * corners[:, 0] * (1 + b * corners[:, 1]) * (1 + c * corners[:, 2]),
                               0.125 * corners[:, 1] * (1 + a * corners[:, 0]) * (1 + c * corners[:, 2]),
                               0.125 * corners[:, 2] * (1 + a * corners[:, 0]) * (1 + b * corners[:, 1])], axis=1)
                dNdx = dN / np.array([dx / 2, dy / 2, dz / 2])
                B = np.zeros((6, 24))
                for n in range(8):
                    x, y, z = dNdx[n]
                    B[0, 3 * n], B[1, 3 * n + 1], B[2, 3 * n + 2] = x, y, z
                    B[3, 3 * n], B[3, 3 * n + 1] = y, x
                    B[4, 3 * n + 1], B[4, 3 * n + 2] = z, y
                    B[5, 3 * n], B[5, 3 * n + 2] = z, x
                K += wa * wb * wc * (dx * dy * dz / 8) * B.T @ D @ B
    return K


def test_hex8_ke0_against_independent_integration():
    m = M.StructuredModel(dim=3, nx=1, ny=1, nz=1, lx=2.0, ly=1.0, lz=0.5, E=70e9, nu=0.3)
    k = Oracle(m).ke0()
    ref2 = _numpy_hex8(2.0, 1.0, 0.5, 70e9, 0.3, 2)
    ref3 = _numpy_hex8(2.0, 1.0, 0.5, 70e9, 0.3, 3)
    scale = np.abs(ref2).max()
    np.testing.assert_allclose(k, ref2, rtol=0, atol=1e-14 * scale)
    # the 2x2x2 rule is exact for a box (test_solid.cpp:163-185)
    np.testing.assert_allclose(ref3, ref2, rtol=0, atol=1e-13 * scale)
    assert np.array_equal(k, k.T)


@pytest.mark.parametrize("dim", [2, 3])
def test_ke0_rigid_body_null_space(dim):
    if dim == 3:
        m = M.StructuredModel(dim=3, nx=1, ny=1, nz=1, lx=1.3, ly=0.7, lz=0.9, E=3.0, nu=0.25)
    else:
        m = M.StructuredModel(dim=2, nx=1, ny=1, lx=1.3, ly=0.7, E=3.0, nu=0.25)
    k = Oracle(m).ke0()
    ev = np.linalg.eigvalsh(k)
    nzero = 6 if dim == 3 else 3
    assert (np.abs(ev) < 1e-12 * ev.max()).sum() == nzero
    assert ev.min() > -1e-12 * ev.max()
    # linear in E
    m2 = M.StructuredModel(**{**m.__dict__, "E": 6.0, "fixed": None, "lower": None, "upper": None, "passive": None})
    np.testing.assert_allclose(Oracle(m2).ke0(), 2 * k, rtol=1e-15, atol=0)


# ---------------------------------------------------------------------------
# Pattern, scatter and reduction (Assembler.cpp:85-225,410-438)
# ---------------------------------------------------------------------------
def _elements(m):
    nnx, nny = m.nx + 1, m.ny + 1
    out = []
    if m.dim == 2:
        for j in range(m.ny):
            for i in range(m.nx):
                n0 = j * nnx + i
                out.append([n0, n0 + 1, n0 + nnx + 1, n0 + nnx])
    else:
        lay = nnx * nny
        for k in range(m.nz):
            for j in range(m.ny):
                for i in range(m.nx):
                    n0 = k * lay + j * nnx + i
                    out.append([n0, n0 + 1, n0 + nnx + 1, n0 + nnx, n0 + lay, n0 + lay + 1, n0 + lay + nnx + 1,
                                n0 + lay + nnx])
    return np.array(out)


@pytest.mark.parametrize("dim", [2, 3])
def test_pattern_and_scatter_against_element_order_accumulation(dim):
    if dim == 3:
        m = M.cantilever_3d(3, 2, 2)
    else:
        m = M.mbb_beam(5, 3)
    o = Oracle(m)
    outer, inner = o.pattern()
    conn = _elements(m)
    d = m.dim
    # neighbourhoods: nodes sharing an element, sorted
    nbr = [set() for _ in range(m.nnodes)]
    for el in conn:
        for a in el:
            nbr[a].update(el.tolist())
    ref_outer = [0]
    ref_inner = []
    for b in range(m.nnodes):
        nb = sorted(nbr[b])
        for kb in range(d):
            ref_inner += [n * d + ka for n in nb for ka in range(d)]
            ref_outer.append(len(ref_inner))
    assert np.array_equal(outer, ref_outer)
    assert np.array_equal(inner, ref_inner)
    # values: sum in ascending element order, each element adding to distinct slots
    rng = np.random.default_rng(7)
    scale = rng.uniform(0.1, 1.0, m.nelem)
    ke = o.ke0()
    pos = {}
    for c in range(m.ndofs):
        for k in range(outer[c], outer[c + 1]):
            pos[(inner[k], c)] = k
    vals = np.zeros(len(inner))
    for e, el in enumerate(conn):
        dofs = np.array([n * d + k for n in el for k in range(d)])
        slots = np.array([[pos[(r, c)] for c in dofs] for r in dofs])
        vals[slots.ravel()] += scale[e] * ke.ravel()
    assert np.array_equal(o.assemble_full(scale), vals)
    # K_ff reduction keeps free rows/cols in order
    ptr, col, val = o.assemble_kff(scale)
    free = o.free()
    assert np.array_equal(free, m.free_dofs())
    K = sp.csc_matrix((vals, inner, outer), shape=(m.ndofs, m.ndofs)).tocsr()
    Kff = K[free][:, free].tocsr()
    Kff.sort_indices()
    assert np.array_equal(ptr, Kff.indptr) and np.array_equal(col, Kff.indices) and np.array_equal(val, Kff.data)


# ---------------------------------------------------------------------------
# Kernels of Multigrid.cpp
# ---------------------------------------------------------------------------
def test_chunked_dot():
    rng = np.random.default_rng(1)
    for n in [10, 4096, 4097, 100000]:
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        d = refcpu.chunked_dot(a, b)
        assert d == refcpu.chunked_dot(a, b)
        assert abs(d - math.fsum(a * b)) <= 1e-13 * np.abs(a * b).sum()
    # chunk structure: partial sums of 4096 in order, then summed in order
    a, b = rng.standard_normal(10000), rng.standard_normal(10000)
    parts = []
    for c in range(0, 10000, 4096):
        s = 0.0
        for i in range(c, min(c + 4096, 10000)):
            s += a[i] * b[i]
        parts.append(s)
    t = 0.0
    for p in parts:
        t += p
    assert refcpu.chunked_dot(a, b) == t


def test_cpqr_against_numpy():
    rng = np.random.default_rng(3)
    for rows, m, rank in [(81, 6, 6), (9, 6, 6), (3, 6, 3), (12, 6, 5), (6, 3, 3)]:
        B = rng.standard_normal((rows, rank)) @ rng.standard_normal((rank, m))
        q, r = refcpu.cpqr(B)
        assert q.shape == (rows, min(rank, rows))
        np.testing.assert_allclose(q.T @ q, np.eye(q.shape[1]), atol=1e-13)
        np.testing.assert_allclose(q @ r, B, atol=1e-12 * np.abs(B).max())
        # R's pivoted diagonal magnitudes match numpy's singular-value-free check
        assert np.linalg.matrix_rank(B) == q.shape[1]


def test_tridiagonal_eigenvalue_against_numpy():
    rng = np.random.default_rng(5)
    for k in [1, 2, 5, 12]:
        d = rng.uniform(0.5, 2.0, k)
        s = rng.uniform(0.01, 0.5, k - 1)
        T = np.diag(d) + np.diag(s, 1) + np.diag(s, -1)
        assert refcpu.tridiag_max(d, s) == pytest.approx(np.linalg.eigvalsh(T).max(), rel=1e-14)


# ---------------------------------------------------------------------------
# AMG hierarchy identities (tests/test_multigrid.cpp:178-255,513-542)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["hex", "quad"])
def test_amg_hierarchy_identities(case):
    if case == "hex":
        m = M.bench_box(16)
    else:
        m = M.bench_box(80, dim=2)
    o = Oracle(m, amg={"coarse_size": 200})
    o.assemble_kff()
    o.amg_setup()
    L = o.amg_levels()
    assert L >= 3
    assert 1.0 < o.amg_operator_complexity() < 3.0
    nsd = 6 if case == "hex" else 3
    for l in range(L - 1):
        rows, cols, *pt = o.amg_matrix(l, "Ptent")
        Pt = csr(rows, cols, *pt)
        B = o.amg_nullspace(l)
        Bc = o.amg_nullspace(l + 1)
        assert B.shape[1] == nsd
        np.testing.assert_allclose(Pt @ Bc, B, atol=1e-12 * np.abs(B).max())
        np.testing.assert_allclose((Pt.T @ Pt).toarray(), np.eye(cols), atol=1e-12)
        A = csr(*o.amg_matrix(l, "A"))
        P = csr(*o.amg_matrix(l, "P"))
        R = csr(*o.amg_matrix(l, "R"))
        assert (abs(R - P.T)).max() == 0
        Ac = csr(*o.amg_matrix(l + 1, "A")).toarray()
        G = (P.T @ A @ P).toarray()
        np.testing.assert_allclose(Ac, 0.5 * (G + G.T), atol=1e-12 * np.abs(G).max())
        assert Ac.shape[0] < A.shape[0]
        # P = (I - omega D^-1 A) P-hat with omega = 4/3 / lambda_max
        lam = o.amg_lambda(l)
        Dinv = sp.diags(1.0 / A.diagonal())
        Pref = Pt - (4.0 / 3.0 / lam) * (Dinv @ A @ Pt)
        np.testing.assert_allclose(P.toarray(), Pref.toarray(), atol=1e-12 * np.abs(Pref).max())
        # lambda_max of D^-1 A is a Lanczos lower bound near the top of the spectrum
        ev = np.linalg.eigvals((Dinv @ A).toarray()).real.max() if A.shape[0] < 3000 else None
        if ev is not None:
            assert 0.7 * ev <= lam <= ev * (1 + 1e-12)


def test_vcycle_symmetric_positive_and_contracting():
    m = M.bench_box(16)
    o = Oracle(m)
    ptr, col, val = o.assemble_kff()
    o.amg_setup()
    A = sp.csr_matrix((val, col, ptr))
    rng = np.random.default_rng(11)
    x, y = rng.standard_normal(o.nfree), rng.standard_normal(o.nfree)
    Mx, My = o.amg_apply(x), o.amg_apply(y)
    assert abs(x @ My - y @ Mx) <= 1e-11 * abs(x @ My)
    assert x @ Mx > 0
    # one cycle reduces the error in the energy norm (test_multigrid.cpp:225-255)
    u = rng.standard_normal(o.nfree)
    b = A @ u
    e1 = u - o.amg_apply(b)
    assert np.sqrt(e1 @ (A @ e1)) < 0.9 * np.sqrt(u @ (A @ u))


def test_reused_hierarchy_equals_fresh_for_scaled_matrix():
    # tests/test_multigrid.cpp:513-542: a uniformly scaled matrix reuses the
    # aggregates and gives the same hierarchy as a fresh setup, bit for bit.
    m = M.bench_box(16)
    a = Oracle(m)
    a.assemble_kff()
    a.amg_setup()
    a.assemble_kff(np.full(m.nelem, 0.5))
    a.amg_setup()
    assert a.amg_reused()
    b = Oracle(m)
    b.assemble_kff(np.full(m.nelem, 0.5))
    b.amg_setup()
    assert not b.amg_reused()
    assert a.amg_levels() == b.amg_levels()
    for l in range(a.amg_levels()):
        for w in ("A", "P"):
            ra, rb = a.amg_matrix(l, w), b.amg_matrix(l, w)
            for u, v in zip(ra[2:], rb[2:]):
                assert np.array_equal(u, v)


@pytest.mark.parametrize("case", ["hex", "quad", "quad_gs"])
def test_amg_pcg_matches_direct_solve(case):
    if case == "hex":
        m = M.cantilever_3d(12, 6, 6)
    else:
        m = M.mbb_beam(60, 20)
    amg = {"smoother": 1} if case == "quad_gs" else {}
    o = Oracle(m, amg=amg)
    rng = np.random.default_rng(2)
    scale = np.where(rng.uniform(size=m.nelem) < 0.5, 1.0, 1e-3)  # SIMP contrast (test_multigrid.cpp:313-358)
    ptr, col, val = o.assemble_kff(scale)
    o.amg_setup()
    A = sp.csr_matrix((val, col, ptr)).tocsc()
    b = m.free_loads()[0]
    ref = spla.spsolve(A, b)
    x, it, rr, cv = o.pcg(b, tol=1e-12, maxit=500)
    assert cv and it < 200
    assert np.linalg.norm(x - ref) <= 1e-7 * np.linalg.norm(ref)
    assert b @ x == pytest.approx(b @ ref, rel=1e-10)
    # warm start from the answer: no iterations (test_multigrid.cpp:360-404)
    x2, it2, _, _ = o.pcg(b, x0=x, tol=1e-8)
    assert it2 == 0 and np.array_equal(x2, x)


# ---------------------------------------------------------------------------
# Filter, sensitivity, OC (DensityFilter.cpp, Sensitivity.cpp, OptimalityCriteria.cpp)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim,radius", [(2, 1.5), (3, 1.5), (3, 2.0)])
def test_filter_brute_force(dim, radius):
    if dim == 3:
        m = M.cantilever_3d(5, 4, 3, radius_cells=radius)
    else:
        m = M.mbb_beam(9, 5, radius_cells=radius)
    o = Oracle(m)
    ptr, col, val, rs = o.filter_matrix()
    c = m.element_centroids()
    d = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    W = np.maximum(0.0, radius - d)
    H = W / W.sum(1, keepdims=True)
    np.testing.assert_allclose(csr(m.nelem, m.nelem, ptr, col, val).toarray(), H, rtol=0, atol=1e-15)
    np.testing.assert_allclose(val.reshape(-1).sum() / m.nelem, 1.0, rtol=1e-14)
    x, v = np.random.default_rng(0).uniform(size=(2, m.nelem))
    assert abs(o.filter(x) @ v - x @ o.filter(v, adjoint=True)) <= 1e-12 * abs(x @ v)
    np.testing.assert_allclose(o.filter(np.ones(m.nelem)), 1.0, rtol=1e-14)


def test_filter_nnz_closed_form():
    # lattice offsets within the radius: 19 at r 1.5 and 27 at r 2.0 in 3-D
    m = M.cantilever_3d(64, 32, 32, radius_cells=2.0)
    nx, ny, nz = 64, 32, 32
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    assert sum((nx - abs(a)) * (ny - abs(b)) * (nz - abs(c)) for a, b, c in offs) == 1678840


def _fd_check(m, rng_seed, tol, elems):
    o = Oracle(m, tol=1e-13, warm_start=False)
    x = np.random.default_rng(rng_seed).uniform(0.3, 0.9, m.nelem)
    ev = o.evaluate(x)
    dc = ev["dc_dx"]
    assert (dc <= 0).all() and (ev["dv_dx"] > 0).all()
    assert ev["energy"].sum() == pytest.approx(0.5 * ev["compliance"], rel=1e-10)
    h = 1e-6
    for e in elems:
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (o.evaluate(xp, False)["compliance"] - o.evaluate(xm, False)["compliance"]) / (2 * h)
        assert abs(fd - dc[e]) <= tol * max(abs(dc[e]), 1e-12 * np.abs(dc).max())


def test_sensitivity_finite_difference_q4():
    m = M.mbb_beam(12, 6)  # tests/test_topopt.cpp:373-435 (12 x 6 Q4)
    _fd_check(m, 4, 1e-5, [0, 7, 30, 71])


def test_sensitivity_finite_difference_hex8():
    m = M.cantilever_3d(6, 3, 2, radius_cells=1.5)  # tests/test_solid.cpp:947-1019
    _fd_check(m, 9, 1e-5, [0, 13, 35])


def test_oc_meets_volume_and_limits():
    m = M.mbb_beam(30, 15)
    o = Oracle(m)
    x = m.initial_design()
    ev = o.evaluate(x)
    xn, info = o.oc(x, ev["dc_dx"], ev["dv_dx"])
    assert info["converged"]
    vol = (np.clip(o.filter(xn), 0, 1) * o.volumes()).sum()
    target = m.volume_fraction * o.volumes().sum()
    assert abs(vol - target) <= 1e-8 * target
    assert (xn >= 0).all() and (xn <= 1).all() and (np.abs(xn - x) <= 0.2 + 1e-15).all()


def test_short_oc_run_improves_design():
    # tests/test_topopt.cpp:670-727 (30 x 15 end to end)
    m = M.mbb_beam(30, 15)
    m.volume_fraction = 0.4
    o = Oracle(m)
    x, hist, n, fc = o.run_oc(m.initial_design(), 40, change_tol=0.01)
    assert hist[0, 2] / fc > 1.5
    assert abs(hist[-1, 4] - 0.4) < 1e-6
