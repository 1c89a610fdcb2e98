"""The modal oracle (oracle/modal.py) pinned on the CPU.

Known answers: the element mass matrices' textbook closed forms, mass
conservation (tests/test_modal.cpp:14-66 of the reference), a dense
generalised eigensolve (test_modal.cpp:90-132), Euler-Bernoulli bending and
fixed-free rod theory (test_modal.cpp:134-230), thickness invariance
(test_modal.cpp:248-280) and the reference's published full-solid first
frequencies and masses of four benchmark decks (docs/results/benchmarks.csv).
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from oracle import modal as OM
from oracle.refcpu import Oracle
from sparlab_amd import deck as D
from sparlab_amd.model import StructuredModel

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
PUBLISHED = json.load(open(os.path.join(GOLD, "benchmarks_published.json")))["cases"]
MODAL_DECKS = ("cantilever_beam", "aerospace_bracket", "bracket_3d", "wing_rib")

# CantileverCase of the reference's tests (tests/TestSupport.hpp:23-31)
L_, H_, T_, E_, RHO = 1.0, 0.1, 0.01, 70.0e9, 2700.0


def cantilever(nx, ny, nu=0.0, thickness=T_):
    m = StructuredModel(dim=2, nx=nx, ny=ny, lx=L_, ly=H_, E=E_, nu=nu, thickness=thickness)
    m.fix(m.select_box(xmax=0.0))
    return m


def operators(m, density=RHO, lumped=False):
    o = Oracle(m)
    free = o.free()
    return OM.kff_csr(o), OM.reduce_free(OM.assemble_mass(m, density, lumped=lumped), free), free


def euler_bernoulli(mode):
    bl = (1.87510407, 4.69409113)[mode - 1]
    return bl * bl / (2 * np.pi * L_ * L_) * np.sqrt(E_ * H_ * H_ / (12.0 * RHO))


def test_element_mass_closed_forms():
    q4 = cantilever(4, 2)
    a = q4.cell[0] * q4.cell[1]
    me = OM.element_mass(q4, RHO)
    ref = RHO * T_ * a / 36.0 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]], dtype=float)
    assert np.allclose(me[0::2, 0::2], ref, rtol=1e-14, atol=0)
    assert np.allclose(me[1::2, 1::2], ref, rtol=1e-14, atol=0)
    assert np.all(me[0::2, 1::2] == 0.0)
    hx = StructuredModel(dim=3, nx=2, ny=3, nz=4, lx=0.2, ly=0.3, lz=0.5)
    v = 0.1 * 0.1 * 0.125
    corners = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
    same = (corners[:, None, :] == corners[None, :, :]).sum(axis=2)
    ref3 = RHO * v / 216.0 * 2.0 ** same
    me3 = OM.element_mass(hx, RHO)
    for c in range(3):
        assert np.allclose(me3[c::3, c::3], ref3, rtol=1e-14, atol=0)
    assert np.allclose(OM.element_mass(hx, RHO, lumped=True).sum(axis=1), RHO * v / 8)


@pytest.mark.parametrize("lumped", [False, True])
def test_mass_conservation(lumped):
    m = cantilever(6, 4)
    M = OM.assemble_mass(m, RHO, lumped=lumped)
    expected = RHO * L_ * H_ * T_
    assert abs(M - M.T).max() == 0.0
    assert M.sum() == pytest.approx(2 * expected, rel=1e-12)
    tx = np.zeros(m.ndofs)
    tx[0::2] = 1.0
    assert tx @ (M @ tx) == pytest.approx(expected, rel=1e-12)
    assert OM.total_mass(m, RHO) == pytest.approx(expected, rel=1e-12)
    free = m.free_dofs()
    assert np.linalg.eigvalsh(OM.reduce_free(M, free).toarray()).min() > 0.0
    scale = np.linspace(0.1, 1.0, m.nelem)
    assert OM.total_mass(m, RHO, scale) == pytest.approx(RHO * T_ * (L_ / 6) * (H_ / 4) * scale.sum(), rel=1e-12)


def test_subspace_iteration_agrees_with_dense_eigensolve():
    m = cantilever(40, 8, nu=0.3)
    K, M, free = operators(m)
    assert K.shape[0] > 400
    r = OM.solve_modal(K, M, num_modes=5, tolerance=1e-12)
    dense = sla.eigh(K.toarray(), M.toarray(), eigvals_only=True)[:5]
    assert np.allclose(r["eigenvalues"], dense, rtol=1e-8, atol=0)
    assert np.all(np.diff(r["eigenvalues"]) >= 0)
    assert r["residuals"].max() < 1e-8
    assert np.allclose(np.einsum("ij,ij->j", r["phi"], M @ r["phi"]), 1.0, rtol=1e-8)


def test_bending_frequencies_match_euler_bernoulli():
    K, M, _ = operators(cantilever(120, 12))
    f = OM.solve_modal(K, M, num_modes=4)["frequencies_hz"]
    f1, f1t = f[0], euler_bernoulli(1)
    assert abs(f1 - f1t) / f1t < 0.01 and f1 < f1t
    f2t = euler_bernoulli(2)
    f2 = f[1:][np.argmin(np.abs(f[1:] - f2t))]
    assert abs(f2 - f2t) / f2t < 0.06 and f2 < f2t


def test_first_axial_mode_matches_rod_theory():
    m = cantilever(160, 16)
    K, M, free = operators(m)
    r = OM.solve_modal(K, M, num_modes=8)
    shapes = np.zeros((m.ndofs, 8))
    shapes[free] = r["phi"]
    ex = (shapes[0::2] ** 2).sum(axis=0)
    ey = (shapes[1::2] ** 2).sum(axis=0)
    ratio = ex / (ex + ey)
    axial = int(np.argmax(ratio))
    assert ratio[axial] > 0.9
    expected = np.sqrt(E_ / RHO) / (4.0 * L_)
    assert abs(r["frequencies_hz"][axial] - expected) / expected < 0.01


def test_frequencies_invariant_under_thickness_scaling():
    fa = OM.solve_modal(*operators(cantilever(40, 8))[:2], num_modes=3)["frequencies_hz"]
    fb = OM.solve_modal(*operators(cantilever(40, 8, thickness=0.3 * T_))[:2], num_modes=3)["frequencies_hz"]
    assert np.allclose(fa, fb, rtol=1e-9, atol=0)


@pytest.mark.parametrize("case", MODAL_DECKS)
def test_published_full_solid_frequency_and_mass(case):
    d = D.load_deck(os.path.join(GOLD, "decks", case + ".json"))
    pub = PUBLISHED[case]
    assert d.modal is not None and d.density > 0
    K, M, _ = operators(d.model, d.density)
    assert OM.total_mass(d.model, d.density) == pytest.approx(pub["mass_solid_kg"], rel=1e-11)
    lam = spla.eigsh(K, k=1, M=M, sigma=0, which="LM", return_eigenvectors=False)
    assert np.sqrt(lam.min()) / (2 * np.pi) == pytest.approx(pub["f1_solid_Hz"], rel=1e-10)
    if case != "aerospace_bracket":  # the largest deck: eigsh alone keeps the CPU suite short
        r = OM.solve_modal(K, M, num_modes=d.modal["num_modes"])
        assert r["frequencies_hz"][0] == pytest.approx(pub["f1_solid_Hz"], rel=1e-10)
        assert r["residuals"].max() < d.modal["residual_tolerance"]
