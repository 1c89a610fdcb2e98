"""slx_modal_solve on the GPU against the modal oracle and the published decks.

The device path (mass products, projections, Ritz rotation and residuals on the
GPU, warm-started AMG-PCG for the shifted systems) is compared with the CPU
restatement (oracle/modal.py: sparse direct solves, scipy eigh) on small Q4 and
Hex8 models, with and without per-element stiffness and mass factors, and with
the reference's published full-solid first frequencies and masses of the four
benchmark decks that enable modal analysis.
"""
import json
import os

import numpy as np
import pytest

from oracle import modal as OM
from oracle.refcpu import Oracle
from sparlab_amd import deck as D
from sparlab_amd.model import StructuredModel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
PUBLISHED = json.load(open(os.path.join(GOLD, "benchmarks_published.json")))["cases"]
RHO = 2700.0


def q4_cantilever(nx=40, ny=8):
    m = StructuredModel(dim=2, nx=nx, ny=ny, lx=1.0, ly=0.1, E=70e9, nu=0.3, thickness=0.01)
    m.fix(m.select_box(xmax=0.0))
    return m


def hex8_cantilever(nx=16, ny=6, nz=5):
    m = StructuredModel(dim=3, nx=nx, ny=ny, nz=nz, lx=0.32, ly=0.12, lz=0.1, E=71.7e9, nu=0.33)
    m.fix(m.select_box(xmax=0.0))
    return m


def oracle_modes(m, num_modes, kscale=None, mscale=None, lumped=False):
    o = Oracle(m)
    free = o.free()
    K = OM.kff_csr(o, kscale)
    M = OM.reduce_free(OM.assemble_mass(m, RHO, mscale, lumped=lumped), free)
    return OM.solve_modal(K, M, num_modes=num_modes, tolerance=1e-12), M, free


def check_against_oracle(m, num_modes, kscale=None, mscale=None, lumped=False):
    from sparlab_amd import native as N
    g = N.HipProblem(m, device=0)
    r = g.modal(RHO, num_modes=num_modes, mass_type="lumped" if lumped else "consistent", stiffness_scale=kscale,
                mass_scale=mscale, shapes=True)
    ref, M, free = oracle_modes(m, num_modes, kscale, mscale, lumped)
    print(f"dim {m.dim}, {len(free)} free DOFs: {r['iterations']} subspace iterations, "
          f"{r['linear_iterations']} CG iterations, f {np.round(r['frequencies_hz'], 3)}", flush=True)
    assert r["converged"] and r["num_modes"] == num_modes
    assert np.allclose(r["eigenvalues_1_per_s2"], ref["eigenvalues"], rtol=1e-9, atol=0)
    assert max(r["eigenpair_residuals"]) < 1e-6
    scale = np.ones(m.nelem) if mscale is None else mscale
    assert r["total_mass_kg"] == pytest.approx(RHO * float(np.prod(m.cell)) * (m.thickness if m.dim == 2 else 1.0)
                                               * scale.sum(), rel=1e-12)
    phi = r["mode_shapes"]
    fixed = np.setdiff1d(np.arange(m.ndofs), free)
    assert np.all(phi[fixed] == 0.0)
    pf = phi[free]
    gram = pf.T @ (M @ pf)
    assert np.allclose(gram, np.eye(num_modes), atol=1e-7)
    # the same modes as the oracle's, up to sign (distinct eigenvalues)
    cross = np.abs(np.einsum("ij,ij->j", pf, M @ ref["phi"]))
    assert np.allclose(cross, 1.0, atol=1e-6)
    return g, r


def test_q4_modes_match_the_oracle():
    check_against_oracle(q4_cantilever(), 5)


def test_small_systems_take_the_dense_eigensolve():
    """At most 400 free DOFs (or a subspace that covers the system): the dense
    generalised eigensolve of ModalAnalysis.cpp:166-179, no subspace iterations
    and no linear solves; more modes than free DOFs are clamped with the
    reference's warning (ModalAnalysis.cpp:153-160)."""
    m = q4_cantilever(20, 4)  # 200 free DOFs
    g, r = check_against_oracle(m, 6)
    assert r["iterations"] == 0 and r["linear_iterations"] == 0 and r["final_relative_change"] == 0.0
    assert r["warnings"] == []
    tiny = q4_cantilever(2, 1)  # 8 free DOFs
    from sparlab_amd import native as N
    gt = N.HipProblem(tiny, device=0)
    with pytest.warns(RuntimeWarning, match="requested 10 modes but the model has only 8 free DOFs; returning 8"):
        rt = gt.modal(RHO, num_modes=10, shapes=True)
    assert rt["num_modes"] == 8 and rt["iterations"] == 0
    assert rt["warnings"] == ["requested 10 modes but the model has only 8 free DOFs; returning 8"]
    ref, M, free = oracle_modes(tiny, 8)
    assert np.allclose(rt["eigenvalues_1_per_s2"], ref["eigenvalues"], rtol=1e-9, atol=0)
    pf = rt["mode_shapes"][free]
    assert np.allclose(pf.T @ (M @ pf), np.eye(8), atol=1e-9)


def test_subspace_solves_without_warm_start():
    """modal warm_start off (linear.warm_start false): every shifted solve starts
    from zero; the modes are the same, the solves take more CG iterations."""
    from sparlab_amd import native as N
    m = q4_cantilever()
    g = N.HipProblem(m, device=0)
    warm = g.modal(RHO, num_modes=3)
    cold = g.modal(RHO, num_modes=3, warm_start=False)
    print(f"linear iterations warm {warm['linear_iterations']}, cold {cold['linear_iterations']}", flush=True)
    assert np.allclose(cold["eigenvalues_1_per_s2"], warm["eigenvalues_1_per_s2"], rtol=1e-9, atol=0)
    assert abs(cold["iterations"] - warm["iterations"]) <= 1
    assert cold["linear_iterations"] > warm["linear_iterations"]


def test_q4_modes_with_element_factors_and_lumped_mass():
    m = q4_cantilever(30, 6)
    rng = np.random.default_rng(20261203)
    ks = rng.uniform(0.3, 1.0, m.nelem)
    ms = rng.uniform(0.3, 1.0, m.nelem)
    check_against_oracle(m, 4, ks, ms)
    check_against_oracle(m, 4, lumped=True)


def test_hex8_modes_match_the_oracle():
    m = hex8_cantilever()
    rng = np.random.default_rng(7)
    ks = rng.uniform(0.5, 1.0, m.nelem)
    check_against_oracle(m, 6)
    check_against_oracle(m, 6, ks, ks)


@pytest.mark.parametrize("case", ("cantilever_beam", "aerospace_bracket", "bracket_3d", "wing_rib"))
def test_published_full_solid_frequency(case):
    from sparlab_amd import native as N
    d = D.load_deck(os.path.join(GOLD, "decks", case + ".json"))
    pub = PUBLISHED[case]
    g = N.HipProblem(d.model, device=0)
    r = g.modal(d.density, **d.modal)
    print(f"deck {case}: f1 {r['frequencies_hz'][0]!r} Hz (published {pub['f1_solid_Hz']!r}), "
          f"{r['iterations']} subspace iterations, {r['linear_iterations']} CG iterations, {r['seconds']:.2f} s",
          flush=True)
    assert r["converged"]
    assert r["frequencies_hz"][0] == pytest.approx(pub["f1_solid_Hz"], rel=1e-9)
    assert r["total_mass_kg"] == pytest.approx(pub["mass_solid_kg"], rel=1e-11)
    assert max(r["eigenpair_residuals"]) < d.modal["residual_tolerance"]
    assert r["warnings"] == []
    # the context re-assembles its own operator afterwards
    solid = g.evaluate(np.ones(d.model.nelem), tol=1e-11, need_gradients=False, warm_start=False)["compliance"]
    assert solid == pytest.approx(pub["full_solid_compliance_J"], rel=1e-8)


def test_modal_errors():
    from sparlab_amd import native as N
    g = N.HipProblem(q4_cantilever(16, 4), device=0)  # 160 free DOFs
    with pytest.raises(N.ModelError):
        g.modal(0.0)
    with pytest.raises(N.ConfigError):
        g.modal(RHO, shift_factor=-0.1)
    with pytest.raises(N.ConfigError):
        g.modal(RHO, num_modes=0)
    # the subspace iteration holds at most 64 vectors; a small system takes the
    # dense eigensolve, which has no such limit
    assert g.modal(RHO, num_modes=33)["num_modes"] == 33
    big = N.HipProblem(q4_cantilever(), device=0)  # 720 free DOFs
    with pytest.raises(N.ConfigError):
        big.modal(RHO, num_modes=33)  # a subspace of 66 > 64
    with pytest.raises(N.ModelError):
        g.modal(RHO, mass_scale=-np.ones(64))


def test_modes_with_the_mixed_precision_hierarchy():
    """The shifted solves take the context's AMG options: with the FP32 V-cycle under FP64
    flexible CG the eigenpairs are the same to the solve tolerance."""
    from sparlab_amd import native as N
    m = hex8_cantilever()
    g = N.HipProblem(m, device=0)
    g.assemble(np.ones(m.nelem))
    g.amg_setup(precision=1)
    r = g.modal(RHO, num_modes=4, linear_tolerance=1e-10)
    ref, _, _ = oracle_modes(m, 4)
    print(f"mixed precision: {r['iterations']} subspace iterations, {r['linear_iterations']} CG iterations", flush=True)
    assert r["converged"]
    assert np.allclose(r["eigenvalues_1_per_s2"], ref["eigenvalues"], rtol=1e-9, atol=0)
