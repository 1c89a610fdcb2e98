"""Interpretation of a density field as a solid (sparlab_amd.interpret) on the CPU.

The grid labelling is checked against the oracle's restatement of the
reference's face-keyed union-find (oracle/modal.py components_by_face,
src/mesh/SubMesh.cpp:38-90), and the masked-grid operators of the retained
structure against the operators of the extracted, renumbered sub-mesh
(extract_element_subset, SubMesh.cpp:96-140): same eigenvalues.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import modal as OM
from oracle.refcpu import Oracle
from sparlab_amd import interpret as I
from sparlab_amd.model import StructuredModel

RHO = 2700.0


def field(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, int(np.prod(shape)))


def reference_interpretation(model, density, threshold=0.5):
    kept = [e for e in range(model.nelem) if density[e] >= threshold]
    comps = OM.components_by_face(model, kept)
    best, bestv = 0, -1.0
    for c, members in enumerate(comps):
        if len(members) > bestv:
            best, bestv = c, len(members)
    return kept, comps, sorted(comps[best])


@pytest.mark.parametrize("dim,seed", [(2, 1), (2, 2), (3, 3), (3, 4)])
def test_components_match_the_union_find_restatement(dim, seed):
    m = (StructuredModel(dim=2, nx=14, ny=9, lx=1.4, ly=0.9, thickness=0.02) if dim == 2 else
         StructuredModel(dim=3, nx=7, ny=6, nz=5, lx=0.7, ly=0.6, lz=0.5))
    rho = field((m.nelem,), seed)
    it = I.interpret_density_as_solid(m, rho, 0.55)
    kept, comps, largest = reference_interpretation(m, rho, 0.55)
    assert it.elements_above_threshold == len(kept)
    assert it.components_above_threshold == len(comps) and len(comps) > 1
    assert np.flatnonzero(it.retained).tolist() == largest
    ve = float(np.prod(m.cell)) * (m.thickness if dim == 2 else 1.0)
    assert it.volume_retained == pytest.approx(len(largest) * ve, rel=1e-14)
    assert it.volume_discarded_as_islands == pytest.approx((len(kept) - len(largest)) * ve, rel=1e-12, abs=1e-18)


def test_corner_contact_is_not_a_connection_and_ties_go_to_the_first_group():
    m = StructuredModel(dim=2, nx=4, ny=4, lx=1.0, ly=1.0)
    rho = np.zeros(16)
    rho[[0, 1, 4, 5]] = 1.0     # a 2 x 2 block at the origin
    rho[[10, 11, 14, 15]] = 1.0  # a 2 x 2 block touching it only at a corner node
    it = I.interpret_density_as_solid(m, rho)
    assert it.components_above_threshold == 2
    assert np.flatnonzero(it.retained).tolist() == [0, 1, 4, 5]
    with pytest.raises(I.InterpretationError):
        I.interpret_density_as_solid(m, np.full(16, 0.2))


@pytest.mark.parametrize("dim", [2, 3])
def test_masked_grid_operators_equal_the_extracted_submesh(dim):
    if dim == 2:
        m = StructuredModel(dim=2, nx=16, ny=8, lx=1.6, ly=0.8, E=70e9, nu=0.3, thickness=0.01)
    else:
        m = StructuredModel(dim=3, nx=8, ny=5, nz=4, lx=0.4, ly=0.25, lz=0.2, E=70e9, nu=0.3)
    m.fix(m.select_box(xmax=0.0))
    rho = field((m.nelem,), 11 + dim)
    rho[m.element_centroids()[:, 0] < m.cell[0]] = 1.0  # keep the clamped column
    it = I.interpret_density_as_solid(m, rho, 0.4)
    sub, used = I.solid_submodel(m, it.retained)
    assert I.supported(m, used)
    ind = it.retained.astype(float)
    o = Oracle(sub)
    K = OM.kff_csr(o, ind)
    M = OM.reduce_free(OM.assemble_mass(sub, RHO, ind), o.free())
    Ks, Ms = OM.submesh_operators(m, o.ke0(), RHO, np.flatnonzero(it.retained), m.fixed)
    assert K.shape == Ks.shape
    a = sla.eigh(K.toarray(), M.toarray(), eigvals_only=True)[:6]
    b = sla.eigh(Ks.toarray(), Ms.toarray(), eigvals_only=True)[:6]
    assert np.allclose(a, b, rtol=1e-10, atol=0)
    assert OM.total_mass(sub, RHO, ind) == pytest.approx(it.volume_retained * RHO, rel=1e-12)
