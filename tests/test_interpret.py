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
from sparlab_amd import deck as D
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


@pytest.mark.parametrize("dim,seed", [(2, 5), (3, 6)])
def test_submesh_numbering_is_first_appearance(dim, seed):
    m = (StructuredModel(dim=2, nx=9, ny=7, lx=0.9, ly=0.7) if dim == 2 else
         StructuredModel(dim=3, nx=5, ny=4, nz=4, lx=0.5, ly=0.4, lz=0.4))
    keep = field((m.nelem,), seed) > 0.4
    order, _ = OM.submesh_node_order(m, np.flatnonzero(keep))
    assert np.array_equal(I.submesh_nodes(m, keep), order)


def test_supports_are_reapplied_on_the_extracted_structure():
    """The deck's boundary conditions evaluated on the retained nodes in the
    sub-mesh numbering (apps/sparlab_topopt.cpp:109): a box keeps its nodes, a
    nearest_node support whose parent node was dropped moves to the nearest
    retained node, node_ids name sub-mesh nodes, and a region that selects no
    retained node is an error."""
    m = StructuredModel(dim=2, nx=6, ny=4, lx=6.0, ly=4.0)
    bcs = [dict(region=D.parse_region({"box": {"xmax": 0.0}}, 2, "a"), comps=[0, 1], name="root"),
           dict(region=D.parse_region({"nearest_node": [6.0, 0.2]}, 2, "b"), comps=[1], name="roller"),
           dict(region=D.parse_region({"node_ids": [5]}, 2, "c"), comps=[0], name="pin")]
    m.fixed = D.apply_constraints(D.Geometry(m), bcs).astype(np.uint8).ravel()
    keep = np.ones(m.nelem, dtype=bool)
    keep[5] = False  # the bottom-right element: the only one touching node (6, 0)
    node = lambda i, j: j * 7 + i  # noqa: E731
    parent = m.fixed.reshape(-1, 2).astype(bool)
    assert parent[node(6, 0), 1] and parent[5, 0]  # the parent's roller and pin
    sub, used = I.solid_submodel(m, keep, bcs)
    f = sub.fixed.reshape(-1, 2).astype(bool)
    assert not used[node(6, 0)] and f[node(6, 0)].all()  # dropped node: every DOF fixed
    # roller: (6, 1) at distance 0.8 beats (5, 0) at 1.02
    assert f[node(6, 1), 1] and not f[node(6, 1), 0]
    # pin: sub-mesh node 5 is parent node (2, 1) (element 1's third corner, first seen there)
    assert I.submesh_nodes(m, keep)[5] == node(2, 1) and f[node(2, 1), 0] and not f[node(2, 1), 1]
    assert not f[5, 0]  # the parent's node 5 keeps no pin
    root = [node(0, j) for j in range(5)]
    assert f[root].all()
    expected = np.zeros_like(f)
    expected[root] = True
    expected[node(6, 1), 1] = expected[node(2, 1), 0] = True
    expected[~used] = True
    assert np.array_equal(f, expected)
    # the parent's own selection is what the call without constraints keeps
    sub0, _ = I.solid_submodel(m, keep)
    f0 = sub0.fixed.reshape(-1, 2).astype(bool)
    assert f0[5, 0] and not f0[node(6, 1), 1]
    bad = bcs + [dict(region=D.parse_region({"box": {"xmin": 10.0}}, 2, "d"), comps=[0], name="far")]
    with pytest.raises(D.DeckError, match="boundary condition 'far' selected no nodes"):
        I.solid_submodel(m, keep, bad)


def test_oracle_face_components_known_answers():
    """The oracle's union-find on hand-built cases: corner or edge contact is no connection,
    a shared face (an edge in 2-D) is one."""
    q = StructuredModel(dim=2, nx=4, ny=4, lx=1.0, ly=1.0)
    assert len(OM.components_by_face(q, [0, 5])) == 2        # diagonal neighbours: a corner
    assert len(OM.components_by_face(q, [0, 1, 5])) == 1     # an L through a shared edge
    assert len(OM.components_by_face(q, [15, 0, 10, 5])) == 4  # a diagonal chain touches only at corners
    h = StructuredModel(dim=3, nx=3, ny=3, nz=3, lx=1.0, ly=1.0, lz=1.0)
    e = lambda i, j, k: (k * 3 + j) * 3 + i
    assert len(OM.components_by_face(h, [e(0, 0, 0), e(1, 1, 0)])) == 2   # a shared edge only
    assert len(OM.components_by_face(h, [e(0, 0, 0), e(1, 1, 1)])) == 2   # a shared corner only
    assert len(OM.components_by_face(h, [e(0, 0, 0), e(0, 0, 1)])) == 1   # a shared face
    assert OM.components_by_face(h, [e(2, 2, 2), e(0, 0, 0), e(1, 0, 0)]) == [[e(2, 2, 2)], [e(0, 0, 0), e(1, 0, 0)]]


def test_oracle_submesh_of_every_element_is_the_full_model():
    """Extracting every element only renumbers the nodes (by first appearance): the sub-mesh
    operators are the full K_ff and M_ff permuted accordingly."""
    m = StructuredModel(dim=2, nx=6, ny=3, lx=0.6, ly=0.3, E=70e9, nu=0.3, thickness=0.01)
    m.fix(m.select_box(xmax=0.0))
    o = Oracle(m)
    K = OM.kff_csr(o)
    M = OM.reduce_free(OM.assemble_mass(m, RHO), o.free())
    Ks, Ms = OM.submesh_operators(m, o.ke0(), RHO, np.arange(m.nelem), m.fixed)
    flat = OM.element_nodes(m).ravel()
    _, first = np.unique(flat, return_index=True)
    order = flat[np.sort(first)]                      # old node of each new node
    dofs = (order[:, None] * 2 + np.arange(2)).ravel()  # old DOF of each new DOF
    fixed = np.asarray(m.fixed, dtype=bool)
    new_free_old = dofs[~fixed[dofs]]                 # old DOF of each new free DOF
    pos = {d: i for i, d in enumerate(o.free())}
    p = np.array([pos[d] for d in new_free_old])
    Kp, Mp = K[p][:, p], M[p][:, p]
    assert abs(Kp - Ks).max() <= 1e-12 * abs(K).max()
    assert abs(Mp - Ms).max() <= 1e-15 * abs(M).max()
