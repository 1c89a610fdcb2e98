"""CPU restatement of the free-vibration analysis of structured models.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

* ``element_mass``: Element::consistent_mass of a Q4 rectangle or a Hex8 brick
  (rho t int N^T N dA, rho int N^T N dV), integrated here with 2-point Gauss
  rules per axis, which are exact for the products of (bi/tri)linear shape
  functions. Row-sum lumping gives rho V / npe per node and component.
* ``assemble_mass``: Assembler::assemble_mass (src/fem/Assembler.cpp:260-330),
  per-element factors included, then the free-free block.
* ``solve_modal``: solve_modal (src/fem/ModalAnalysis.cpp:120-418): the
  starting subspace (ModalAnalysis.cpp:52-86), shift-invert subspace iteration
  with a sparse direct factorisation (the reference's SimplicialLDLT path),
  Rayleigh-Ritz through scipy.linalg.eigh (lowest_pencil_pairs,
  ModalAnalysis.cpp:27-50), the relative-change and eigenpair-residual
  stopping rule, M-normalisation and total mass.

* ``components_by_face`` / ``submesh_operators``: components_by_face and
  extract_element_subset (src/mesh/SubMesh.cpp:38-140), pinned in
  tests/test_interpret.py on hand-built contact cases (corner, edge, face) and
  on the every-element sub-mesh, which must equal the permuted full operators.

The modal functions are pinned in tests/test_modal_oracle.py against: the reference's published
full-solid first frequencies and masses of the benchmark decks
(docs/results/benchmarks.csv), shift-invert Lanczos (scipy eigsh) and dense
generalised eigensolves, Euler-Bernoulli and fixed-free rod theory
(tests/test_modal.cpp:134-230), and mass conservation (test_modal.cpp:14-66).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_G = 1.0 / np.sqrt(3.0)


def _corners(dim):
    """Natural coordinates of the element corners in the local node order
    (src/mesh/StructuredMesh.hpp: counter-clockwise, bottom layer first)."""
    q = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
    if dim == 2:
        return np.array(q, dtype=float)
    return np.array([(a, b, -1) for a, b in q] + [(a, b, 1) for a, b in q], dtype=float)


def element_mass(model, density, lumped=False):
    """The element mass matrix (npe*dim square, DOF = node*dim + component)."""
    dim = model.dim
    cell = np.array(model.cell[:dim], dtype=float)
    corners = _corners(dim)
    npe = len(corners)
    vol = float(np.prod(cell)) * (model.thickness if dim == 2 else 1.0)
    if lumped:
        return np.kron(np.eye(npe), np.eye(dim)) * density * vol / npe
    detj = vol / 2.0 ** dim  # d(x)/d(xi) = h / 2 per axis (times t in 2-D)
    mn = np.zeros((npe, npe))
    pts = np.array(np.meshgrid(*[[-_G, _G]] * dim, indexing="ij")).reshape(dim, -1).T
    for xi in pts:
        n = np.prod(1.0 + corners * xi, axis=1) / 2.0 ** dim
        mn += np.outer(n, n) * detj  # weights 1
    return np.kron(mn, np.eye(dim)) * density


def element_nodes(model):
    nnx, nny = model.nx + 1, model.ny + 1
    e = np.arange(model.nelem)
    ex = e % model.nx
    if model.dim == 2:
        ey = e // model.nx
        n0 = ey * nnx + ex
        return np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx], axis=1)
    ey = (e // model.nx) % model.ny
    ez = e // (model.nx * model.ny)
    lay = nnx * nny
    n0 = (ez * nny + ey) * nnx + ex
    return np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx, n0 + lay, n0 + lay + 1, n0 + lay + nnx + 1, n0 + lay + nnx],
                    axis=1)


def assemble_mass(model, density, scale=None, lumped=False):
    """Full M (ndofs x ndofs CSR); scale: per-element factors."""
    dim = model.dim
    me = element_mass(model, density, lumped)
    nodes = element_nodes(model)
    dofs = (nodes[:, :, None] * dim + np.arange(dim)[None, None, :]).reshape(model.nelem, -1)
    s = np.ones(model.nelem) if scale is None else np.asarray(scale, dtype=float)
    rows = np.repeat(dofs, dofs.shape[1], axis=1).ravel()
    cols = np.tile(dofs, (1, dofs.shape[1])).ravel()
    vals = (s[:, None, None] * me[None]).reshape(model.nelem, -1).ravel()
    return sp.csr_matrix((vals, (rows, cols)), shape=(model.ndofs, model.ndofs))


def reduce_free(a, free):
    return a[free][:, free].tocsr()


def starting_subspace(kd, md, q):
    """ModalAnalysis.cpp:52-86 (q <= n: the seeded random columns are never reached)."""
    n = len(kd)
    x = np.zeros((n, q))
    x[:, 0] = md
    ratio = np.where(kd > 0.0, md / np.where(kd > 0.0, kd, 1.0), 0.0)
    order = np.argsort(-ratio, kind="stable")
    for col in range(1, q):
        x[order[col - 1], col] = 1.0
    return x


def solve_modal(K, M, num_modes=6, max_iterations=200, tolerance=1e-10, residual_tolerance=1e-6):
    """Lowest eigenpairs of (K, M) (free-free sparse matrices). Returns a dict with
    eigenvalues, frequencies_hz, residuals, phi (n x m, M-orthonormal), iterations."""
    n = K.shape[0]
    m = min(num_modes, n)
    q = min(n, max(2 * m, m + 8))
    K = K.tocsc()
    M = M.tocsr()
    if q >= n or n <= 400:
        lam, vec = sla.eigh(K.toarray(), M.toarray())
        lam, phi, iters = lam[:m], vec[:, :m], 0
    else:
        lu = spla.splu(K)
        x = starting_subspace(K.diagonal(), M.diagonal(), q)
        prev = np.full(m, np.finfo(float).max)
        converged = False
        for it in range(1, max_iterations + 1):
            xbar = lu.solve(M @ x)
            kr = xbar.T @ (K @ xbar)
            mr = xbar.T @ (M @ xbar)
            mu, ritz = sla.eigh(0.5 * (kr + kr.T), 0.5 * (mr + mr.T))
            x = xbar @ ritz
            change = float(np.max(np.abs(mu[:m] - prev) / np.maximum(np.abs(mu[:m]), 1e-300)))
            prev = mu[:m].copy()
            mx = M @ x[:, :m]
            res = np.linalg.norm(K @ x[:, :m] - mx * mu[:m], axis=0) / np.maximum(
                np.linalg.norm(mx * mu[:m], axis=0), 1e-300)
            iters = it
            if change < tolerance and res.max() < residual_tolerance:
                converged = True
                break
        if not converged:
            raise RuntimeError("oracle subspace iteration did not converge")
        lam, phi = prev, x[:, :m]
    phi = phi / np.sqrt(np.einsum("ij,ij->j", phi, M @ phi))
    mphi = M @ phi
    res = np.linalg.norm(K @ phi - mphi * lam, axis=0) / np.maximum(np.linalg.norm(mphi * lam, axis=0), 1e-300)
    omega = np.sqrt(np.maximum(lam, 0.0))
    return dict(eigenvalues=lam, frequencies_hz=omega / (2.0 * np.pi), residuals=res, phi=phi, iterations=iters,
                subspace_size=q)


def total_mass(model, density, scale=None):
    """sum(M) / dim (ModalAnalysis.cpp:139-150)."""
    return float(assemble_mass(model, density, scale).sum()) / model.dim


def kff_csr(oracle, scale=None):
    """The oracle's assembled K_ff as a scipy matrix."""
    ptr, col, val = oracle.assemble_kff(scale)
    n = len(ptr) - 1
    return sp.csr_matrix((val, col, ptr), shape=(n, n))


def components_by_face(model, subset):
    """components_by_face (src/mesh/SubMesh.cpp:38-90): union-find over the
    elements of `subset` (ascending) that share a face (an edge in 2-D), keyed by
    the face's sorted corner nodes; components in order of first member."""
    nodes = element_nodes(model)
    faces = ([(0, 1), (1, 2), (2, 3), (3, 0)] if model.dim == 2 else
             [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)])
    parent = list(range(len(subset)))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    owner = {}
    for s, e in enumerate(subset):
        for f in faces:
            key = tuple(sorted(int(nodes[e, a]) for a in f))
            if key in owner:
                ra, rb = find(owner[key]), find(s)
                if ra != rb:
                    parent[rb] = ra
            else:
                owner[key] = s
    comps, index = [], {}
    for s, e in enumerate(subset):
        r = find(s)
        if r not in index:
            index[r] = len(comps)
            comps.append([])
        comps[index[r]].append(int(e))
    return comps


def submesh_node_order(model, elements):
    """extract_element_subset's node numbering (src/mesh/SubMesh.cpp:110-122):
    the parent node of each sub-mesh node, and the parent-to-sub map (-1: dropped)."""
    nodes = element_nodes(model)[np.asarray(elements)]
    old_to_new = -np.ones(model.nnodes, dtype=np.int64)
    order = []
    for n in nodes.ravel():
        if old_to_new[n] < 0:
            old_to_new[n] = len(order)
            order.append(n)
    return np.array(order), old_to_new


def submesh_operators(model, ke0, density, elements, fixed):
    """K_ff and M_ff of the mesh made of `elements` alone, nodes renumbered by
    first appearance (extract_element_subset, src/mesh/SubMesh.cpp:96-140), with
    the given fixed DOF mask (over the parent's DOFs) restricted to the kept nodes."""
    dim = model.dim
    nodes = element_nodes(model)[np.asarray(elements)]
    order, old_to_new = submesh_node_order(model, elements)
    conn = old_to_new[nodes]
    nn = len(order)
    dofs = (conn[:, :, None] * dim + np.arange(dim)[None, None, :]).reshape(len(elements), -1)
    rows = np.repeat(dofs, dofs.shape[1], axis=1).ravel()
    cols = np.tile(dofs, (1, dofs.shape[1])).ravel()
    me = element_mass(model, density)
    K = sp.csr_matrix((np.tile(np.asarray(ke0).ravel(), len(elements)), (rows, cols)), shape=(nn * dim, nn * dim))
    M = sp.csr_matrix((np.tile(me.ravel(), len(elements)), (rows, cols)), shape=(nn * dim, nn * dim))
    fx = np.asarray(fixed, dtype=bool).reshape(model.nnodes, dim)[order].ravel()
    free = np.flatnonzero(~fx)
    return reduce_free(K, free), reduce_free(M, free)
