"""The optimised density field read as a solid body, and its free vibration.

interpret_density_as_solid follows src/mesh/SubMesh.cpp:172-260: elements
with a physical density at or above the threshold are kept, the largest
face-connected group of them (edge-connected in 2-D; components_by_face,
SubMesh.cpp:38-90) is retained, ties going to the group whose lowest element
index comes first, and the volumes above the threshold, retained and
discarded as islands are reported.

The reference then builds a new mesh of the retained elements, renumbers its
nodes in order of first appearance (SubMesh.cpp:98-130) and re-applies the
deck's supports to it (apps/sparlab_topopt.cpp:92-130). The supports are
re-applied here the same way: each boundary condition's region is evaluated on
the retained nodes in that numbering, with the tolerance of their bounding box,
so a nearest_node support picks the nearest retained node and node_ids name
sub-mesh nodes. The operators then come out of the parent grid directly:
per-element stiffness and mass factors of 1 on retained elements and 0
elsewhere, and every DOF of a node that no retained element touches fixed.
K_ff and M_ff are the sub-model's matrices with the same unknowns in the same
relative order, so the eigenpairs are the same.
"""
from __future__ import annotations

import dataclasses

import numpy as np
from scipy import ndimage

from .model import StructuredModel


class InterpretationError(RuntimeError):
    """No element reaches the threshold (the reference's MeshError)."""


@dataclasses.dataclass
class Interpretation:
    threshold: float
    elements_above_threshold: int
    elements_retained: int
    components_above_threshold: int
    volume_above_threshold: float
    volume_retained: float
    volume_discarded_as_islands: float
    retained: np.ndarray  # bool per element

    def as_dict(self) -> dict:
        d = dataclasses.asdict(self)
        d.pop("retained")
        return d


def element_volume(model: StructuredModel) -> float:
    return float(np.prod(model.cell)) * (model.thickness if model.dim == 2 else 1.0)


def _grid_shape(model: StructuredModel) -> tuple:
    return (model.ny, model.nx) if model.dim == 2 else (model.nz, model.ny, model.nx)


def interpret_density_as_solid(model: StructuredModel, density, threshold: float = 0.5,
                               largest_component_only: bool = True) -> Interpretation:
    density = np.asarray(density, dtype=float)
    if density.shape != (model.nelem,):
        raise ValueError(f"density vector has length {density.size} but the mesh has {model.nelem} elements")
    if not (0.0 < threshold < 1.0):
        raise ValueError(f"density threshold must lie strictly between 0 and 1 (got {threshold})")
    kept = density >= threshold
    nk = int(kept.sum())
    if nk == 0:
        raise InterpretationError(f"no element reaches the density threshold {threshold} (max density is "
                                  f"{density.max()}); lower the threshold or increase the volume fraction")
    # face neighbours only (the default cross-shaped structure); labels follow
    # the first element of each group in index order
    labels, ncomp = ndimage.label(kept.reshape(_grid_shape(model)))
    labels = labels.ravel()
    ve = element_volume(model)
    retained = kept
    if largest_component_only and ncomp > 1:
        counts = np.bincount(labels, minlength=ncomp + 1)[1:]
        best = int(np.argmax(counts)) + 1  # first of the largest (uniform element volumes)
        retained = labels == best
    nr = int(retained.sum())
    return Interpretation(threshold=threshold, elements_above_threshold=nk, elements_retained=nr,
                          components_above_threshold=int(ncomp), volume_above_threshold=nk * ve,
                          volume_retained=nr * ve, volume_discarded_as_islands=(nk - nr) * ve, retained=retained)


def element_nodes(model: StructuredModel) -> np.ndarray:
    """(nelem, npe) node indices in the local order of src/mesh/StructuredMesh.hpp."""
    nnx, nny = model.nx + 1, model.ny + 1
    e = np.arange(model.nelem)
    ex = e % model.nx
    if model.dim == 2:
        n0 = (e // model.nx) * nnx + ex
        return np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx], axis=1)
    ey = (e // model.nx) % model.ny
    ez = e // (model.nx * model.ny)
    lay = nnx * nny
    n0 = (ez * nny + ey) * nnx + ex
    return np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx, n0 + lay, n0 + lay + 1, n0 + lay + nnx + 1, n0 + lay + nnx],
                    axis=1)


def submesh_nodes(model: StructuredModel, retained) -> np.ndarray:
    """The parent node of each sub-mesh node, in the order extract_element_subset
    numbers them (SubMesh.cpp:110-122): retained elements ascending, each
    element's nodes in its local order, a node numbered where it first appears."""
    conn = element_nodes(model)[np.asarray(retained, dtype=bool)].ravel()
    _, first = np.unique(conn, return_index=True)
    return conn[np.sort(first)]


def solid_submodel(model: StructuredModel, retained, constraints=None,
                   origin=(0.0, 0.0, 0.0)) -> tuple[StructuredModel, np.ndarray]:
    """The parent grid with every DOF of the nodes outside the retained elements
    fixed; returns (model, retained-node mask). With ``constraints`` (a deck's
    parsed boundary conditions) the supports are re-applied on the sub-mesh as
    the reference does; without, the parent's fixed DOFs on retained nodes are
    kept. A condition that selects no retained node raises DeckError."""
    from . import deck as D

    retained = np.asarray(retained, dtype=bool)
    used = np.zeros(model.nnodes, dtype=bool)
    used[element_nodes(model)[retained].ravel()] = True
    if constraints is None:
        fixed = np.asarray(model.fixed, dtype=np.uint8).reshape(model.nnodes, model.dim).copy()
    else:
        node_map = submesh_nodes(model, retained)
        coords = model.node_coords()[node_map] + np.asarray(origin, dtype=float)[: model.dim]
        sub_fixed = D.apply_constraints(D.Geometry.of_nodes(model.dim, coords), constraints)
        fixed = np.zeros((model.nnodes, model.dim), dtype=np.uint8)
        fixed[node_map] = sub_fixed
    fixed[~used] = 1
    sub = dataclasses.replace(model, fixed=fixed.ravel(), loads=None, weights=None)
    return sub, used


def supported(model: StructuredModel, used) -> bool:
    """Whether any support of the parent acts on a retained node."""
    parent = np.asarray(model.fixed, dtype=bool).reshape(model.nnodes, model.dim)
    return bool(parent[np.asarray(used, dtype=bool)].any())
