"""Structured SIMP problems: mesh geometry, supports, loads and design domain.

This is the deck-level half of the host interface: it turns the settings a
SparLab JSON deck carries (``mesh``, ``boundary_conditions``, ``load_cases``,
``topology``) into the arrays the native library takes (a fixed-DOF mask, full
load vectors, design bounds). The selector semantics follow the reference:

* a box selector keeps nodes inside the box grown by 1e-9 times the mesh
  diagonal (src/fem/Selector.cpp:12-17,49-54);
* a nearest-node selector keeps the first node at the smallest squared
  distance (src/fem/Selector.cpp:117-130);
* a point load with ``distribution: total`` divides its force over the
  selected nodes (src/fem/BoundaryConditions.cpp:46-53);
* load-case weights are normalised to sum to one (src/fem/FemModel.cpp:412-420).

Node and element numbering follow src/mesh/StructuredMesh.hpp:6-27 (i fastest,
then j, then k).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class StructuredModel:
    dim: int
    nx: int
    ny: int
    nz: int = 0
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0
    E: float = 1.0
    nu: float = 0.3
    thickness: float = 1.0
    fixed: np.ndarray | None = None      # uint8 per DOF
    loads: np.ndarray | None = None      # (cases, ndofs) float64
    weights: np.ndarray | None = None    # (cases,) raw weights
    volume_fraction: float = 0.5
    lower: np.ndarray | None = None
    upper: np.ndarray | None = None
    passive: np.ndarray | None = None
    filter_radius: float = 0.0           # metres; 0 disables the filter
    penalty: float = 3.0
    emin_ratio: float = 1e-9

    def __post_init__(self):
        if self.dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        if self.dim == 2:
            self.nz = 0
        if self.fixed is None:
            self.fixed = np.zeros(self.ndofs, dtype=np.uint8)
        ne = self.nelem
        if self.lower is None:
            self.lower = np.zeros(ne)
        if self.upper is None:
            self.upper = np.ones(ne)
        if self.passive is None:
            self.passive = np.zeros(ne, dtype=np.uint8)

    # ---- sizes -----------------------------------------------------------
    @property
    def nnodes(self) -> int:
        if self.dim == 2:
            return (self.nx + 1) * (self.ny + 1)
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    @property
    def nelem(self) -> int:
        return self.nx * self.ny * (self.nz if self.dim == 3 else 1)

    @property
    def ndofs(self) -> int:
        return self.nnodes * self.dim

    @property
    def cell(self) -> tuple:
        if self.dim == 2:
            return (self.lx / self.nx, self.ly / self.ny)
        return (self.lx / self.nx, self.ly / self.ny, self.lz / self.nz)

    def mean_element_size(self) -> float:
        """Side of the square / cube of the mean cell measure (src/mesh/Mesh.cpp:398-406)."""
        c = self.cell
        return math.sqrt(c[0] * c[1]) if self.dim == 2 else float(np.cbrt(c[0] * c[1] * c[2]))

    def node_coords(self) -> np.ndarray:
        """(nnodes, dim) coordinates, i fastest (StructuredMesh.cpp:157-216)."""
        c = self.cell
        i = np.arange(self.nx + 1) * c[0]
        j = np.arange(self.ny + 1) * c[1]
        if self.dim == 2:
            X, Y = np.meshgrid(i, j, indexing="xy")
            return np.stack([X.ravel(), Y.ravel()], axis=1)
        k = np.arange(self.nz + 1) * c[2]
        Z, Y, X = np.meshgrid(k, j, i, indexing="ij")
        return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def element_centroids(self) -> np.ndarray:
        c = self.cell
        xs = (np.arange(self.nx) + 0.5) * c[0]
        ys = (np.arange(self.ny) + 0.5) * c[1]
        if self.dim == 2:
            X, Y = np.meshgrid(xs, ys, indexing="xy")
            return np.stack([X.ravel(), Y.ravel()], axis=1)
        zs = (np.arange(self.nz) + 0.5) * c[2]
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    # ---- selectors ---------------------------------------------------------
    def _tolerance(self) -> float:
        ext = [self.lx, self.ly] + ([self.lz] if self.dim == 3 else [])
        return 1e-9 * max(math.sqrt(sum(e * e for e in ext)), 1e-30)

    def select_box(self, xmin=-np.inf, xmax=np.inf, ymin=-np.inf, ymax=np.inf,
                   zmin=-np.inf, zmax=np.inf) -> np.ndarray:
        x = self.node_coords()
        t = self._tolerance()
        keep = (x[:, 0] >= xmin - t) & (x[:, 0] <= xmax + t) & (x[:, 1] >= ymin - t) & (x[:, 1] <= ymax + t)
        if self.dim == 3:
            keep &= (x[:, 2] >= zmin - t) & (x[:, 2] <= zmax + t)
        return np.nonzero(keep)[0]

    def select_circle(self, centre, radius) -> np.ndarray:
        x = self.node_coords()
        d = np.hypot(x[:, 0] - centre[0], x[:, 1] - centre[1])
        return np.nonzero(d <= radius + self._tolerance())[0]

    def select_annulus(self, centre, inner, outer) -> np.ndarray:
        x = self.node_coords()
        d = np.hypot(x[:, 0] - centre[0], x[:, 1] - centre[1])
        t = self._tolerance()
        return np.nonzero((d >= inner - t) & (d <= outer + t))[0]

    def select_nearest(self, point) -> np.ndarray:
        x = self.node_coords()
        p = np.zeros(self.dim)
        p[: len(point)] = np.asarray(point[: self.dim], dtype=float)
        d2 = ((x - p) ** 2).sum(axis=1)
        return np.array([int(np.argmin(d2))])  # argmin returns the first minimum

    def elements_in_circle(self, centre, radius) -> np.ndarray:
        c = self.element_centroids()
        d = np.hypot(c[:, 0] - centre[0], c[:, 1] - centre[1])
        return np.nonzero(d <= radius + self._tolerance())[0]

    # ---- supports and loads ------------------------------------------------
    def fix(self, nodes, comps=None):
        comps = range(self.dim) if comps is None else comps
        for k in comps:
            self.fixed[np.asarray(nodes) * self.dim + k] = 1

    def add_load_case(self, weight=1.0):
        row = np.zeros((1, self.ndofs))
        self.loads = row if self.loads is None else np.vstack([self.loads, row])
        w = np.array([weight], dtype=float)
        self.weights = w if self.weights is None else np.concatenate([self.weights, w])
        return self.loads.shape[0] - 1

    def add_point_load(self, case, nodes, force, total=True):
        nodes = np.asarray(nodes)
        scale = 1.0 / len(nodes) if total else 1.0
        for n in nodes:
            for k in range(self.dim):
                self.loads[case, n * self.dim + k] += scale * force[k]

    def set_passive(self, elements, solid: bool, void_density=0.0):
        v = 1.0 if solid else void_density
        self.lower[elements] = v
        self.upper[elements] = v
        self.passive[elements] = 1

    def initial_design(self, density=None) -> np.ndarray:
        x = np.full(self.nelem, self.volume_fraction if density is None else density, dtype=float)
        x[self.passive.astype(bool)] = self.lower[self.passive.astype(bool)]
        return x

    def free_dofs(self) -> np.ndarray:
        return np.nonzero(self.fixed == 0)[0].astype(np.int32)

    def free_loads(self) -> np.ndarray:
        return np.ascontiguousarray(self.loads[:, self.free_dofs()])


# ---------------------------------------------------------------------------
# Problems of the benchmark protocol and the workload configurations
# ---------------------------------------------------------------------------

def bench_box(nx: int, dim: int = 3) -> StructuredModel:
    """The sparlab_bench model (apps/sparlab_bench.cpp:87-127,241-243):
    nx x nx/2 (x nx/4) box, lx = 2, ly = 1, lz = 0.5, E = 70 GPa, nu = 0.3,
    x = 0 clamped, -1000 N along y at the node nearest (lx, 0, 0)."""
    ny = max(1, round(nx / 2))
    nz = max(1, ny // 2) if dim == 3 else 0
    m = StructuredModel(dim=dim, nx=nx, ny=ny, nz=nz, lx=2.0, ly=1.0, lz=0.5, E=70e9, nu=0.3,
                        thickness=1.0 if dim == 3 else 0.01)
    m.fix(m.select_box(xmax=0.0))
    c = m.add_load_case()
    m.add_point_load(c, m.select_nearest([2.0, 0.0, 0.0]), [0.0, -1000.0, 0.0], total=False)
    return m


def mbb_beam(nx=120, ny=40, radius_cells=1.5) -> StructuredModel:
    """MBB half-beam (configs/benchmarks/mbb_beam.json at 120 x 40 unit cells):
    u_x = 0 on x = 0, roller u_y = 0 at (nx, 0), unit load (0, -1) at (0, ny)."""
    m = StructuredModel(dim=2, nx=nx, ny=ny, lx=float(nx), ly=float(ny), E=1.0, nu=0.3, thickness=1.0,
                        volume_fraction=0.5)
    m.fix(m.select_box(xmax=0.0), comps=[0])
    m.fix(m.select_nearest([float(nx), 0.0]), comps=[1])
    c = m.add_load_case()
    m.add_point_load(c, m.select_nearest([0.0, float(ny)]), [0.0, -1.0])
    m.filter_radius = radius_cells * m.mean_element_size()
    return m


def cantilever_3d(nx, ny, nz, volume_fraction=0.3, radius_cells=2.0) -> StructuredModel:
    """Hex8 cantilever of unit cells: x = 0 face clamped, unit total load (0, -1, 0)
    over the free-end bottom edge (box xmin = nx, ymax = 0), as the tip_bottom_edge
    region of configs/benchmarks/bracket_3d_large.json."""
    m = StructuredModel(dim=3, nx=nx, ny=ny, nz=nz, lx=float(nx), ly=float(ny), lz=float(nz), E=1.0, nu=0.3,
                        volume_fraction=volume_fraction)
    m.fix(m.select_box(xmax=0.0))
    c = m.add_load_case()
    m.add_point_load(c, m.select_box(xmin=float(nx), ymax=0.0), [0.0, -1.0, 0.0])
    m.filter_radius = radius_cells * m.mean_element_size()
    return m


def seeded_density(nelem: int, seed: int = 20261203, low=0.001, high=1.0) -> np.ndarray:
    """Densities uniform in [low, high) from a fixed seed (the build's choice)."""
    return np.random.default_rng(seed).uniform(low, high, nelem)
