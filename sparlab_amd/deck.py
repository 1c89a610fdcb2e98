"""SparLab JSON decks for the structured SIMP path.

Reads the decks ``sparlab_topopt`` takes (src/io/Config.cpp:538-1730) and turns
the part this library runs -- a structured Q4 / Hex8 mesh, supports, point
loads, tractions and pressures on the boundary, load-case weights, the SIMP
law, the density filter, passive regions and the optimality-criteria loop with
its penalty continuation -- into a
:class:`~sparlab_amd.model.StructuredModel` and the ``run_oc`` options. Settings
outside that path (other meshes, body loads, robust projection, MMA,
stress or buckling constraints, ...) raise :class:`DeckError` naming the key,
rather than being ignored.

The deck syntax follows the reference's JSON reader (src/io/Json.cpp:50-80):
standard JSON plus ``//`` line comments and ``/* */`` block comments.

Region semantics follow src/fem/Selector.cpp:
* tolerance: the selector's own ``tolerance`` if positive, else 1e-9 times the
  mesh bounding-box diagonal (Selector.cpp:12-18,48);
* node regions test node coordinates, element regions element centroids
  (Selector.cpp:139-141,189-191);
* circle / annulus distances are measured in the plane normal to ``axis``
  (default z; Selector.cpp:38-43, Config.cpp:32-39);
* ``nearest_node`` keeps the first node at the smallest squared distance
  (Selector.cpp:117-130);
* ``any_of`` ORs its members, ``invert`` complements the union
  (Selector.cpp:143-148, Config.cpp:170-191).
"""
from __future__ import annotations

import dataclasses
import json
import math
import os

import numpy as np

from .model import StructuredModel


class DeckError(ValueError):
    """A deck that is malformed, or asks for something outside this path."""


# ---------------------------------------------------------------------------
# JSON with comments
# ---------------------------------------------------------------------------
def strip_comments(text: str) -> str:
    """Remove // and /* */ comments outside string literals (src/io/Json.cpp:50-80)."""
    out = []
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif c == "/" and i + 1 < n and text[i + 1] == "/":
            while i < n and text[i] != "\n":
                i += 1
        elif c == "/" and i + 1 < n and text[i + 1] == "*":
            end = text.find("*/", i + 2)
            if end < 0:
                raise DeckError("unterminated block comment")
            # keep the line count so that JSON errors point at the right line
            out.append("\n" * text.count("\n", i, end))
            i = end + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


def read_deck(path: str) -> dict:
    with open(path, "r", encoding="utf-8") as f:
        text = f.read()
    try:
        doc = json.loads(strip_comments(text))
    except json.JSONDecodeError as e:
        raise DeckError(f"{path}: {e}") from None
    if not isinstance(doc, dict):
        raise DeckError(f"{path}: the top-level configuration must be a JSON object")
    return doc


# ---------------------------------------------------------------------------
# Regions
# ---------------------------------------------------------------------------
_PRIMITIVES = ("all", "box", "circle", "annulus", "sphere", "node_ids", "element_ids", "nearest_node", "group")
_AXES = {"x": 0, "y": 1, "z": 2}


def _vec3(v, dim, what):
    if not isinstance(v, (list, tuple)) or len(v) != dim or not all(isinstance(a, (int, float)) for a in v):
        raise DeckError(f"'{what}' must be a list of {dim} numbers")
    p = np.zeros(3)
    p[:dim] = v
    return p


def _primitive(node: dict, dim: int, where: str) -> dict:
    if not isinstance(node, dict):
        raise DeckError(f"'{where}' must be an object")
    names = [k for k in _PRIMITIVES if k in node]
    if not names:
        raise DeckError(f"'{where}' does not name a region primitive; expected one of "
                        + ", ".join(_PRIMITIVES) + " (or an 'any_of' list)")
    if len(names) > 1:
        raise DeckError(f"'{where}' names more than one region primitive; use 'any_of' to combine several regions")
    kind = names[0]
    sel = {"kind": kind, "tolerance": float(node.get("tolerance", 0.0))}
    body = node[kind]
    inf = math.inf
    if kind == "box":
        b = body
        if dim == 2 and ("zmin" in b or "zmax" in b):
            raise DeckError(f"'{where}.box' sets zmin/zmax but the mesh is two-dimensional")
        sel["lo"] = np.array([b.get("xmin", -inf), b.get("ymin", -inf), b.get("zmin", -inf)], dtype=float)
        sel["hi"] = np.array([b.get("xmax", inf), b.get("ymax", inf), b.get("zmax", inf)], dtype=float)
        if np.any(sel["lo"] > sel["hi"]):
            raise DeckError(f"'{where}.box' has an empty interval")
    elif kind in ("circle", "annulus", "sphere"):
        sel["center"] = _vec3(body.get("center"), dim, f"{where}.{kind}.center")
        sel["radius"] = float(body.get("radius", 0.0))
        if not sel["radius"] > 0.0:
            raise DeckError(f"'{where}.{kind}.radius' must be positive")
        if kind != "sphere":
            axis = body.get("axis", "z")
            if axis not in _AXES:
                raise DeckError(f"'{where}.{kind}.axis' must be \"x\", \"y\" or \"z\", got \"{axis}\"")
            sel["axis"] = _AXES[axis]
            if dim == 2 and sel["axis"] != 2:
                raise DeckError(f"'{where}.{kind}.axis' must be \"z\" on a two-dimensional mesh")
        if kind == "annulus":
            sel["inner"] = float(body.get("inner_radius", 0.0))
            if sel["inner"] >= sel["radius"]:
                raise DeckError(f"'{where}.annulus' needs inner_radius strictly below radius")
    elif kind in ("node_ids", "element_ids"):
        ids = np.asarray(body, dtype=np.int64)
        if ids.size == 0:
            raise DeckError(f"'{where}.{kind}' is empty")
        sel["ids"] = ids
    elif kind == "nearest_node":
        sel["point"] = _vec3(body, dim, f"{where}.nearest_node")
    elif kind == "group":
        raise DeckError(f"'{where}.group' names a mesh group; structured meshes carry none")
    return sel


def parse_region(node: dict, dim: int, where: str) -> dict:
    """Config.cpp:170-191: a list of primitives (``any_of`` or one) and ``invert``."""
    if not isinstance(node, dict):
        raise DeckError(f"required region '{where}' is missing")
    members = node.get("any_of")
    if members:
        prims = [_primitive(m, dim, f"{where}.any_of[{i}]") for i, m in enumerate(members)]
    else:
        prims = [_primitive(node, dim, where)]
    return {"name": node.get("name", ""), "invert": bool(node.get("invert", False)), "members": prims}


def _contains(sel: dict, x: np.ndarray, tol: float) -> np.ndarray:
    """Selector::contains (Selector.cpp:46-71) over points x[:, 3]."""
    t = sel["tolerance"] if sel["tolerance"] > 0.0 else tol
    kind = sel["kind"]
    if kind == "all":
        return np.ones(len(x), dtype=bool)
    if kind == "box":
        return np.all((x >= sel["lo"] - t) & (x <= sel["hi"] + t), axis=1)
    d = x - sel["center"]
    if kind == "sphere":
        return np.sqrt((d * d).sum(axis=1)) <= sel["radius"] + t
    d[:, sel["axis"]] = 0.0
    r = np.sqrt((d * d).sum(axis=1))
    if kind == "circle":
        return r <= sel["radius"] + t
    return (r >= sel["inner"] - t) & (r <= sel["radius"] + t)


class Geometry:
    """Node coordinates and element centroids of a structured mesh with its origin."""

    def __init__(self, model: StructuredModel, origin=(0.0, 0.0, 0.0)):
        self.dim = model.dim
        o = np.asarray(origin, dtype=float)
        nodes = model.node_coords()
        cents = model.element_centroids()
        self.nodes = np.zeros((len(nodes), 3))
        self.nodes[:, : self.dim] = nodes + o[: self.dim]
        self.cents = np.zeros((len(cents), 3))
        self.cents[:, : self.dim] = cents + o[: self.dim]
        ext = [model.lx, model.ly] + ([model.lz] if model.dim == 3 else [])
        self.tol = 1.0e-9 * max(math.sqrt(sum(e * e for e in ext)), 1.0e-30)

    @classmethod
    def of_nodes(cls, dim: int, nodes: np.ndarray) -> "Geometry":
        """An unstructured node set (the nodes of an extracted sub-mesh, in its
        numbering), its tolerance from its own bounding box (Selector.cpp:12-18)."""
        g = cls.__new__(cls)
        g.dim = dim
        g.nodes = np.zeros((len(nodes), 3))
        g.nodes[:, :dim] = np.asarray(nodes, dtype=float)[:, :dim]
        g.cents = np.zeros((0, 3))
        ext = g.nodes.max(axis=0) - g.nodes.min(axis=0) if len(nodes) else np.zeros(3)
        g.tol = 1.0e-9 * max(float(np.sqrt((ext[:dim] ** 2).sum())), 1.0e-30)
        return g

    def nodes_of(self, region: dict, where: str) -> np.ndarray:
        """SelectorGroup::select_nodes (Selector.cpp:73-148)."""
        hit = np.zeros(len(self.nodes), dtype=bool)
        for sel in region["members"]:
            k = sel["kind"]
            if k == "element_ids":
                raise DeckError(f"region '{where}' uses element_ids where a node region is expected")
            if k == "node_ids":
                if sel["ids"].min() < 0 or sel["ids"].max() >= len(self.nodes):
                    raise DeckError(f"region '{where}' lists a node id outside [0, {len(self.nodes) - 1}]")
                hit[sel["ids"]] = True
            elif k == "nearest_node":
                d2 = ((self.nodes - sel["point"]) ** 2).sum(axis=1)
                hit[int(np.argmin(d2))] = True  # the first minimum
            else:
                hit |= _contains(sel, self.nodes, self.tol)
        return np.nonzero(hit != region["invert"])[0]

    def elements_of(self, region: dict, where: str) -> np.ndarray:
        """SelectorGroup::select_elements (Selector.cpp:150-197): centroids."""
        hit = np.zeros(len(self.cents), dtype=bool)
        for sel in region["members"]:
            k = sel["kind"]
            if k in ("node_ids", "nearest_node"):
                raise DeckError(f"region '{where}' uses a node-only selector where an element region is expected")
            if k == "element_ids":
                if sel["ids"].min() < 0 or sel["ids"].max() >= len(self.cents):
                    raise DeckError(f"region '{where}' lists an element id outside [0, {len(self.cents) - 1}]")
                hit[sel["ids"]] = True
            else:
                hit |= _contains(sel, self.cents, self.tol)
        return np.nonzero(hit != region["invert"])[0]


def apply_constraints(geo: Geometry, constraints: list) -> np.ndarray:
    """apply_constraints (BoundaryConditions.cpp:93-130): the fixed components,
    [nodes, dim] bool, of the parsed boundary conditions on a node set; a
    condition that selects no node is an error, as there."""
    fixed = np.zeros((len(geo.nodes), geo.dim), dtype=bool)
    for bc in constraints:
        nodes = geo.nodes_of(bc["region"], bc["name"])
        if len(nodes) == 0:
            raise DeckError(f"boundary condition '{bc['name']}' selected no nodes; check its coordinates against "
                            "the mesh extents (an unconstrained model yields a singular stiffness matrix)")
        for c in bc["comps"]:
            fixed[nodes, c] = True
    return fixed


def boundary_faces(m: StructuredModel):
    """The faces on the boundary of a structured grid: (nodes[F, k], outward
    normals[F, 3], measures[F]), k = 2 (edges) or 4 (faces). A face's measure
    includes the thickness in 2-D (the area the traction acts on)."""
    nx, ny = m.nx, m.ny
    if m.dim == 2:
        hx, hy = m.lx / nx, m.ly / ny
        node = lambda i, j: j * (nx + 1) + i  # noqa: E731
        faces, normals, meas = [], [], []
        for i in range(nx):
            faces += [(node(i, 0), node(i + 1, 0)), (node(i, ny), node(i + 1, ny))]
            normals += [(0.0, -1.0, 0.0), (0.0, 1.0, 0.0)]
            meas += [hx * m.thickness] * 2
        for j in range(ny):
            faces += [(node(0, j), node(0, j + 1)), (node(nx, j), node(nx, j + 1))]
            normals += [(-1.0, 0.0, 0.0), (1.0, 0.0, 0.0)]
            meas += [hy * m.thickness] * 2
        return np.array(faces, dtype=np.int64), np.array(normals), np.array(meas)
    nz = m.nz
    hx, hy, hz = m.lx / nx, m.ly / ny, m.lz / nz
    node = lambda i, j, k: (k * (ny + 1) + j) * (nx + 1) + i  # noqa: E731
    faces, normals, meas = [], [], []
    for a in range(nx):
        for b in range(ny):
            for k, sgn in ((0, -1.0), (nz, 1.0)):
                faces.append((node(a, b, k), node(a + 1, b, k), node(a + 1, b + 1, k), node(a, b + 1, k)))
                normals.append((0.0, 0.0, sgn))
                meas.append(hx * hy)
    for a in range(nx):
        for c in range(nz):
            for j, sgn in ((0, -1.0), (ny, 1.0)):
                faces.append((node(a, j, c), node(a + 1, j, c), node(a + 1, j, c + 1), node(a, j, c + 1)))
                normals.append((0.0, sgn, 0.0))
                meas.append(hx * hz)
    for b in range(ny):
        for c in range(nz):
            for i, sgn in ((0, -1.0), (nx, 1.0)):
                faces.append((node(i, b, c), node(i, b + 1, c), node(i, b + 1, c + 1), node(i, b, c + 1)))
                normals.append((sgn, 0.0, 0.0))
                meas.append(hy * hz)
    return np.array(faces, dtype=np.int64), np.array(normals), np.array(meas)


def add_face_load(m: StructuredModel, case: int, nodes_in: np.ndarray, traction=None, pressure=None) -> int:
    """Consistent nodal loads of a uniform traction [Pa] or pressure [Pa] on the
    boundary faces whose nodes all lie in the region (BoundaryConditions.cpp:63-77,
    159-232): on a straight edge or a flat rectangular face the linear / bilinear
    shape functions give each node measure / (nodes per face); a pressure acts
    against the outward normal (FaceGeometry.cpp:174-191). Returns the faces matched."""
    faces, normals, meas = boundary_faces(m)
    mask = np.zeros(m.nnodes, dtype=bool)
    mask[nodes_in] = True
    sel = np.all(mask[faces], axis=1)
    k = faces.shape[1]
    for f, nrm, a in zip(faces[sel], normals[sel], meas[sel]):
        force = np.asarray(traction, dtype=float)[: m.dim] * a if traction is not None else -pressure * a * nrm[: m.dim]
        for n in f:
            m.loads[case, n * m.dim: (n + 1) * m.dim] += force / k
    return int(sel.sum())


# ---------------------------------------------------------------------------
# The deck
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Deck:
    name: str
    description: str
    source: str
    model: StructuredModel
    origin: tuple
    case_names: list
    case_weights: list          # as written in the deck
    tolerance: float            # relative residual of the linear solves
    filter_type: str
    filter_radius: float        # metres, 0 when the filter is off
    initial_density: float
    passive_solid: int
    passive_void: int
    run: dict                   # keyword arguments of HipProblem.run_oc
    history_stride: int = 1
    density: float = 0.0        # kg/m^3 (material.density)
    modal: dict | None = None   # keyword arguments of HipProblem.modal when modal.enabled, else None
    modal_flags: dict | None = None  # analyse_optimised_topology, compare_mass_matched_baseline
    interpretation_threshold: float = 0.5  # optimizer.interpretation_threshold (Config.cpp:1571)
    skipped_reports: list = dataclasses.field(default_factory=list)  # report-only settings not computed
    amg: dict = dataclasses.field(default_factory=dict)  # solver.linear.amg keys set by the deck (library names)
    constraints: list = dataclasses.field(default_factory=list)  # parsed boundary conditions (region, comps, name)

    @property
    def domain_volume(self) -> float:
        t = self.model.thickness if self.model.dim == 2 else 1.0
        return self.model.lx * self.model.ly * (self.model.lz if self.model.dim == 3 else t)

    def initial_design(self) -> np.ndarray:
        """DesignDomain's initial vector (DesignDomain.cpp:29-67): the start density, passive values."""
        x = np.full(self.model.nelem, self.initial_density)
        p = self.model.passive.astype(bool)
        x[p] = self.model.lower[p]
        return x


def _unsupported(path: str, what: str):
    raise DeckError(f"'{path}' asks for {what}, which this structured SIMP path does not run")


def _number(d: dict, key: str, default, where: str):
    v = d.get(key, default)
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise DeckError(f"'{where}.{key}' must be a number")
    return v


_SMOOTHERS = {"chebyshev": 0, "gauss_seidel": 1, "symmetric_gauss_seidel": 1}


def _amg_options(lin: dict) -> dict:
    """solver.linear.amg (Config.cpp:1097-1110) as keyword arguments of the
    library's multigrid options, only the keys the deck sets; checked as
    AmgPreconditioner's constructor checks them (Multigrid.cpp:899-925)."""
    a = lin.get("amg", {})
    if not isinstance(a, dict):
        raise DeckError("'solver.linear.amg' must be an object")
    where = "solver.linear.amg"
    out = {}
    for key in ("strength_threshold", "chebyshev_ratio", "prolongator_damping", "coarse_pivot_tolerance"):
        if key in a:
            out[key] = float(_number(a, key, 0.0, where))
    for key in ("max_levels", "coarse_size", "smoother_degree", "lanczos_steps"):
        if key in a:
            v = _number(a, key, 0, where)
            if v != int(v):
                raise DeckError(f"'{where}.{key}' must be an integer")
            out[key] = int(v)
    if "smoother" in a:
        text = str(a["smoother"])
        if text not in _SMOOTHERS:
            raise DeckError(f"unknown multigrid smoother '{text}' (expected chebyshev|gauss_seidel)")
        out["smoother"] = _SMOOTHERS[text]
    if "reuse_aggregates" in a:
        if not isinstance(a["reuse_aggregates"], bool):
            raise DeckError(f"'{where}.reuse_aggregates' must be true or false")
        out["reuse_aggregates"] = int(a["reuse_aggregates"])
    if "strength_threshold" in out and not 0.0 <= out["strength_threshold"] < 1.0:
        raise DeckError("multigrid strength_threshold must lie in [0, 1)")
    if out.get("max_levels", 1) < 1:
        raise DeckError("multigrid max_levels must be at least 1")
    if out.get("coarse_size", 1) < 1:
        raise DeckError("multigrid coarse_size must be positive")
    if out.get("smoother_degree", 1) < 1:
        raise DeckError("multigrid smoother_degree must be at least 1")
    if not out.get("chebyshev_ratio", 2.0) > 1.0:
        raise DeckError("multigrid chebyshev_ratio must exceed 1")
    if not 0.0 < out.get("prolongator_damping", 1.0) < 2.0:
        raise DeckError("multigrid prolongator_damping must lie in (0, 2)")
    if out.get("lanczos_steps", 2) < 2:
        raise DeckError("multigrid lanczos_steps must be at least 2")
    if out.get("max_levels", 1) > 16:
        _unsupported(f"{where}.max_levels", "more than 16 multigrid levels")
    return out


def parse_deck(doc: dict, source: str = "", overrides: dict | None = None) -> Deck:
    """Configuration for sparlab_topopt (src/io/Config.cpp:538-1730, apps/sparlab_topopt.cpp:195-262).

    ``overrides`` takes the command-line overrides of sparlab_topopt:
    volume_fraction, penalty, filter_radius, filter_radius_elements, filter_type,
    max_iterations, nx, ny, nz, youngs_modulus, load_weights.
    """
    ov = dict(overrides or {})
    name = doc.get("name", "case")
    mesh = doc.get("mesh")
    if not isinstance(mesh, dict):
        raise DeckError("'mesh' is missing")
    mtype = mesh.get("type", "structured_quad")
    if mtype not in ("structured_quad", "structured_hex"):
        _unsupported("mesh.type", f"a {mtype} mesh")
    dim = 2 if mtype == "structured_quad" else 3
    nx = int(ov.get("nx") or mesh.get("nx", 0))
    ny = int(ov.get("ny") or mesh.get("ny", 0))
    nz = int(ov.get("nz") or mesh.get("nz", 0)) if dim == 3 else 0
    if dim == 2 and ov.get("nz"):
        raise DeckError("--nz applies to a solid (structured_hex) mesh only")
    if nx <= 0 or ny <= 0 or (dim == 3 and nz <= 0):
        raise DeckError("'mesh' needs positive nx, ny" + (", nz" if dim == 3 else ""))
    lx = float(mesh.get("lx", 0.0))
    ly = float(mesh.get("ly", 0.0))
    lz = float(mesh.get("lz", 0.0)) if dim == 3 else 1.0
    if not (lx > 0 and ly > 0 and lz > 0):
        raise DeckError("'mesh' needs positive lx, ly" + (", lz" if dim == 3 else ""))
    origin = (float(mesh.get("x0", 0.0)), float(mesh.get("y0", 0.0)), float(mesh.get("z0", 0.0)))

    mat = doc.get("material")
    if not isinstance(mat, dict):
        raise DeckError("'material' is missing")
    if "material_regions" in doc:
        _unsupported("material_regions", "a second material")
    E = float(ov.get("youngs_modulus") or _number(mat, "youngs_modulus", 0.0, "material"))
    nu = float(_number(mat, "poisson_ratio", 0.3, "material"))
    if not E > 0:
        raise DeckError("'material.youngs_modulus' must be positive")

    model_node = doc.get("model", {})
    thickness = float(model_node.get("thickness", 1.0))
    if dim == 3 and thickness != 1.0:
        raise DeckError("'model.thickness' does not apply to a solid mesh; remove the key")
    state = model_node.get("stress_state", "plane_stress" if dim == 2 else "three_dimensional")
    if dim == 2 and state != "plane_stress":
        _unsupported("model.stress_state", f"the stress state {state}")
    if dim == 3 and state not in ("three_dimensional", "3d"):
        raise DeckError(f"'model.stress_state' = \"{state}\" does not fit a structured_hex mesh")
    points = model_node.get("integration", {}).get("stiffness_points", 2)
    if points != 2:
        _unsupported("model.integration.stiffness_points", f"{points}-point stiffness integration")

    topo = doc.get("topology", {})
    if not topo.get("enabled", False):
        raise DeckError("'topology.enabled' is not true; this driver runs topology optimisation decks")
    vf = float(ov.get("volume_fraction") or topo.get("volume_fraction", 0.4))
    if not (0.0 < vf <= 1.0):
        raise DeckError(f"volume fraction must lie in (0, 1] (got {vf})")
    simp = topo.get("simp", {})
    penalty = float(ov.get("penalty") or simp.get("penalty", 3.0))
    emin = float(simp.get("emin_ratio", 1e-9))

    m = StructuredModel(dim=dim, nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, E=E, nu=nu, thickness=thickness,
                        volume_fraction=vf, penalty=penalty, emin_ratio=emin)
    geo = Geometry(m, origin)

    # supports (Config.cpp:879-941)
    bcs = doc.get("boundary_conditions") or []
    if not bcs:
        raise DeckError("'boundary_conditions' is missing or empty; an unconstrained model has a singular "
                        "stiffness matrix")
    constraints = []
    for i, bc in enumerate(bcs):
        where = f"boundary_conditions[{i}]"
        fix = bc.get("fix") or []
        if not fix:
            raise DeckError(f"'{where}.fix' must list the components to constrain")
        comps = []
        for c in fix:
            if c == "z" and dim == 2:
                raise DeckError(f"'{where}.fix' fixes \"z\" but the mesh is two-dimensional")
            if c not in ("x", "y", "z"):
                raise DeckError(f"'{where}.fix' must hold \"x\", \"y\"" + (", \"z\"" if dim == 3 else "")
                                + f", got \"{c}\"")
            comps.append(_AXES[c])
        value = bc.get("value")
        if value is not None and any(float(v) != 0.0 for v in value):
            _unsupported(f"{where}.value", "a prescribed non-zero displacement")
        spec = dict(region=parse_region(bc.get("region"), dim, f"{where}.region"), comps=comps,
                    name=bc.get("name", f"bc{i}"))
        constraints.append(spec)
    m.fixed = apply_constraints(geo, constraints).astype(np.uint8).ravel()

    # load cases (Config.cpp:943-1070)
    cases = doc.get("load_cases") or []
    if not cases:
        raise DeckError("'load_cases' is missing or empty; define at least one")
    names, weights = [], []
    w_over = ov.get("load_weights")
    if w_over is not None and len(w_over) != len(cases):
        raise DeckError(f"--load-weights supplied {len(w_over)} weight(s) but the deck defines {len(cases)} "
                        "load case(s)")
    for li, lc in enumerate(cases):
        where = f"load_cases[{li}]"
        for key in ("body_forces", "centrifugal", "temperature", "springs", "prescribed_displacements"):
            if lc.get(key):
                _unsupported(f"{where}.{key}", key.replace("_", " "))
        g = lc.get("gravity")
        if g is not None and any(float(v) != 0.0 for v in g):
            _unsupported(f"{where}.gravity", "self-weight")
        cname = lc.get("name", f"case{li}")
        w = float(w_over[li]) if w_over is not None else float(lc.get("weight", 1.0))
        if not w >= 0.0:
            raise DeckError(f"'{where}.weight' must be non-negative, got {w}")
        case = m.add_load_case(w)
        for pi, pl in enumerate(lc.get("point_loads") or []):
            pw = f"{where}.point_loads[{pi}]"
            if "moment" in pl:
                _unsupported(f"{pw}.moment", "a point moment")
            force = _vec3(pl.get("force"), dim, f"{pw}.force")[:dim]
            dist = pl.get("distribution", "total")
            if dist not in ("total", "per_node"):
                raise DeckError(f"'{pw}.distribution' must be 'total' or 'per_node', got '{dist}'")
            lname = pl.get("name", f"{cname}_point{pi}")
            nodes = geo.nodes_of(parse_region(pl.get("region"), dim, f"{pw}.region"), lname)
            if len(nodes) == 0:
                raise DeckError(f"point load '{lname}' selected no nodes")
            m.add_point_load(case, nodes, force, total=(dist == "total"))
        for ti, tr in enumerate(lc.get("tractions") or []):
            tw = f"{where}.tractions[{ti}]"
            t = _vec3(tr.get("traction"), dim, f"{tw}.traction")
            tname = tr.get("name", f"{cname}_traction{ti}")
            nodes = geo.nodes_of(parse_region(tr.get("region"), dim, f"{tw}.region"), tname)
            if add_face_load(m, case, nodes, traction=t) == 0:
                raise DeckError(f"traction region '{tname}' in load case '{cname}' matched no boundary "
                                + ("edge" if dim == 2 else "face"))
        for pi, pr in enumerate(lc.get("pressures") or []):
            pw = f"{where}.pressures[{pi}]"
            pval = pr.get("pressure")
            if isinstance(pval, bool) or not isinstance(pval, (int, float)):
                raise DeckError(f"'{pw}.pressure' must be a number")
            pname = pr.get("name", f"{cname}_pressure{pi}")
            nodes = geo.nodes_of(parse_region(pr.get("region"), dim, f"{pw}.region"), pname)
            if add_face_load(m, case, nodes, pressure=float(pval)) == 0:
                raise DeckError(f"pressure region '{pname}' in load case '{cname}' matched no boundary "
                                + ("edge" if dim == 2 else "face"))
        names.append(cname)
        weights.append(w)
    if sum(weights) <= 0.0:
        raise DeckError("the load-case weights sum to zero")

    # the optimisation problem (Config.cpp:1542-1730)
    for key, what in (("stress", "a stress constraint"),
                      ("overhang", "an overhang filter"), ("buckling_constraint", "a buckling constraint"),
                      ("length_scale", "a length-scale filter")):
        node = topo.get(key)
        if isinstance(node, dict) and node.get("enabled", False):
            _unsupported(f"topology.{key}", what)
    opt = topo.get("optimizer", {})
    method = opt.get("method", "oc")
    if method != "oc":
        _unsupported("topology.optimizer.method", f"the {method} update")
    # Heaviside projection with beta continuation (Config.cpp:1594-1617, Projection.cpp:10-56)
    pnode = dict(topo.get("projection", {}) or {})
    for key in ("projection", "robust"):  # the CLI's switches (apps/sparlab_topopt.cpp:300-318)
        if key in ov:
            pnode["enabled" if key == "projection" else "robust"] = bool(ov[key])
    if ov.get("robust"):
        pnode["enabled"] = True
    if "beta_max" in ov:
        pnode["beta_max"] = float(ov["beta_max"])
    projection = None
    if pnode.get("enabled", False):
        robust = bool(pnode.get("robust", False))
        projection = dict(eta=float(pnode.get("eta", 0.5)), beta_start=float(pnode.get("beta_start", 1.0)),
                          beta_max=float(pnode.get("beta_max", 32.0)), beta_factor=float(pnode.get("beta_factor", 2.0)),
                          beta_interval=int(pnode.get("beta_interval", 50)),
                          advance_on_convergence=bool(pnode.get("advance_on_convergence", True)), robust=robust,
                          robust_delta=float(pnode.get("robust_delta", 0.1)),
                          robust_volume_interval=int(pnode.get("robust_volume_interval", 1)),
                          # a report on a non-robust design (TopologyOptimizer.cpp:396-420); robust runs ignore it
                          erosion_check=bool(pnode.get("erosion_check", False)) and not robust)
        pj = projection
        if not (0.0 < pj["eta"] < 1.0):
            raise DeckError(f"topology.projection.eta must lie in (0, 1), got {pj['eta']}")
        if not (pj["beta_start"] > 0.0 and math.isfinite(pj["beta_start"])):
            raise DeckError(f"topology.projection.beta_start must be positive, got {pj['beta_start']}")
        if not (pj["beta_max"] >= pj["beta_start"] and math.isfinite(pj["beta_max"])):
            raise DeckError(f"topology.projection.beta_max ({pj['beta_max']}) must be at least beta_start "
                            f"({pj['beta_start']})")
        if pj["beta_max"] > 1024.0:
            raise DeckError(f"topology.projection.beta_max = {pj['beta_max']} is beyond any useful sharpness")
        if not pj["beta_factor"] > 1.0:
            raise DeckError(f"topology.projection.beta_factor must exceed 1, got {pj['beta_factor']}")
        if pj["beta_interval"] < 1:
            raise DeckError(f"topology.projection.beta_interval must be at least 1, got {pj['beta_interval']}")
        d = pj["robust_delta"]
        if (robust or pj["erosion_check"]) and not (d > 0.0 and pj["eta"] - d > 0.0 and pj["eta"] + d < 1.0):
            raise DeckError(f"topology.projection.robust_delta = {d} must be positive and keep eta +- delta inside "
                            f"(0, 1) (eta = {pj['eta']})")
        if robust and pj["robust_volume_interval"] < 1:
            raise DeckError("topology.projection.robust_volume_interval must be at least 1")
    filt = topo.get("filter", {})
    ftype = str(ov.get("filter_type") or filt.get("type", "density")).lower()
    if ftype == "off":
        ftype = "none"
    if ftype == "sensitivity":
        _unsupported("topology.filter.type", "the sensitivity filter")
    if ftype not in ("none", "density"):
        raise DeckError(f"unknown filter type '{ftype}' (expected none|density|sensitivity)")
    radius = float(filt.get("radius", 0.0))
    radius_elements = float(filt.get("radius_elements", 1.5))
    if "filter_radius" in ov:
        radius, radius_elements = float(ov["filter_radius"]), 0.0
    if "filter_radius_elements" in ov:
        radius, radius_elements = 0.0, float(ov["filter_radius_elements"])
    if ftype == "none":
        radius = 0.0
    elif not radius > 0.0:
        # Configuration::resolved_filter_radius (Config.cpp:360-370)
        if not radius_elements > 0.0:
            raise DeckError("topology.filter needs either 'radius' (metres) or a positive 'radius_elements'")
        radius = radius_elements * m.mean_element_size()
    m.filter_radius = radius

    # passive regions (Config.cpp:1710-1727, DesignDomain.cpp:40-67)
    tag = np.zeros(m.nelem, dtype=np.int8)  # 0 free, 1 solid, 2 void
    for pi, pr in enumerate(topo.get("passive_regions") or []):
        pname = pr.get("name", f"passive{pi}")
        kind = pr.get("type", "solid")
        if kind not in ("solid", "void"):
            raise DeckError(f"'topology.passive_regions[{pi}].type' must be 'solid' or 'void', got '{kind}'")
        solid = kind == "solid"
        density = float(pr.get("density", 0.0))
        if not solid and not (0.0 <= density < 1.0):
            raise DeckError(f"passive void region '{pname}' needs a void density in [0, 1) (got {density})")
        els = geo.elements_of(parse_region(pr.get("region"), dim, f"topology.passive_regions[{pi}].region"), pname)
        if len(els) == 0:
            raise DeckError(f"passive region '{pname}' selected no elements; check its geometry against the design "
                            "domain extents")
        t = 1 if solid else 2
        clash = (tag[els] != 0) & (tag[els] != t)
        if clash.any():
            raise DeckError(f"element {int(els[clash][0])} is claimed by both a passive solid and a passive void "
                            f"region (last one: '{pname}'); make the regions disjoint")
        tag[els] = t
        m.set_passive(els, solid, void_density=density)
    if not (tag == 0).any():
        raise DeckError("every element is passive; there is nothing left to optimise")
    vol = (lx * ly * thickness) if dim == 2 else lx * ly * lz
    ev = vol / m.nelem  # every cell of a structured mesh has the same measure
    target = vf * vol
    solid_volume = float((tag == 1).sum()) * ev
    if solid_volume > target:
        raise DeckError(f"the passive solid regions already occupy {solid_volume} m^3, more than the volume target "
                        f"of {target} m^3")
    if solid_volume + float((tag == 0).sum()) * ev < target:
        raise DeckError(f"the volume target {target} m^3 is unreachable with the passive void regions")

    init = float(topo.get("initial_density", -1.0))
    start = init if init >= 0.0 else vf
    if not (0.0 <= start <= 1.0):
        raise DeckError(f"initial density must lie in [0, 1] (got {start})")

    # TopologyOptimizerOptions (Config.cpp:1561-1576)
    run = dict(
        max_iterations=int(ov.get("max_iterations") or opt.get("max_iterations", 200)),
        change_tol=float(opt.get("change_tolerance", 1e-2)),
        objective_tol=float(opt.get("objective_tolerance", 0.0)),
        window=int(opt.get("objective_window", 5)),
        cont_steps=int(opt.get("continuation_steps", 1)),
        penalty_start=float(opt.get("penalty_start", 1.0)),
        cont_iterations=int(opt.get("continuation_iterations", 25)),
        move=float(opt.get("move_limit", 0.2)),
        damping=float(opt.get("damping", 0.5)),
        projection=projection,
    )
    lin = doc.get("solver", {}).get("linear", {})
    # an iterative deck's CG stops at iterative_tolerance (LinearSolver.cpp:265, Config.cpp:1080-1084); a direct
    # one's residual_tolerance is the acceptance bound of its factorisation, used here as the CG tolerance
    if str(lin.get("type", "simplicial_ldlt")).lower() in ("amg_cg", "conjugate_gradient", "cg", "auto"):
        tol = float(lin.get("iterative_tolerance", 1e-12))
    else:
        tol = float(lin.get("residual_tolerance", 1e-8))
    amg = _amg_options(lin)
    # linear.warm_start (Config.cpp:1086): solves start from the previous solution
    warm = lin.get("warm_start", True)
    if not isinstance(warm, bool):
        raise DeckError("'solver.linear.warm_start' must be true or false")
    run["warm_start"] = warm
    density = float(doc.get("material", {}).get("density", 0.0))
    thr = float(opt.get("interpretation_threshold", 0.5))
    # checks that only report on the final design and do not change the optimisation
    skipped = [f"topology.{k}" for k in ("length_scale_check", "overhang_check")
               if isinstance(topo.get(k), dict) and topo[k].get("enabled", False)]
    if not (0.0 < thr < 1.0):
        raise DeckError("optimizer.interpretation_threshold must lie in (0, 1)")
    # modal block (Config.cpp:1118-1143)
    mnode = dict(doc.get("modal", {}))
    if "modes" in ov:  # --modes n (apps/sparlab_topopt.cpp:295-297)
        mnode["enabled"] = True
        mnode["num_modes"] = int(ov["modes"])
    modal = flags = None
    if mnode.get("enabled", False):
        if not density > 0.0:
            raise DeckError("modal analysis is enabled but material.density is zero; set a positive density [kg/m^3]")
        mass_type = str(mnode.get("mass_type", "consistent")).lower()
        if mass_type not in ("consistent", "lumped"):
            raise DeckError(f"unknown modal.mass_type '{mass_type}' (consistent or lumped)")
        if float(mnode.get("shift_factor", 0.0)) != 0.0:
            _unsupported("modal.shift_factor", "a shifted eigenproblem")
        modal = dict(num_modes=int(mnode.get("num_modes", 6)), mass_type=mass_type,
                     max_iterations=int(mnode.get("max_iterations", 200)),
                     tolerance=float(mnode.get("tolerance", 1e-10)),
                     residual_tolerance=float(mnode.get("residual_tolerance", 1e-6)),
                     rigid_body_ratio=float(mnode.get("rigid_body_ratio", 1e-10)),
                     linear_tolerance=tol, warm_start=warm)
        if modal["num_modes"] < 1:
            raise DeckError("modal analysis requires num_modes >= 1")
        flags = dict(analyse_optimised_topology=bool(mnode.get("analyse_optimised_topology", True)),
                     compare_mass_matched_baseline=bool(mnode.get("compare_mass_matched_baseline", True)))
    return Deck(name=name, description=doc.get("description", ""), source=source, model=m, origin=origin,
                case_names=names, case_weights=weights, tolerance=tol, filter_type=ftype, filter_radius=radius,
                initial_density=start, passive_solid=int((tag == 1).sum()), passive_void=int((tag == 2).sum()),
                run=run, history_stride=int(opt.get("history_stride", 1)), density=density, modal=modal,
                modal_flags=flags, interpretation_threshold=thr, skipped_reports=skipped, amg=amg,
                constraints=constraints)


def load_deck(path: str, overrides: dict | None = None) -> Deck:
    return parse_deck(read_deck(path), source=os.path.abspath(path), overrides=overrides)
