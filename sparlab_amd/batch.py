"""Batches of independent SIMP solves on one GPU.

A batch -- many designs of different sizes, or the load cases of one design --
is assembled as one block-diagonal K_ff (each design's K_ff assembled on the
device by its own HipProblem) and solved by the segmented CG of the native
library: one multigrid hierarchy for the whole batch (block-diagonal, since
aggregates never cross blocks that share no entry), each segment with its own
step lengths and convergence test, every kernel launch covering all segments.
The reference solves designs and load cases one at a time
(src/topopt/Sensitivity.cpp:135-176); per segment the result is the same solve
to the same tolerance.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import model as M
from . import native as N


def mbb_sizes(count=32, seed=4242, lo=60, hi=240):
    """Uneven MBB half-beam sizes: nx uniform in [lo, hi], ny = nx / 3 rounded to even."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for nx in rng.integers(lo, hi + 1, size=count):
        ny = max(2, int(round(nx / 3.0 / 2.0)) * 2)
        out.append((int(nx), ny))
    return out


def bracket_2d(nx=240, ny=160):
    """Two-bolt bracket with a pin lug, Q4 plane stress (the geometry of the
    reference's aerospace_bracket deck, unit loads): bolt holes fixed, three
    load cases on the lug ring -- vertical, lateral and torsional (tangential
    tractions of unit total moment)."""
    lx, ly = 0.30, 0.20
    m = M.StructuredModel(dim=2, nx=nx, ny=ny, lx=lx, ly=ly, E=1.0, nu=0.33, thickness=1.0,
                          volume_fraction=0.35, filter_radius=1.5 * lx / nx)
    bolts = np.concatenate([m.select_circle((0.030, 0.045), 0.0105), m.select_circle((0.030, 0.155), 0.0105)])
    m.fix(bolts)
    ring = m.select_annulus((0.270, 0.100), 0.0075, 0.0105)
    c0 = m.add_load_case()
    m.add_point_load(c0, ring, (0.0, -1.0))
    c1 = m.add_load_case()
    m.add_point_load(c1, ring, (1.0, 0.0))
    c2 = m.add_load_case()
    xy = m.node_coords()[ring]
    d = xy - np.array([0.270, 0.100])
    r2 = (d * d).sum(axis=1)
    t = np.stack([-d[:, 1], d[:, 0]], axis=1) / r2.sum()  # sum of r x t over the ring = 1
    for k, node in enumerate(ring):
        m.loads[c2, 2 * node:2 * node + 2] += t[k]
    m.weights = np.ones(3) / 3.0
    void = np.concatenate([m.elements_in_circle((0.030, 0.045), 0.0075), m.elements_in_circle((0.030, 0.155), 0.0075),
                           m.elements_in_circle((0.270, 0.100), 0.0075)])
    m.set_passive(void, solid=False)
    return m


@dataclasses.dataclass
class Batch:
    """A block-diagonal system: CSR, DOF layout, segment offsets, right-hand sides."""
    ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    coords: np.ndarray
    unknowns: np.ndarray
    seg_ptr: np.ndarray
    b: np.ndarray
    dim: int
    labels: list


def assemble(parts, dim, ctx=None):
    """parts: (model, element factors, load cases to solve). Each design's K_ff is
    assembled on the device; segment per (design, load case)."""
    ctx = ctx or N.Context(0)
    ptrs, cols, vals, xs, us, bs, seg, labels = [], [], [], [], [], [], [0], []
    row0 = node0 = nnz0 = 0
    for m, fac, cases in parts:
        g = N.HipProblem(m, ctx=ctx)
        g.assemble(fac)
        p, c, v = g.kff()
        free = g.free()
        xyz = m.node_coords()[:, :dim]
        loads = m.free_loads()
        for case in cases:
            ptrs.append(p[:-1].astype(np.int64) + nnz0)
            cols.append(c + row0)
            vals.append(v)
            xs.append(xyz)
            us.append(free + node0 * dim)
            bs.append(loads[case])
            nnz0 += len(c)
            row0 += len(free)
            node0 += m.nnodes
            seg.append(row0)
            labels.append((m.nx, m.ny, case))
    ptr = np.concatenate(ptrs + [np.array([nnz0])])
    if nnz0 >= 2 ** 31:
        raise ValueError("batch too large for 32-bit CSR indices")
    return Batch(ptr=ptr.astype(np.int32), col=np.concatenate(cols).astype(np.int32), val=np.concatenate(vals),
                 coords=np.concatenate(xs), unknowns=np.concatenate(us).astype(np.int32),
                 seg_ptr=np.array(seg, dtype=np.int32), b=np.concatenate(bs), dim=dim, labels=labels)


def _factors(m, seed, penalty, emin):
    x = M.seeded_density(m.nelem, seed=seed)
    return emin + (1.0 - emin) * x ** penalty


def mbb_designs(count=32, seed=4242, density_seed=20261203):
    """`count` MBB designs of seeded uneven sizes and the seed of each one's densities."""
    return [(M.mbb_beam(nx, ny), density_seed + 1 + k) for k, (nx, ny) in enumerate(mbb_sizes(count, seed))]


def c4_parts(count=32, seed=4242, density_seed=20261203, penalty=3.0, emin=1e-9, bracket=True):
    """The batched workload: the bracket's three load cases (when `bracket`) plus
    `count` MBB designs, each with seeded densities in [0.001, 1] through the SIMP law."""
    parts = []
    if bracket:
        br = bracket_2d()
        parts.append((br, _factors(br, density_seed, penalty, emin), [0, 1, 2]))
    for m, s in mbb_designs(count, seed, density_seed):
        parts.append((m, _factors(m, s, penalty, emin), [0]))
    return parts


def solver_for(batch: Batch, tol=1e-8, ctx=None, **amg):
    s = N.HipAmgCgSolver(tolerance=tol, ctx=ctx, **amg)
    s.set_layout(batch.dim, batch.coords, batch.unknowns, dofs_per_node=batch.dim)
    s.factorize(batch.ptr, batch.col, batch.val)
    return s
