"""Benchmark of the SIMP linear-solve path on one MI355X.

Workload (the metric's configuration): 3-D cantilever, Hex8 128 x 64 x 64
unit cells (524 288 elements, 1 622 400 free DOFs), x = 0 face clamped, unit
total load on the free-end bottom edge, E = 1, nu = 0.3, SIMP p = 3,
Emin/E0 = 1e-9, density filter r = 2 cells, volume fraction 0.2, OC updates
from the uniform design x0 = 0.2.

One step = one SIMP iteration with every input resident in HBM: density
filter, SIMP factors, K_ff assembly, AMG numeric setup (aggregates reused),
PCG from the warm start to the true relative residual 1e-8, compliance,
element sensitivities, filter adjoints and the OC update (the loop body of
TopologyOptimizer::run_oc, which ends with one re-evaluation of the final
design). One load case: one solve per step, plus that final evaluation.

Two runs of the same loop: the headline "mixed+learned" (FP32 AMG hierarchy and
V-cycle under FP64 flexible CG, learned warm start trained on the warmup
iterations' solves) and the plain FP64 GPU path (the drop-in: FP64 everywhere,
warm start from the previous solution). `value` is the headline run's solves
per second; the FP64 run's figure and the ratio sit beside it.

The JSON line also carries the roofline of the dominant kernel (the one with
the largest total time among those timed with HIP events on the library's
stream over the timed region), the others beside it, two variants of the
headline run (bfloat16 coarse operators; coarse operators reused across
evaluations), the headline's agreement with the FP64 oracle at the last timed
design, and the CPU baseline: full refcpu SIMP iterations of the same workload
on the host, min and median over repeats, at all host threads and at 4.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Hex8 1.6M-DOF AMG-PCG solves/s @rtol 1e-8: mixed+learned vs FP64 GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64, vector and matrix


def cpu_baseline(model, x_start, tol, thread_counts, repeats=3):
    """refcpu (the C++17/OpenMP restatement of the reference; Eigen is absent) on the host, timed as
    apps/sparlab_bench.cpp:12-16,240-310 times the reference: per thread count, the aggregation on the
    initial design and one warm-up SIMP iteration (untimed), then `repeats` SIMP iterations of the same
    workload from the design of the last timed GPU iteration, each = evaluate (filter, SIMP, K_ff
    assembly, AMG numeric setup with reused aggregates, PCG warm-started from the previous solution to
    tol, sensitivities, filter adjoints) + the OC update. A solve = assembly + numeric setup + PCG
    (timed inside refcpu); min and median over the repeats."""
    from oracle import refcpu

    def log(msg):
        print(f"bench[cpu]: {msg}", file=sys.stderr, flush=True)

    out = {}
    for nt in thread_counts:
        refcpu.set_num_threads(nt)
        log(f"{nt} threads: warm-up SIMP iteration")
        t0 = time.perf_counter()
        o = refcpu.Oracle(model, tol=tol)
        prime_aggregates(o, model)
        x = np.array(x_start, dtype=np.float64)
        ev = o.evaluate(x)
        x, _ = o.oc(x, ev["dc_dx"], ev["dv_dx"])
        log(f"{nt} threads: model, filter and warm-up iteration {time.perf_counter() - t0:.1f} s")
        it_s, solve_s, pcg_s, its = [], [], [], []
        for rep in range(repeats):
            t1 = time.perf_counter()
            ev = o.evaluate(x)
            xn, _ = o.oc(x, ev["dc_dx"], ev["dv_dx"])
            it_s.append(time.perf_counter() - t1)
            log(f"{nt} threads: repeat {rep + 1}: {it_s[-1]:.2f} s, {ev['iterations']} CG iterations")
            sec = o.last_evaluate_seconds()
            solve_s.append(sec["assemble"] + sec["setup"] + sec["pcg"])
            pcg_s.append(sec["pcg"])
            its.append(ev["iterations"])
            x = xn
        log(f"{nt} threads: SIMP iterations {[round(v, 2) for v in it_s]} s, solves {[round(float(v), 2) for v in solve_s]} s, "
            f"PCG {[round(float(v), 2) for v in pcg_s]} s, CG its {its}")
        out[nt] = {"solves_per_s": 1.0 / min(solve_s), "solves_per_s_median": 1.0 / float(np.median(solve_s)),
                   "pcg_only_solves_per_s": 1.0 / min(pcg_s),
                   "pcg_only_solves_per_s_median": 1.0 / float(np.median(pcg_s)),
                   "simp_iterations_per_s": 1.0 / min(it_s),
                   "simp_iterations_per_s_median": 1.0 / float(np.median(it_s)),
                   "pcg_iterations_per_solve": float(np.mean(its)), "repeats": repeats}
        del o
    nmax = max(thread_counts)
    best = out[nmax]
    return {"value": best["solves_per_s"], "unit": "solves/s", "cores": nmax, "kind": "port",
            "sample": f"refcpu (C++17/OpenMP restatement of the reference; Eigen absent, so it stands in for the "
                      f"reference's threaded Eigen path) on the host: {repeats} SIMP iterations of the same "
                      f"workload (128x64x64 unless --size says otherwise) from the design of the last timed GPU "
                      f"iteration, after one untimed warm-up iteration, at {nmax} and at 4 threads; a solve = K_ff "
                      f"assembly + AMG numeric setup (aggregates reused) + PCG warm-started from the previous "
                      f"solution to rtol {tol}; value = 1 / the fastest solve at {nmax} threads (min over the "
                      f"repeats, as sparlab_bench reports), medians beside it",
            "threads": {str(k): v for k, v in out.items()}}


def prime_aggregates(o, model):
    """The first AMG setup (aggregation) on the initial design, as the GPU runs and the
    reference's loop do; later setups reuse the aggregates."""
    x0 = model.initial_design()
    o.assemble_kff(model.emin_ratio + (1.0 - model.emin_ratio) * np.clip(o.filter(x0), 0.0, 1.0) ** model.penalty)
    o.amg_setup()


def oracle_reference(model, x):
    """refcpu's evaluation of design x at rtol 1e-10 (1e-9 when the capped solve cannot reach
    1e-10 at high density contrast): compliance, free displacements and sensitivities."""
    from oracle import refcpu

    for tol_ref in (1e-10, 1e-9):
        t0 = time.perf_counter()
        print(f"bench[oracle]: refcpu evaluation at rtol {tol_ref:g} of the last timed design", file=sys.stderr,
              flush=True)
        try:
            o = refcpu.Oracle(model, tol=tol_ref, max_iterations=400)
            prime_aggregates(o, model)
            eo = o.evaluate(x, want_u=True)
        except refcpu.OracleError as e:
            print(f"bench[oracle]: rtol {tol_ref:g} not reached ({e}), {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
            continue
        print(f"bench[oracle]: {eo['iterations']} CG iterations, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
              flush=True)
        return {"tol": tol_ref, "compliance": eo["compliance"], "u": eo["u"][0][model.free_dofs()],
                "dc": eo["dc_dx"], "iterations": eo["iterations"]}
    return None


def cpu_side(directory, size):
    """The host-only part of the bench in its own process (no GPU runtime, whose threads and
    OpenMP runtime would share the cores with refcpu's): the oracle reference at the design in
    directory/x.npy and the CPU baseline, written to directory/cpu.npz and cpu.json."""
    from oracle import refcpu
    from sparlab_amd import model as M

    nx, ny, nz = (int(v) for v in size.split(","))
    m = M.cantilever_3d(nx, ny, nz, volume_fraction=0.2, radius_cells=2.0)
    x = np.load(os.path.join(directory, "x.npy"))
    tol = float(open(os.path.join(directory, "tol.txt")).read())
    ref = oracle_reference(m, x)
    if ref is not None:
        np.savez(os.path.join(directory, "cpu.npz"), u=ref["u"], dc=ref["dc"])
    nproc = refcpu.num_threads()
    base = cpu_baseline(m, x, tol, sorted({nproc, min(4, nproc)}, reverse=True))
    meta = {"baseline": base, "reference": None if ref is None else
            {k: ref[k] for k in ("tol", "compliance", "iterations")}}
    with open(os.path.join(directory, "cpu.json"), "w") as f:
        json.dump(meta, f)


def run_cpu_side(m, g, x, tol, size):
    """Runs cpu_side in a child process and compares the headline path (problem g, rtol tol)
    with the oracle reference at design x; returns (agreement, cpu_baseline)."""
    import subprocess
    import tempfile

    d = tempfile.mkdtemp(prefix="sparlab_bench_")
    np.save(os.path.join(d, "x.npy"), x)
    with open(os.path.join(d, "tol.txt"), "w") as f:
        f.write(repr(tol))
    r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--cpu-side", d, "--size", size])
    if r.returncode != 0 or not os.path.exists(os.path.join(d, "cpu.json")):
        print(f"bench: the CPU side ended with status {r.returncode}", file=sys.stderr, flush=True)
        return None, None
    meta = json.load(open(os.path.join(d, "cpu.json")))
    agree = None
    ref = meta["reference"]
    if ref is not None:
        arr = np.load(os.path.join(d, "cpu.npz"))
        eg = g.evaluate(x, tol=tol)
        ug = g.solution(0)
        uo, dco = arr["u"], arr["dc"]
        agree = {"compliance_rel": abs(eg["compliance"] - ref["compliance"]) / abs(ref["compliance"]),
                 "u_rel": float(np.linalg.norm(ug - uo) / np.linalg.norm(uo)),
                 "dc_rel": float(np.linalg.norm(eg["dc_dx"] - dco) / np.linalg.norm(dco)),
                 "dc_max_rel": float(np.abs(eg["dc_dx"] - dco).max() / np.abs(dco).max()),
                 "oracle_cg_iterations": ref["iterations"],
                 "reference": f"refcpu at rtol {ref['tol']:g} against the headline path at rtol {tol}, at the "
                              f"design of the last timed SIMP iteration"}
    for f in ("x.npy", "tol.txt", "cpu.npz", "cpu.json"):
        if os.path.exists(os.path.join(d, f)):
            os.remove(os.path.join(d, f))
    os.rmdir(d)
    return agree, meta["baseline"]


def simp_run(N, m, x0, warmup, steps, precision, learned, tol, label, coarse_reuse=0):
    """Warmup iterations (they train the network when learned), then `steps` timed ones."""
    t0 = time.perf_counter()
    g = N.HipProblem(m, device=0)
    g.assemble(1e-9 + (1 - 1e-9) * np.clip(g.filter(x0), 0, 1) ** 3)
    # first setup: aggregation and patterns, untimed
    g.amg_setup(precision=precision, coarse_reuse=coarse_reuse)
    print(f"bench[{label}]: problem and first AMG setup {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    if learned:
        g.warm_start(True, train_solves=warmup, min_samples=4, steps_per_solve=16, buffer=8, learning_rate=1e-3)
    x1, h1, _, _ = g.run_oc(x0, warmup, change_tol=0.0, tol=tol)
    # one more OC update so the timed loop starts at a design no solve has seen
    ev = g.evaluate(x1, tol=tol)
    x1, _ = g.oc(x1, ev["dc_dx"], ev["dv_dx"])
    print(f"bench[{label}]: warmup {warmup} iterations, CG its {h1[:, 11].astype(int).tolist()}", file=sys.stderr,
          flush=True)
    g.ctx.timing(True)
    g.ctx.synchronize()
    t0 = time.perf_counter()
    x2, h2, done, c2 = g.run_oc(x1, steps, change_tol=0.0, tol=tol)
    g.ctx.synchronize()
    elapsed = time.perf_counter() - t0
    timers = {name: g.ctx.timer(name) for name in ("spmv", "galerkin_ap", "galerkin_raw", "assemble")}
    g.ctx.timing(False)
    its = h2[:, 11].astype(int).tolist()
    print(f"bench[{label}]: {done} timed iterations in {elapsed:.3f} s, CG its {its}, c = {c2:.8e}",
          file=sys.stderr, flush=True)
    st = g.amg_stats()
    phases = {k: round(1e3 * getattr(st, k), 2) for k in ("setup_seconds", "check_seconds", "spectral_seconds",
                                                          "smoothing_seconds", "galerkin_seconds",
                                                          "coarse_seconds")}
    print(f"bench[{label}]: last AMG setup (ms) {phases}, unknowns {[st.unknowns[i] for i in range(st.levels)]}",
          file=sys.stderr, flush=True)
    return dict(problem=g, x_first=x1, x=x2, steps=done, elapsed=elapsed, iterations=its, compliance=c2,
                timers=timers, phases=phases, warm=g.warm_start_stats() if learned else None)


def level_info(g, level, which):
    import ctypes as C

    from sparlab_amd import native as N

    info = (C.c_longlong * 3)()
    g.ctx.check(N.load().slx_amg_level_info(g.ctx.h, level, which, info))
    return list(info)


def measured_traffic(name, sizes):
    """HBM bytes per launch from the committed rocprofv3 counter summary, when its sizes (three
    integers: nnz of A, B, C for a product; free DOFs, elements, nodes for the operator) match the run."""
    path = os.path.join(ROOT, "profiles", "traffic_level0_kernels_cantilever_128x64x64.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#") or parts[0] != name or len(parts) != 8:
                continue
            if tuple(int(v) for v in parts[1:4]) != tuple(sizes):
                return None
            fk, wk, fs, ws = float(parts[4]), float(parts[5]), float(parts[6]), float(parts[7])
            return fk * 1024 * fs + wk * 1024 * ws
    return None


def smoother_roofline(g, fp32):
    """The level-1 Chebyshev step (CSR with shared column lists) of the same hierarchy, timed with
    HIP events around 50 back-to-back launches; priced in HBM bytes: 4 (FP32) or 8 (FP64) per value,
    4 per shared column index, 8 per row (row pointer and column base) and 7 vector words per
    row (r in and out, d in and out, x read and written, D^-1)."""
    rows, _, nnz = level_info(g, 1, 0)
    nodes, ncol, lanes = level_info(g, 1, 8)
    word = 4 if fp32 else 8
    launches = 50
    avg = g.time_smoother(1, launches)
    g.amg_apply(np.ones(g.nfree))  # the level's vectors back to a V-cycle's state
    nbytes = word * nnz + 4 * (ncol if ncol else nnz) + 8 * rows + 7 * word * rows
    ach = nbytes / (avg * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": (f"k_node_spmv<{-lanes}> (level-1 Chebyshev step, {-lanes} lanes per node" if lanes < 0 else
                       f"k_cheb_step<{lanes}> (level-1 Chebyshev step, {lanes} lanes per row")
                      + f", {'FP32' if fp32 else 'FP64'} CSR with shared column lists)",
            "bytes_per_launch": nbytes, "avg_ms": avg, "launches": launches, "total_ms": avg * launches,
            "sizes": {"rows": rows, "nnz": nnz, "nodes": nodes, "shared_columns": ncol}}


def mf_smoother_roofline(g, m, fp32, vcycles):
    """The level-0 Chebyshev step through the matrix-free operator (k_mf_hex8 mode 3), the kernel
    with the largest share of a SIMP step: timed with HIP events around 50 back-to-back launches.
    Algorithmic HBM bytes per launch: 7 vector words per free DOF (d, r and x read; r, d and x
    written; D^-1 read), one word per element factor and one 4-byte map entry per node (the node's
    first free index and free-component mask).
    Launches in the timed region: 2 (degree - 1) = 4 per V-cycle."""
    word = 4 if fp32 else 8
    launches = 50
    avg = g.time_smoother(0, launches)
    g.amg_apply(np.ones(g.nfree))  # the level's vectors back to a V-cycle's state
    nnodes = m.nnodes
    nbytes = 7 * word * g.nfree + word * m.nelem + 4 * nnodes
    ach = nbytes / (avg * 1e-3) / 1e9
    in_run = 4 * vcycles
    tile = "x".join(str(v) for v in g.mf_tile())
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(f"mf_cheb{32 if fp32 else 64}_{tile}", (g.nfree, m.nelem, nnodes)),
            "kernel": f"k_mf_hex8<{'float' if fp32 else 'double'},3> (level-0 Chebyshev step, matrix-free, "
                      f"{tile}-node tiles)",
            "bytes_per_launch": nbytes, "avg_ms": avg, "launches": in_run, "total_ms": avg * in_run,
            "launches_timed": launches, "sizes": {"free_dofs": g.nfree, "elements": m.nelem, "nodes": nnodes}}


def rooflines(g, m, timers):
    """Roofline of each kernel timed with HIP events over the timed region, largest total first.

    galerkin_ap / galerkin_raw: the level-0 row products A P and R (A P), HBM-bound;
    algorithmic bytes = 8 (nnz A + nnz B + nnz C) values + 16 per run (descriptor and
    entry range) + 4 per entry (block positions) + 12 per output node (order, run
    pointer, row pointer) + 4 per output row (A row pointer).
    spmv: the matrix-free K_ff product of CG (with x.y partials), FP64 compute-bound;
    1200 flop per element (Ke0 u_e: 576 FMA, the factor: 24 more).
    """
    out = []
    n, nnz = g.nfree, g.kff_nnz
    sizes = {"P": level_info(g, 0, 1)[2], "R": level_info(g, 0, 3)[2], "AP": level_info(g, 0, 4)[2],
             "RAW": level_info(g, 0, 5), "plan_ap": level_info(g, 0, 6), "plan_raw": level_info(g, 0, 7)}
    for name, (ms, launches) in timers.items():
        if not launches or name == "assemble":
            continue
        avg = ms / launches
        if name == "spmv":
            flops = 1200.0 * m.nelem
            ach = flops / (avg * 1e-3) / 1e12
            out.append({"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": ach / FP64_PEAK_TFLOPS, "traffic": None,
                        "kernel": "k_mf_hex8<double,1> (CG q = K_ff p matrix-free, FP64 vector FMA)",
                        "flops_per_launch": flops, "avg_ms": avg, "launches": launches, "total_ms": ms})
            continue
        if name == "galerkin_ap":
            a_nnz, b_nnz, c_nnz, rows = nnz, sizes["P"], sizes["AP"], n
            nodes, runs, entries = sizes["plan_ap"]
            label = "k_row_product<3,3> (level-0 A P, FP64)"
        else:
            a_nnz, b_nnz, c_nnz, rows = sizes["R"], sizes["AP"], sizes["RAW"][2], sizes["RAW"][0]
            nodes, runs, entries = sizes["plan_raw"]
            label = "k_row_product<6,3> (level-0 R (A P), FP64)"
        nbytes = 8 * (a_nnz + b_nnz + c_nnz) + 16 * runs + 4 * entries + 12 * nodes + 4 * rows
        ach = nbytes / (avg * 1e-3) / 1e9
        sz = {"nnz_a": a_nnz, "nnz_b": b_nnz, "nnz_c": c_nnz, "nodes": nodes, "runs": runs, "entries": entries}
        out.append({"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                    "traffic": measured_traffic(name, (a_nnz, b_nnz, c_nnz)), "kernel": label,
                    "bytes_per_launch": nbytes,
                    "avg_ms": avg, "launches": launches, "total_ms": ms, "sizes": sz})
    out.sort(key=lambda r: -r["total_ms"])
    return out


def cases_leg(args, prec, tol, log):
    """The bracket's three load cases (vertical, lateral, torsional) as one block
    right-hand side against the same evaluation solving them one at a time: each
    timed step is one evaluation from a zero start (filter, SIMP, assembly, AMG
    numeric setup, the three solves, sensitivities); solve_seconds is the PCG part."""
    from sparlab_amd import batch as B
    from sparlab_amd import model as M
    from sparlab_amd import native as N

    br = B.bracket_2d()
    x = M.seeded_density(br.nelem, seed=20261203)
    out = {}
    for mode in ("block", "one_at_a_time"):
        if mode == "one_at_a_time":
            os.environ["SPARLAB_NO_BLOCK_CASES"] = "1"
        ctx = N.Context(0)
        g = N.HipProblem(br, ctx=ctx)
        g.assemble()
        g.amg_setup(precision=prec)
        for _ in range(max(2, args.warmup // 4)):
            g.evaluate(x, tol=tol, warm_start=False)
        ctx.synchronize()
        solve, its, comp = 0.0, 0, None
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ev = g.evaluate(x, tol=tol, warm_start=False)
            solve += ev["solve_seconds"]
            its = ev["iterations"]
            comp = ev["compliance"]
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        out[mode] = dict(eval_ms=1e3 * wall, solve_ms=1e3 * solve / args.steps, iterations=its, compliance=comp)
        log(f"bench[cases]: {mode}: evaluation {1e3 * wall:.2f} ms, three solves {1e3 * solve / args.steps:.2f} ms, "
            f"{its} PCG iterations in all, compliance {comp!r}")
        ctx.close()
        os.environ.pop("SPARLAB_NO_BLOCK_CASES", None)
    # The two runs use separate contexts, and the mixed path's FP32 V-cycle picks its
    # CSR lane counts per context by timing, so their summation orders (and the last
    # bits) can differ; both solve to the same rtol.
    rel = abs(out["block"]["compliance"] - out["one_at_a_time"]["compliance"]) / abs(out["one_at_a_time"]["compliance"])
    out["compliance_rel_diff"] = rel
    assert rel <= 1e-9, f"block and case-by-case compliances differ by {rel:.3g} (relative)"
    assert abs(out["block"]["iterations"] - out["one_at_a_time"]["iterations"]) <= 3, "iteration counts differ"
    return br, out


def designs_leg(args, prec, tol, log):
    """32 MBB designs of seeded uneven sizes as one block-diagonal system and one
    segmented CG, against the same designs solved one at a time."""
    from sparlab_amd import batch as B
    from sparlab_amd import native as N

    ctx = N.Context(0)
    t0 = time.perf_counter()
    parts = B.c4_parts(bracket=False)  # the MBB designs (the bracket's cases run as a block)
    bt = B.assemble(parts, 2, ctx=ctx)
    nseg = len(bt.seg_ptr) - 1
    log(f"bench[designs]: {nseg} segments, {len(bt.b)} unknowns, {len(bt.val)} nonzeros, assembled in "
        f"{time.perf_counter() - t0:.1f} s")
    s = B.solver_for(bt, tol=tol, ctx=ctx, precision=prec)  # first setup: aggregation, untimed
    log("bench[designs]: hierarchy built")
    for _ in range(max(1, args.warmup // 4)):
        s.solve_segmented(bt.seg_ptr, bt.b)
    ctx.synchronize()
    log("bench[designs]: warm")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.factorize(bt.ptr, bt.col, bt.val)  # values upload + numeric setup (aggregates reused)
    ctx.synchronize()
    t_setup = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x, res = s.solve_segmented(bt.seg_ptr, bt.b)
    ctx.synchronize()
    t_pcg = (time.perf_counter() - t0) / args.steps
    its = [r[0] for r in res]
    assert all(r[2] for r in res), "a segment did not converge"
    log(f"bench[designs]: setup {1e3 * t_setup:.1f} ms, PCG {1e3 * t_pcg:.1f} ms per batch; iterations {its}")
    # the same designs one at a time: a context and a hierarchy each (untimed), the solves timed
    import scipy.sparse as sp

    K = sp.csr_matrix((bt.val, bt.col, bt.ptr))
    singles = []
    for k in range(nseg):
        a, b = int(bt.seg_ptr[k]), int(bt.seg_ptr[k + 1])
        Kg = K[a:b, a:b].tocsr()
        nodes = bt.unknowns[a:b] // 2
        n0 = int(nodes.min())
        one = N.HipAmgCgSolver(tolerance=tol, ctx=N.Context(0), precision=prec)
        one.set_layout(2, bt.coords[n0:int(nodes.max()) + 1], bt.unknowns[a:b] - 2 * n0, dofs_per_node=2)
        one.factorize(Kg.indptr, Kg.indices, Kg.data)
        singles.append((one, bt.b[a:b]))
    for one, bg in singles:
        one.solve(bg)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for one, bg in singles:
            one.solve(bg)
    t_single = (time.perf_counter() - t0) / args.steps
    for one, _ in singles:
        one.ctx.close()
    log(f"bench[designs]: one at a time {1e3 * t_single:.1f} ms per batch")
    return dict(segments=nseg, unknowns=int(len(bt.b)), nonzeros=int(len(bt.val)), pcg_ms=1e3 * t_pcg,
                setup_ms=1e3 * t_setup, iterations=its, one_at_a_time_pcg_ms=1e3 * t_single,
                speedup_vs_one_at_a_time=t_single / t_pcg)


def batch_main(args):
    """The batched workload: a 240 x 160 two-bolt bracket's three load cases as one
    block right-hand side. The leg of 32 MBB designs as one segmented CG
    (designs_leg) is switched off."""
    tol = 1e-8
    prec = 1 if args.precision == "mixed" else 0

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    br, cases = cases_leg(args, prec, tol, log)
    blk, one = cases["block"], cases["one_at_a_time"]
    if os.environ.get("SPARLAB_BATCH_BENCH") == "1":
        # switched off: a profiled run (rocprofv3 --kernel-trace) of the batch bench with this leg
        # enabled ended in a GPU memory fault whose cause has not been found (the load-case leg had
        # run under the profiler without one); the segmented solve stays covered by its tests
        raise SystemExit("bench: the 32-design segmented leg is switched off (a profiled run of it faulted the "
                         "GPU, cause not found); the batch bench runs the load-case leg only")
    designs = None
    solves = 3 + (designs["segments"] if designs else 0)
    ms = blk["solve_ms"] + (designs["pcg_ms"] if designs else 0.0)
    out = {
        "metric": "batched Q4 AMG-PCG solves/s @rtol 1e-8 (bracket 240x160, 3 load cases as a block RHS"
                  + (" + 32 MBB designs)" if designs else ")"),
        "value": solves / (1e-3 * ms), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if prec else "f64", "data": "synthetic",
        "config": {"workload": f"two-bolt bracket Q4 240x160 ({br.ndofs} DOFs), vertical, lateral and torsional "
                               f"unit loads, seeded densities, SIMP p 3; one block PCG for the three cases"
                               + ("; 32 MBB designs (numpy PCG64(4242) sizes, nx 60..240) as one segmented CG"
                                  if designs else ""),
                   "cases_block_solve_ms": blk["solve_ms"], "cases_one_at_a_time_solve_ms": one["solve_ms"],
                   "cases_block_speedup": one["solve_ms"] / blk["solve_ms"],
                   "cases_block_evaluation_ms": blk["eval_ms"], "cases_one_at_a_time_evaluation_ms": one["eval_ms"],
                   "cases_iterations": blk["iterations"], "cases_compliance": blk["compliance"],
                   "cases_compliance_rel_diff_one_at_a_time": cases["compliance_rel_diff"],
                   "designs": designs},
    }
    print(json.dumps(out))


def warmstart_main(args):
    """The learned warm start's workload: Hex8 96 x 48 x 48 cantilever (unit cells,
    volume fraction 0.3, filter r 2), 200 SIMP/OC iterations from the uniform
    design at rtol 1e-8 on the mixed hierarchy, once from the plain warm start
    (the previous solution, as the reference), once from the previous solution
    plus the K-optimal multiple of the previous change (the learned start's
    machinery without the network), and once with the network trained online on
    the first 40 iterations' solves. The iterations after training are compared:
    PCG iterations, wall time and the K-norm distance of each starting point from
    the converged solution."""
    from sparlab_amd import model as M
    from sparlab_amd import native as N

    nx, ny, nz = 96, 48, 48
    total, train, tol = 200, 40, 1e-8

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    m = M.cantilever_3d(nx, ny, nz, volume_fraction=0.3, radius_cells=2.0)
    runs = {}
    for label in ("plain", "previous_change", "learned"):
        ctx = N.Context(0)
        g = N.HipProblem(m, ctx=ctx)
        x0 = m.initial_design()
        g.assemble(1e-9 + (1 - 1e-9) * np.clip(g.filter(x0), 0, 1) ** 3)
        g.amg_setup(precision=1)
        if label == "learned":
            g.warm_start(True, train_solves=train, min_samples=4, steps_per_solve=16, buffer=8, learning_rate=1e-3)
        elif label == "previous_change":  # no training and no prediction: the extrapolation alone
            g.warm_start(True, train_solves=0, min_samples=1 << 30)
        ctx.synchronize()
        t0 = time.perf_counter()
        x1, h1, _, _ = g.run_oc(x0, train, change_tol=0.0, tol=tol)
        ctx.synchronize()
        t_train = time.perf_counter() - t0
        log(f"bench[warmstart/{label}]: first {train} iterations in {t_train:.2f} s, CG its "
            f"{h1[:, 11].astype(int).tolist()}")
        t0 = time.perf_counter()
        x2, h2, done, c2 = g.run_oc(x1, total - train, change_tol=0.0, tol=tol)
        ctx.synchronize()
        t_rest = time.perf_counter() - t0
        its = h2[:, 11].astype(int)
        st = g.warm_start_stats() if label != "plain" else None
        log(f"bench[warmstart/{label}]: iterations {train + 1}..{train + done} in {t_rest:.2f} s, "
            f"{int(its.sum())} CG iterations, c = {c2:.10e}")
        runs[label] = dict(seconds=t_rest, steps=done, cg_iterations=int(its.sum()), compliance=c2,
                           first_seconds=t_train, stats=st)
        ctx.close()
    pl, pc, le = runs["plain"], runs["previous_change"], runs["learned"]
    st = le["stats"]
    k = max(st["predictions"], 1)
    out = {
        "metric": "learned warm start: SIMP iterations/s after training (Hex8 96x48x48, rtol 1e-8)",
        "value": le["steps"] / le["seconds"], "unit": "SIMP iterations/s", "n_gpus": 1, "steps": le["steps"],
        "warmup": train, "ms_per_step": 1e3 * le["seconds"] / le["steps"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"cantilever Hex8 {nx}x{ny}x{nz}, {total} SIMP/OC iterations from x0 = 0.3, "
                               f"network trained on the first {train} solves, mixed hierarchy, rtol {tol}",
                   "plain_simp_iterations_per_s": pl["steps"] / pl["seconds"],
                   "previous_change_simp_iterations_per_s": pc["steps"] / pc["seconds"],
                   "speedup_vs_plain": pl["seconds"] / le["seconds"],
                   "speedup_vs_previous_change": pc["seconds"] / le["seconds"],
                   "cg_iterations_after_training": {"learned": le["cg_iterations"],
                                                    "previous_change": pc["cg_iterations"],
                                                    "plain": pl["cg_iterations"]},
                   "final_compliance": {"learned": le["compliance"], "previous_change": pc["compliance"],
                                        "plain": pl["compliance"]},
                   "first_iterations_seconds": {"learned": le["first_seconds"], "previous_change": pc["first_seconds"],
                                                "plain": pl["first_seconds"]},
                   "knorm_error_ratio_vs_previous_solution": {
                       "combination": st["energy_ratio_sum"] / k, "net_alone": st["net_ratio_sum"] / k,
                       "previous_change_alone": st["secant_ratio_sum"] / k},
                   "train_seconds": st["train_seconds"], "predict_seconds": st["predict_seconds"],
                   "predictions": st["predictions"], "first_loss": st["first_loss"], "last_loss": st["last_loss"]},
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed SIMP iterations; with the learned warm start their solves train the network")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--size", type=str, default="128,64,64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="headline run without the learned warm start")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-bf16", action="store_true", help="skip the bfloat16 coarse-operator variant run")
    ap.add_argument("--no-reuse", action="store_true", help="skip the coarse-operator reuse variant run")
    ap.add_argument("--workload", choices=["simp", "batch", "warmstart"], default="simp",
                    help="simp: the headline SIMP loop; batch: the batched load-case and design solves; "
                         "warmstart: the learned warm start's 200-iteration run against the plain warm start")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference reading for the HBM byte counters)")
    ap.add_argument("--cpu-side", type=str, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_side:
        cpu_side(args.cpu_side, args.size)
        return
    if args.copy_calibration:
        import torch

        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs the path on a single MI355X and has no "
              f"multi-GPU mode (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        batch_main(args)
        return
    if args.workload == "warmstart":
        warmstart_main(args)
        return

    from sparlab_amd import model as M
    from sparlab_amd import native as N

    nx, ny, nz = (int(v) for v in args.size.split(","))
    m = M.cantilever_3d(nx, ny, nz, volume_fraction=0.2, radius_cells=2.0)
    tol = 1e-8
    x0 = m.initial_design()
    head = simp_run(N, m, x0, args.warmup, args.steps, 1 if args.precision == "mixed" else 0,
                    not args.no_learned, tol, "headline")
    base = simp_run(N, m, x0, args.warmup, args.steps, 0, False, tol, "fp64") if not args.no_fp64 else None
    # the bfloat16 coarse-operator storage variant of the headline run
    b16 = (simp_run(N, m, x0, args.warmup, args.steps, 2, not args.no_learned, tol, "bf16-coarse")
           if not args.no_bf16 and args.precision == "mixed" else None)
    # the headline run with the coarse operators reused across evaluations (coarse_reuse 4)
    reuse = (simp_run(N, m, x0, args.warmup, args.steps, 1, not args.no_learned, tol, "coarse-reuse", coarse_reuse=4)
             if not args.no_reuse and args.precision == "mixed" else None)

    # agreement with the FP64 path at the headline run's final design (per field)
    agree = None
    if base is not None:
        gh, gb = head["problem"], base["problem"]
        eh = gh.evaluate(head["x"], tol=tol)
        eb = gb.evaluate(head["x"], tol=1e-10)
        uh, ub = gh.solution(0), gb.solution(0)
        agree = {"compliance_rel": abs(eh["compliance"] - eb["compliance"]) / abs(eb["compliance"]),
                 "dc_rel": float(np.linalg.norm(eh["dc_dx"] - eb["dc_dx"]) / np.linalg.norm(eb["dc_dx"])),
                 "u_rel": float(np.linalg.norm(uh - ub) / np.linalg.norm(ub)),
                 "reference": "FP64 GPU path (parity-tested against the oracle) at rtol 1e-10"}

    g = head["problem"]
    n, nnz = g.nfree, g.kff_nnz
    solves = head["steps"] + 1  # run_oc re-evaluates the final design
    value = solves / head["elapsed"]
    roofs = rooflines(g, m, head["timers"])
    # V-cycles of the timed region: one per CG iteration and one for each solve's first residual
    vcycles = int(np.sum(head["iterations"])) + len(head["iterations"])
    roofs.append(mf_smoother_roofline(g, m, args.precision == "mixed", vcycles))
    sm = smoother_roofline(g, args.precision == "mixed")
    if sm is not None:
        roofs.append(sm)
    roofs.sort(key=lambda r: -r["total_ms"])
    config = {"workload": f"cantilever Hex8 {nx}x{ny}x{nz} SIMP/OC from x0 = 0.2, {n} free DOFs, rtol {tol}; "
                          f"{'FP32 V-cycle under FP64 flexible CG' if args.precision == 'mixed' else 'FP64'}"
                          f"{' + learned warm start' if not args.no_learned else ''}",
              "elements": m.nelem, "free_dofs": n, "kff_nnz": nnz, "precision": args.precision,
              "learned_warm_start": not args.no_learned,
              "simp_iterations_per_s": head["steps"] / head["elapsed"],
              "pcg_iterations_per_solve": float(np.mean(head["iterations"])),
              "amg_setup_ms": head["phases"]["setup_seconds"]}
    if head["warm"]:
        w = head["warm"]
        k = max(1, w["predictions"])
        config["warm_start"] = {"train_solves": args.warmup, "train_steps": w["train_steps"],
                                "loss_first": w["first_loss"], "loss_last": w["last_loss"],
                                "knorm_error_ratio": w["energy_ratio_sum"] / k,
                                "knorm_error_ratio_net_alone": w["net_ratio_sum"] / k,
                                "knorm_error_ratio_previous_change_alone": w["secant_ratio_sum"] / k}
    if b16 is not None:
        config["bf16_coarse_variant"] = {"solves_per_s": (b16["steps"] + 1) / b16["elapsed"],
                                         "pcg_iterations_per_solve": float(np.mean(b16["iterations"])),
                                         "compliance_rel_to_headline": abs(b16["compliance"] - head["compliance"])
                                         / abs(head["compliance"])}
        b16["problem"].ctx.close()
    if reuse is not None:
        config["coarse_reuse_variant"] = {
            "coarse_reuse": 4, "solves_per_s": (reuse["steps"] + 1) / reuse["elapsed"],
            "simp_iterations_per_s": reuse["steps"] / reuse["elapsed"],
            "pcg_iterations_per_solve": float(np.mean(reuse["iterations"])),
            "compliance_rel_to_headline": abs(reuse["compliance"] - head["compliance"]) / abs(head["compliance"]),
            "note": "coarse operators (A_1 and below) of the last full numeric setup kept while the PCG iteration "
                    "count stays within 20 % of the count after it; level 0 refreshed every evaluation"}
        if base is not None:
            config["coarse_reuse_variant"]["speedup_vs_fp64_gpu"] = (
                (reuse["steps"] + 1) / reuse["elapsed"] / ((base["steps"] + 1) / base["elapsed"]))
        reuse["problem"].ctx.close()
    if base is not None:
        bsolves = base["steps"] + 1
        config["fp64_gpu"] = {"solves_per_s": bsolves / base["elapsed"],
                              "simp_iterations_per_s": base["steps"] / base["elapsed"],
                              "ms_per_step": 1e3 * base["elapsed"] / base["steps"],
                              "pcg_iterations_per_solve": float(np.mean(base["iterations"])),
                              "amg_setup_ms": base["phases"]["setup_seconds"]}
        config["speedup_vs_fp64_gpu"] = value / (bsolves / base["elapsed"])
        config["fp64_agreement"] = agree
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["elapsed"] / head["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": roofs[0] if roofs else None,
        "rooflines_other": roofs[1:],
    }
    if not args.no_cpu_baseline:
        print("bench: oracle agreement and the CPU baseline (child process)", file=sys.stderr, flush=True)
        agree_o, cpu = run_cpu_side(m, g, head["x"], tol, args.size)
        config["oracle_agreement"] = agree_o
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
