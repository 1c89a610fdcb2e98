"""Benchmark of the SIMP linear-solve path on one MI355X.

Workload (the metric's configuration): 3-D cantilever, Hex8 128 x 64 x 64
unit cells (524 288 elements, 1 622 400 free DOFs), x = 0 face clamped, unit
total load on the free-end bottom edge, E = 1, nu = 0.3, SIMP p = 3,
Emin/E0 = 1e-9, density filter r = 2 cells, volume fraction 0.2; designs are
seeded densities in [0.001, 1] perturbed deterministically from step to step.

One step = one design evaluation of the hot path with every input resident in
HBM: density filter, SIMP factors, K_ff assembly, AMG numeric setup (aggregates
reused), PCG from the previous solution to the true relative residual 1e-8,
compliance, element sensitivities and the filter adjoint. `value` is solves
per second (one load case: one solve per step).

The JSON line also carries the roofline of the dominant kernel (the FP64 CSR
operator product q = K_ff p of PCG), timed with HIP events on the library's
stream over the timed region, and the CPU baseline: the refcpu restatement of
the reference algorithm timed on the host on one solve of the same workload.
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


def designs(x0, k):
    """Design of step k: the seeded field scaled by a smooth deterministic factor."""
    idx = np.arange(x0.size)
    return np.clip(x0 * (1.0 + 0.05 * np.sin(0.7 * k + 0.013 * idx)), 0.001, 1.0)


def cpu_baseline(model, x, tol, gpu_iterations):
    """refcpu on the host: one solve of the workload (assembly, numeric AMG setup
    with reused aggregates, PCG to tol), after one untimed first setup."""
    from oracle import refcpu

    o = refcpu.Oracle(model, tol=tol)
    fac = 1e-9 + (1 - 1e-9) * np.clip(o.filter(x), 0, 1) ** 3
    o.assemble_kff(fac)
    o.amg_setup()  # first setup: aggregation, untimed
    x1 = designs(x, 1)
    fac1 = 1e-9 + (1 - 1e-9) * np.clip(o.filter(x1), 0, 1) ** 3
    t0 = time.perf_counter()
    o.assemble_kff(fac1)
    o.amg_setup()
    b = model.free_loads()[0]
    _, its, _, _ = o.pcg(b, tol=tol, maxit=max(1, gpu_iterations))
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": refcpu.num_threads(), "kind": "port",
            "sample": f"refcpu (C++17/OpenMP restatement of the reference, Eigen absent) on the host: one solve of "
                      f"the same 128x64x64 workload = K_ff assembly + AMG numeric setup with reused aggregates + "
                      f"PCG from zero for {its} iterations (the GPU's per-solve count) to rtol {tol}; "
                      f"{dt:.2f} s, after one untimed first setup"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--size", type=str, default="128,64,64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference reading for the HBM byte counters)")
    args = ap.parse_args()
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

    from sparlab_amd import model as M
    from sparlab_amd import native as N

    nx, ny, nz = (int(v) for v in args.size.split(","))
    m = M.cantilever_3d(nx, ny, nz, volume_fraction=0.2, radius_cells=2.0)
    tol = 1e-8
    prec = 1 if args.precision == "mixed" else 0
    print(f"bench: building the problem {nx}x{ny}x{nz} ({m.nelem} elements)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    g = N.HipProblem(m, device=0)
    x0 = M.seeded_density(m.nelem)
    print(f"bench: model ready in {time.perf_counter() - t0:.1f} s ({g.nfree} free DOFs, K_ff nnz {g.kff_nnz})",
          file=sys.stderr, flush=True)
    # first setup (aggregation, patterns): part of building the solver, untimed
    t0 = time.perf_counter()
    ev = None
    g.assemble(1e-9 + (1 - 1e-9) * np.clip(g.filter(x0), 0, 1) ** 3)
    g.amg_setup(precision=prec)
    print(f"bench: first AMG setup {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    def step(k):
        return g.evaluate(designs(x0, k), tol=tol)

    for w in range(args.warmup):
        ev = step(w)
        print(f"bench: warmup {w}: c = {ev['compliance']:.6e}, {ev['iterations']} CG its", file=sys.stderr,
              flush=True)
    st = g.amg_stats()
    phases = {k: round(1e3 * getattr(st, k), 2) for k in ("setup_seconds", "check_seconds", "spectral_seconds",
                                                          "aggregation_seconds", "smoothing_seconds",
                                                          "galerkin_seconds", "coarse_seconds")}
    print(f"bench: AMG setup phases (ms) {phases}, levels {st.levels}, unknowns "
          f"{[st.unknowns[i] for i in range(st.levels)]}, node-product levels {st.node_product_levels}",
          file=sys.stderr, flush=True)
    g.ctx.timing(True)
    g.ctx.synchronize()
    its = []
    asm = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev = step(args.warmup + k)
        its.append(ev["iterations"])
        asm.append(ev["assemble_seconds"])
        print(f"bench: step {k}: c = {ev['compliance']:.6e}, {ev['iterations']} CG its", file=sys.stderr, flush=True)
    g.ctx.synchronize()
    elapsed = time.perf_counter() - t0
    spmv_ms, spmv_n = g.ctx.timer("spmv")
    g.ctx.timing(False)
    n = g.nfree
    nnz = g.kff_nnz
    spmv_bytes = 12 * nnz + 4 * (n + 1) + 16 * n
    avg_ms = spmv_ms / max(spmv_n, 1)
    achieved = spmv_bytes / (avg_ms * 1e-3) / 1e9 if spmv_n else 0.0
    value = args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if prec else "f64",
        "data": "synthetic",
        "config": {"workload": f"cantilever Hex8 {nx}x{ny}x{nz}, {n} free DOFs, rtol {tol}, "
                               f"{'FP32 V-cycle under FP64 flexible CG' if prec else 'FP64'}",
                   "elements": m.nelem, "free_dofs": n, "kff_nnz": nnz, "precision": args.precision,
                   "pcg_iterations_per_solve": float(np.mean(its)),
                   "setup_ms_per_step": 1e3 * float(np.mean(asm))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_spmv<16,3,double> (PCG q = K_ff p, FP64 CSR)",
                     "bytes_per_launch": spmv_bytes, "avg_ms": avg_ms, "launches": spmv_n},
    }
    if not args.no_cpu_baseline:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(m, x0, tol, int(round(np.mean(its))))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
