"""Command-line driver for SIMP topology optimisation on one GPU.

    python -m sparlab_amd.topopt --config deck.json [--output dir] [overrides]

The counterpart of the reference's ``sparlab_topopt`` (apps/sparlab_topopt.cpp)
for the decks this library runs (structured Q4 / Hex8, point loads, density
filter, OC with penalty continuation; see :mod:`sparlab_amd.deck`). The deck is
read unchanged; the linear solves run on the GPU's AMG-preconditioned CG
whatever ``solver.linear.type`` the deck names, to the deck's
``residual_tolerance`` unless ``--tolerance`` says otherwise.

Outputs in the output directory (src/io/ResultWriter.cpp):

* ``history.csv``: one row per iteration, the reference's columns
  (ResultWriter.cpp:1428-1470);
* ``density_final.csv``: per element centroid, volume, design, physical
  density, stiffness factor, passive tag and strain energy
  (ResultWriter.cpp:1474-1515);
* ``summary.json``: provenance, mesh, material, ``optimization_setup``,
  ``optimization_result`` (ResultWriter.cpp:2560-2810) and
  ``mass_stiffness_comparison`` (apps/sparlab_topopt.cpp:438-466,873-930)
  and, when the deck enables ``modal``, ``modal_initial_solid``,
  ``modal_mass_matched_baseline`` (2-D) and ``modal_optimised_topology`` with
  the interpretation of the final density field (apps/sparlab_topopt.cpp:490-640).
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import sys
import time

import numpy as np

from . import deck as D
from . import interpret as I

_OVERRIDES = (("volume-fraction", float), ("penalty", float), ("filter-radius", float),
              ("filter-radius-elements", float), ("filter-type", str), ("max-iterations", int), ("nx", int),
              ("ny", int), ("nz", int), ("youngs-modulus", float))
_PRECISION = {"fp64": 0, "mixed": 1, "bf16": 2}


def _fmt(v) -> str:
    """CsvWriter's 17 significant digits (include/sparlab/io/CsvWriter.hpp:21-22)."""
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return format(float(v), ".17g")


def write_csv(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f, lineterminator="\n")
        w.writerow(header)
        for r in rows:
            w.writerow([_fmt(v) for v in r])


def _num(v):
    """JSON has no infinity: the reference's writer emits null for it."""
    v = float(v)
    return v if math.isfinite(v) else None


def optimised_topology_modes(dk: D.Deck, phys, device, log) -> dict:
    """The final density thresholded and reduced to its largest face-connected
    group (src/mesh/SubMesh.cpp:172-260), and that solid's lowest modes
    (apps/sparlab_topopt.cpp:490-560,635-643). A structure the threshold cannot
    produce, or one no support holds, is reported instead of analysed."""
    from . import native as N
    m = dk.model
    out = {}
    try:
        interp = I.interpret_density_as_solid(m, phys, dk.interpretation_threshold)
    except I.InterpretationError as e:
        log(f"{dk.name}: the density field cannot be interpreted as a solid: {e}")
        out["interpreted_solid_analysis"] = {"analysis_failed": True, "reason": f"no interpreted structure: {e}"}
        return out
    out["topology_interpretation"] = interp.as_dict()
    block = {"num_elements": interp.elements_retained,
             "num_nodes": int(len(I.submesh_nodes(m, interp.retained))),
             "mass_kg": interp.volume_retained * dk.density, "loads_applied": False}
    out["interpreted_solid_analysis"] = block
    try:  # the deck's supports re-applied on the extracted structure (apps/sparlab_topopt.cpp:109)
        sub, used = I.solid_submodel(m, interp.retained, dk.constraints, dk.origin)
    except D.DeckError as e:
        block.update(analysis_failed=True, reason=str(e))
        log(f"{dk.name}: skipping the modal analysis of the optimised topology: {e}")
        return out
    ind = interp.retained.astype(float)
    try:
        gs = N.HipProblem(sub, device=device)
        modes = gs.modal(dk.density, stiffness_scale=ind, mass_scale=ind, **dk.modal)
    except N.SparlabError as e:
        block.update(analysis_failed=True, reason=str(e))
        log(f"{dk.name}: the interpreted structure could not be analysed: {e}")
        return out
    out["modal_optimised_topology"] = {k: v for k, v in modes.items() if k != "seconds"}
    log(f"{dk.name}: f1 topology {modes['frequencies_hz'][0]:.10g} Hz (mass {modes['total_mass_kg']:.10g} kg, "
        f"{interp.elements_retained} elements), {modes['seconds']:.2f} s")
    return out


def _multigrid_options(amg: dict) -> dict:
    """The multigrid options in effect, named as ResultWriter.cpp:1694-1707 names them."""
    from . import native as N

    o = N.amg_options(**amg)
    out = {k: getattr(o, k) for k in ("strength_threshold", "max_levels", "coarse_size", "smoother_degree",
                                      "chebyshev_ratio", "prolongator_damping", "lanczos_steps",
                                      "coarse_pivot_tolerance")}
    out["smoother"] = "gauss_seidel" if o.smoother == 1 else "chebyshev"
    out["reuse_aggregates"] = bool(o.reuse_aggregates)
    return out


def run(dk: D.Deck, output: str | None = None, device: int = 0, precision: str = "fp64",
        tolerance: float | None = None, tag: str = "", write_csv_files: bool = True, log=print) -> dict:
    """Optimise one deck on the GPU and write its outputs; returns the summary."""
    from . import native as N

    m = dk.model
    tol = dk.tolerance if tolerance is None else tolerance
    t_all = time.perf_counter()
    timings = {}  # the reference's TimingLedger keys (apps/sparlab_topopt.cpp:404-860)
    t0 = time.perf_counter()
    g = N.HipProblem(m, device=device)
    x0 = dk.initial_design()
    if precision != "fp64" or dk.amg:
        # the options (precision, the deck's solver.linear.amg) take effect at a
        # setup: run it on the first iteration's operator, where the loop's first
        # setup would aggregate anyway
        p0 = dk.run["penalty_start"] if dk.run["cont_steps"] > 1 else m.penalty
        g.assemble(m.emin_ratio + (1.0 - m.emin_ratio) * np.clip(g.filter(x0), 0.0, 1.0) ** p0)
        g.amg_setup(precision=_PRECISION[precision], **dk.amg)
    log(f"{dk.name}: {m.nelem} elements, {g.nfree} free DOFs, {len(dk.case_names)} load case(s), "
        f"filter radius {dk.filter_radius:.6g} m, precision {precision}")
    timings["model_build"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, hist, iters, comp = g.run_oc(x0, dk.run["max_iterations"], change_tol=dk.run["change_tol"],
                                    objective_tol=dk.run["objective_tol"], window=dk.run["window"],
                                    cont_steps=dk.run["cont_steps"], penalty_start=dk.run["penalty_start"],
                                    cont_iterations=dk.run["cont_iterations"], move=dk.run["move"],
                                    damping=dk.run["damping"], tol=tol, projection=dk.run["projection"],
                                    warm_start=dk.run.get("warm_start", True))
    timings["optimization"] = time.perf_counter() - t0
    res = dict(g.last_run)
    case_c = g.load_case_compliance()
    st = g.amg_stats()
    log(f"{dk.name}: {iters} iterations ({res['stop_reason']}), compliance {comp:.10g} J, "
        f"volume fraction {res['volume_fraction']:.10g}, {res['total_seconds']:.2f} s")

    # the final design's fields for density_final.csv (the run's last evaluation, repeated)
    pj = dk.run["projection"]
    robust = bool(pj and pj["robust"])
    beta = res["final_beta"] if pj else 0.0
    eta = pj["eta"] if pj else 0.5
    fin = g.evaluate(x, tol=tol, need_gradients=False, beta=beta, eta=eta)
    filtered = np.clip(g.filter(x), 0.0, 1.0)
    phys = filtered
    if pj:
        # the projected physical density (src/topopt/Projection.cpp:58-63)
        a = math.tanh(beta * eta)
        phys = (a + np.tanh(beta * (filtered - eta))) / (a + math.tanh(beta * (1.0 - eta)))
    factor = m.emin_ratio + phys ** m.penalty * (1.0 - m.emin_ratio)
    # the full solid domain (apps/sparlab_topopt.cpp:438-466)
    t0 = time.perf_counter()
    solid = g.evaluate(np.ones(m.nelem), tol=tol, need_gradients=False, warm_start=False)["compliance"]
    timings["solid_reference"] = time.perf_counter() - t0
    vf = res["volume_fraction"]

    wsum = sum(dk.case_weights)
    vol = dk.domain_volume
    target = m.volume_fraction * vol
    summary = {
        "provenance": {
            "code": "sparlab_amd", "config_source": dk.source, "case": dk.name + (("_" + tag) if tag else ""),
            "description": dk.description, "units": "SI: m, N, Pa, kg, kg/m^3, Hz, J",
            "timestamp_utc": time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime()),
            "device": f"HIP device {device} (gfx950)",
        },
        "mesh": {
            "source": "structured_quad" if m.dim == 2 else "structured_hex",
            "element_type": "Quad4" if m.dim == 2 else "Hex8", "dim": m.dim, "num_nodes": m.nnodes,
            "num_elements": m.nelem, "num_dofs": m.ndofs, "num_free_dofs": g.nfree,
            "num_prescribed_dofs": int(m.fixed.sum()), "thickness_m": m.thickness if m.dim == 2 else 1.0,
            "domain_volume_m3": vol,
            "structured_grid": {"nx": m.nx, "ny": m.ny, **({"nz": m.nz} if m.dim == 3 else {}), "lx_m": m.lx,
                                "ly_m": m.ly, **({"lz_m": m.lz} if m.dim == 3 else {})},
        },
        "material": {"youngs_modulus_Pa": m.E, "poisson_ratio": m.nu},
        "optimization_setup": {
            "load_case_names": dk.case_names, "load_case_weights": dk.case_weights,
            "load_case_weights_normalised": [w / wsum for w in dk.case_weights],
            "volume_fraction_target": m.volume_fraction, "volume_target_m3": target,
            "initial_density": dk.initial_density, "domain_volume_m3": vol, "num_design_variables": m.nelem,
            "num_free_variables": m.nelem - dk.passive_solid - dk.passive_void,
            "num_passive_solid": dk.passive_solid, "num_passive_void": dk.passive_void,
            "filter_type": dk.filter_type, "filter_radius_m": dk.filter_radius, "simp_penalty": m.penalty,
            "simp_emin_ratio": m.emin_ratio, "move_limit": dk.run["move"], "damping": dk.run["damping"],
            "continuation_steps": dk.run["cont_steps"], "penalty_start": dk.run["penalty_start"],
            "continuation_iterations": dk.run["cont_iterations"], "max_iterations": dk.run["max_iterations"],
            "change_tolerance": dk.run["change_tol"], "objective_tolerance": dk.run["objective_tol"],
            "objective_window": dk.run["window"], "method": "oc",
        },
        "optimization_result": {
            "converged": bool(res["converged"]), "stop_reason": res["stop_reason"],
            "final_design_change": _num(res["final_design_change"]),
            "final_relative_objective_change": _num(res["final_objective_change"]),
            "iterations": iters, "linear_solves": res["linear_solves"],
            "linear_solver": {
                "solver": "amg_cg (HIP, gfx950)", "precision": precision, "residual_tolerance": tol,
                "multigrid_options": _multigrid_options(dk.amg),
                "iterative_iterations_total": res["linear_iterations"],
                "iterative_iterations_per_solve": res["linear_iterations"] / max(res["linear_solves"], 1),
                "multigrid": {
                    "levels": st.levels,
                    "level_sizes": [{"unknowns": int(st.unknowns[i]), "nonzeros": int(st.nonzeros[i]),
                                     "lambda_max_Dinv_A": st.lambda_max[i]} for i in range(st.levels)],
                    "operator_complexity": st.operator_complexity, "grid_complexity": st.grid_complexity,
                    "near_null_space_dimension": st.near_null_space_dimension,
                    "coarse_pivot_ratio": st.coarse_pivot_ratio,
                    "last_setup_seconds": st.setup_seconds,
                    "aggregates_reused_in_last_setup": bool(st.reused_aggregates),
                },
            },
            "compliance_J": comp, "load_case_compliance_J": [float(c) for c in case_c],
            "volume_m3": res["volume"], "volume_fraction": vf,
            "volume_constraint_relative_violation": (res["volume"] - target) / target,
            "volume_constraint_satisfied": (res["volume"] - target) / target <= 1.0e-6,
            "grey_level": res["grey_level"], "total_seconds": res["total_seconds"],
            "seconds_per_iteration": res["total_seconds"] / iters if iters else 0.0, "warnings": [],
        },
        "mass_stiffness_comparison": {
            "full_solid_compliance_J": solid, "optimised_compliance_J": comp,
            "compliance_penalty_vs_full_solid": comp / solid if solid > 0 else 0.0,
        },
    }
    if pj:
        # ResultWriter.cpp:2640-2673 (setup) and 2753-2764 (result)
        summary["optimization_setup"]["projection"] = dict(pj, enabled=True)
        summary["optimization_result"]["projection"] = {
            "final_beta": beta, "eta": eta,
            "filtered_grey_level": float(4.0 * np.sum(filtered * (1.0 - filtered)) / m.nelem),
            "note": "grey_level is that of the projected (physical) density; filtered_grey_level that of the "
                    "density before projection"}
        if robust or pj["erosion_check"]:
            # RobustRecord (ResultWriter.cpp:2766-2800): the three designs at eta + delta, eta, eta - delta
            d = pj["robust_delta"]
            if robust:
                ce, ve, cd, vd = (res["compliance_eroded"], res["volume_fraction_eroded"], res["compliance_dilated"],
                                  res["volume_fraction_dilated"])
            else:
                # the erosion check of a non-robust design (TopologyOptimizer.cpp:396-420)
                er = g.evaluate(x, tol=tol, need_gradients=False, beta=beta, eta=eta + d)
                di = g.evaluate(x, tol=tol, need_gradients=False, beta=beta, eta=eta - d)
                ce, ve, cd, vd = er["compliance"], er["volume_fraction"], di["compliance"], di["volume_fraction"]
            block = {"designs": [
                {"design": "eroded", "eta": eta + d, "compliance_J": ce, "volume_fraction": ve},
                {"design": "intermediate (blueprint)", "eta": eta, "compliance_J": comp, "volume_fraction": vf},
                {"design": "dilated", "eta": eta - d, "compliance_J": cd, "volume_fraction": vd}]}
            if robust:
                block["dilated_target_fraction"] = res["dilated_target_fraction"]
                block["note"] = ("compliance_J and volume_fraction below are the blueprint's; the optimiser "
                                 "minimised the eroded design's compliance")
                summary["optimization_result"]["robust"] = block
            else:
                block["note"] = ("a design optimised without the robust formulation, evaluated once at the eroded and "
                                 "dilated thresholds eta +- robust_delta of the final beta: its compliance if the part "
                                 "came out uniformly thinner or thicker than drawn")
                summary["optimization_result"]["erosion_check"] = block
    for key in dk.skipped_reports:
        summary["optimization_result"]["warnings"].append(f"{key} is a report on the final design that this driver "
                                                          "does not compute")
    if len(hist):
        summary["optimization_result"]["initial_compliance_J"] = float(hist[0, N.H_COMPLIANCE])
        summary["optimization_result"]["compliance_reduction_factor"] = (
            float(hist[0, N.H_COMPLIANCE]) / comp if comp > 0 else 0.0)
    if m.dim == 2:
        plate = solid / vf if vf > 0 else 0.0
        summary["mass_stiffness_comparison"]["equal_mass_uniform_plate_compliance_J"] = plate
        summary["mass_stiffness_comparison"]["stiffness_gain_over_equal_mass_plate"] = plate / comp if comp > 0 else 0.0
    if dk.modal is not None:
        t0 = time.perf_counter()
        # the full solid domain's lowest modes (apps/sparlab_topopt.cpp:612-620)
        solid_modes = g.modal(dk.density, **dk.modal)
        summary["modal_initial_solid"] = {k: v for k, v in solid_modes.items() if k != "seconds"}
        log(f"{dk.name}: f1 solid {solid_modes['frequencies_hz'][0]:.10g} Hz (mass {solid_modes['total_mass_kg']:.10g} "
            f"kg), {solid_modes['iterations']} subspace iterations, {solid_modes['linear_iterations']} CG iterations, "
            f"{solid_modes['seconds']:.2f} s")
        if m.dim == 2 and dk.modal_flags["compare_mass_matched_baseline"]:
            # the equal-mass uniform plate (thickness times the achieved volume fraction,
            # apps/sparlab_topopt.cpp:622-633): K and M both scale with the thickness
            thin = g.modal(dk.density, stiffness_scale=np.full(m.nelem, vf), mass_scale=np.full(m.nelem, vf),
                           **dk.modal)
            summary["modal_mass_matched_baseline"] = {k: v for k, v in thin.items() if k != "seconds"}
        if dk.modal_flags["analyse_optimised_topology"]:
            topo = optimised_topology_modes(dk, phys, device, log)
            summary.update(topo)
        timings["modal_analysis"] = time.perf_counter() - t0

    t_out = time.perf_counter()
    if output:
        os.makedirs(output, exist_ok=True)
        if write_csv_files:
            # beta only for a projected run, the robust columns only for a robust one
            extra = ([N.H_BETA] if pj else []) + ([N.H_ERODED_VF, N.H_DILATED_VF, N.H_DILATED_TARGET] if robust else [])
            write_csv(os.path.join(output, "history.csv"), list(N.HISTORY_COLUMNS[:N.H_BETA]) +
                      [N.HISTORY_COLUMNS[i] for i in extra],
                      [([int(r[0])] + list(r[1:N.H_BISECT]) + [int(r[N.H_BISECT]), int(r[N.H_VCONV])]
                        + [r[N.H_SECONDS], int(r[N.H_LINEAR])] + [r[i] for i in extra]) for r in hist])
            cents = D.Geometry(m, dk.origin).cents[:, : m.dim]
            ev = vol / m.nelem
            tags = np.zeros(m.nelem, dtype=int)
            pas = m.passive.astype(bool)
            tags[pas & (m.lower >= 1.0)] = 1
            tags[pas & (m.lower < 1.0)] = 2
            header = ["element"] + ["cx[m]", "cy[m]", "cz[m]"][: m.dim] + (["area[m2]"] if m.dim == 2 else []) + [
                "volume[m3]", "design_x[-]", "physical_density[-]", "stiffness_factor[-]", "passive_tag[-]",
                "strain_energy[J]"]
            area = m.lx * m.ly / m.nelem
            rows = []
            for e in range(m.nelem):
                rows.append([e] + list(cents[e]) + ([area] if m.dim == 2 else []) +
                            [ev, x[e], phys[e], factor[e], int(tags[e]), fin["energy"][e]])
            write_csv(os.path.join(output, "density_final.csv"), header, rows)
        timings["output"] = time.perf_counter() - t_out
        timings["total"] = time.perf_counter() - t_all
        summary["timings_s"] = dict(timings)
        with open(os.path.join(output, "summary.json"), "w") as f:
            json.dump(summary, f, indent=2)
        log(f"{dk.name}: wrote {output}")
    timings["total"] = time.perf_counter() - t_all
    summary["timings_s"] = dict(timings)
    summary["_seconds_wall"] = timings["total"]
    return summary


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m sparlab_amd.topopt",
                                 description="SIMP topology optimisation of a SparLab deck on one GPU")
    ap.add_argument("--config", required=True, help="input deck with a topology section")
    ap.add_argument("--output", help="output directory (default results/<case>)")
    for name, typ in _OVERRIDES:
        ap.add_argument("--" + name, type=typ)
    ap.add_argument("--load-weights", help="override the per-load-case weights, w1,w2,...")
    ap.add_argument("--tag", default="", help="suffix appended to the case name in summary.json")
    ap.add_argument("--no-csv", action="store_true", help="skip the CSV outputs")
    ap.add_argument("--no-vtk", action="store_true", help="accepted for compatibility; no VTK is written")
    ap.add_argument("--solver", help="accepted for compatibility; the GPU solver is always amg_cg")
    ap.add_argument("--precision", choices=sorted(_PRECISION), default="fp64",
                    help="multigrid hierarchy precision (fp64: the drop-in path)")
    ap.add_argument("--tolerance", type=float, help="relative residual of the CG solves (default: the deck's)")
    ap.add_argument("--device", type=int, default=0)
    # apps/sparlab_topopt.cpp:165-180,290-330
    ap.add_argument("--modes", type=int, help="enable modal analysis with n modes")
    ap.add_argument("--projection", action="store_true", help="switch the Heaviside projection on")
    ap.add_argument("--no-projection", action="store_true", help="switch the deck's projection off")
    ap.add_argument("--beta-max", type=float, help="override topology.projection.beta_max")
    ap.add_argument("--robust", action="store_true", help="robust (eroded / dilated) formulation, with the projection")
    ap.add_argument("--no-robust", action="store_true", help="switch the deck's robust formulation off")
    a = ap.parse_args(argv)
    ov = {}
    for name, _ in _OVERRIDES:
        v = getattr(a, name.replace("-", "_"))
        if v is not None:
            ov[name.replace("-", "_")] = v
    if a.load_weights:
        ov["load_weights"] = [float(w) for w in a.load_weights.split(",")]
    if a.modes is not None:
        ov["modes"] = a.modes
    if a.projection or a.robust:
        ov["projection"] = True
    if a.no_projection:
        ov["projection"] = False
    if a.beta_max is not None:
        ov["beta_max"] = a.beta_max
    if a.robust:
        ov["robust"] = True
    if a.no_robust:
        ov["robust"] = False
    if a.solver and a.solver != "amg_cg":
        print(f"note: solver.linear.type {a.solver} is replaced by the GPU's amg_cg", file=sys.stderr)
    try:
        dk = D.load_deck(a.config, overrides=ov)
    except D.DeckError as e:
        print(f"configuration error: {e}", file=sys.stderr)
        return 2
    out = a.output or os.path.join("results", dk.name + (("_" + a.tag) if a.tag else ""))
    run(dk, out, device=a.device, precision=a.precision, tolerance=a.tolerance, tag=a.tag,
        write_csv_files=not a.no_csv, log=lambda s: print(s, file=sys.stderr, flush=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
