"""The deck driver on the GPU against the reference's published benchmark table.

Each of the five benchmark decks (tests/golden/decks) is optimised end to end
by sparlab_amd.topopt -- deck reader, GPU assembly, AMG-PCG, sensitivities,
filter, OC with continuation, outputs -- and compared with the reference's
published row (docs/results/benchmarks.csv, solved there with a direct
solver): iteration count, stop reason, final compliance, achieved volume
fraction, grey level, initial and full-solid compliance.
"""
import csv
import json
import os

import numpy as np
import pytest

from sparlab_amd import deck as D

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = ("mbb_beam", "cantilever_beam", "bracket_3d", "aerospace_bracket", "wing_rib", "mbb_beam_projected",
         "bracket_3d_projected", "mbb_beam_robust", "bracket_3d_large")
PUBLISHED = json.load(open(os.path.join(GOLD, "benchmarks_published.json")))["cases"]


def deck(case):
    return D.load_deck(os.path.join(GOLD, "decks", case + ".json"))


@pytest.mark.parametrize("case", CASES)
def test_initial_and_solid_compliance_on_gpu(case):
    from sparlab_amd import native as N
    d = deck(case)
    m = d.model
    pub = PUBLISHED[case]
    print(f"deck {case}: initial and full-solid compliance", flush=True)
    g = N.HipProblem(m, device=0)
    solid = g.evaluate(np.ones(m.nelem), tol=1e-11, need_gradients=False, warm_start=False)["compliance"]
    assert solid == pytest.approx(pub["full_solid_compliance_J"], rel=1e-8)
    p0 = d.run["penalty_start"] if d.run["cont_steps"] > 1 else m.penalty
    pj = d.run["projection"]
    # the first history row: the eroded design (eta + delta) in a robust run
    eta0 = (pj["eta"] + (pj["robust_delta"] if pj["robust"] else 0.0)) if pj else 0.5
    c0 = g.evaluate(d.initial_design(), tol=1e-11, need_gradients=False, warm_start=False, penalty=p0,
                    beta=pj["beta_start"] if pj else 0.0, eta=eta0)["compliance"]
    assert c0 == pytest.approx(pub["initial_compliance_J"], rel=1e-8)


@pytest.mark.parametrize("case", CASES)
def test_deck_run_matches_the_published_benchmark(case, tmp_path):
    from sparlab_amd import native as N
    from sparlab_amd.topopt import run
    d = deck(case)
    pub = PUBLISHED[case]
    print(f"deck {case}: full optimisation ({pub['iterations']} iterations published)", flush=True)
    s = run(d, str(tmp_path), log=lambda t: print(t, flush=True))
    r = s["optimization_result"]
    print(f"deck {case}: {r['iterations']} iterations, {r['stop_reason']}, compliance {r['compliance_J']!r} "
          f"(published {pub['compliance_J']!r}), {r['total_seconds']:.2f} s", flush=True)
    assert r["stop_reason"] == pub["stop_reason"]
    assert r["iterations"] == pub["iterations"]
    assert r["compliance_J"] == pytest.approx(pub["compliance_J"], rel=1e-8)
    assert r["volume_fraction"] == pytest.approx(pub["volume_fraction_achieved"], abs=1e-8)
    assert r["grey_level"] == pytest.approx(pub["grey_level"], rel=1e-5)
    assert r["initial_compliance_J"] == pytest.approx(pub["initial_compliance_J"], rel=1e-7)
    c = s["mass_stiffness_comparison"]
    assert c["full_solid_compliance_J"] == pytest.approx(pub["full_solid_compliance_J"], rel=1e-7)
    # the weighted per-case compliances add up to the objective
    w = np.array(s["optimization_setup"]["load_case_weights_normalised"])
    assert float(w @ np.array(r["load_case_compliance_J"])) == pytest.approx(r["compliance_J"], rel=1e-12)

    # outputs
    rows = list(csv.reader(open(tmp_path / "history.csv")))
    pj = d.run["projection"]
    robust = bool(pj and pj["robust"])
    assert rows[0] == list(N.HISTORY_COLUMNS[: N.H_BETA + (1 if pj else 0)]) + (
        list(N.HISTORY_COLUMNS[N.H_ERODED_VF:]) if robust else [])
    if robust:
        des = r["robust"]["designs"]
        assert des[0]["compliance_J"] > des[1]["compliance_J"] > des[2]["compliance_J"]
        assert des[0]["volume_fraction"] < des[1]["volume_fraction"] < des[2]["volume_fraction"]
        assert des[1]["compliance_J"] == r["compliance_J"]
    if pj:
        assert r["projection"]["final_beta"] == pub["final_beta"]
        assert float(rows[-1][rows[0].index("beta[-]")]) == pub["final_beta"]
    assert len(rows) == 1 + r["iterations"]
    assert [int(x[0]) for x in rows[1:]] == list(range(1, r["iterations"] + 1))
    assert float(rows[1][2]) == r["initial_compliance_J"]
    dens = list(csv.reader(open(tmp_path / "density_final.csv")))
    assert len(dens) == 1 + d.model.nelem
    vol_col = dens[0].index("volume[m3]")
    rho_col = dens[0].index("physical_density[-]")
    v = sum(float(x[vol_col]) * float(x[rho_col]) for x in dens[1:])
    assert v == pytest.approx(r["volume_m3"], rel=1e-9)
    saved = json.load(open(tmp_path / "summary.json"))
    assert saved["optimization_result"]["stop_reason"] == r["stop_reason"]
    # the reference's timing ledger keys and the multigrid block's coarse pivot ratio
    # (ResultWriter.cpp:1664-1683,2917)
    t = saved["timings_s"]
    assert {"model_build", "optimization", "solid_reference", "output", "total"} <= set(t)
    assert 0.0 < t["optimization"] <= t["total"]
    mg = saved["optimization_result"]["linear_solver"]["multigrid"]
    assert 0.0 < mg["coarse_pivot_ratio"] <= 1.0
    # the full solid domain's modes (published f1_solid_Hz, mass_solid_kg)
    if d.modal is not None:
        modes = saved["modal_initial_solid"]
        # the eigenvalue-change tolerance of the deck bounds the agreement (1e-10 by default, 1e-8 in some decks);
        # the published masses carry the reference's summation rounding (1e-11 over 110 k elements)
        ftol = max(1e-9, d.modal["tolerance"])
        assert modes["frequencies_hz"][0] == pytest.approx(pub["f1_solid_Hz"], rel=ftol)
        assert modes["total_mass_kg"] == pytest.approx(pub["mass_solid_kg"], rel=1e-10)
        if d.model.dim == 2 and d.modal_flags["compare_mass_matched_baseline"]:
            thin = saved["modal_mass_matched_baseline"]
            assert thin["frequencies_hz"] == pytest.approx(modes["frequencies_hz"], rel=1e-8)
            assert thin["total_mass_kg"] == pytest.approx(modes["total_mass_kg"] * r["volume_fraction"], rel=1e-12)
        if d.modal_flags["analyse_optimised_topology"]:
            # the thresholded, largest face-connected structure (published f1_topology_Hz, mass_topology_kg)
            topo = saved["modal_optimised_topology"]
            print(f"deck {case}: f1 topology {topo['frequencies_hz'][0]!r} Hz (published {pub['f1_topology_Hz']!r}), "
                  f"mass {topo['total_mass_kg']!r} kg (published {pub['mass_topology_kg']!r})", flush=True)
            assert topo["total_mass_kg"] == pytest.approx(pub["mass_topology_kg"], rel=1e-10)
            assert saved["interpreted_solid_analysis"]["mass_kg"] == pytest.approx(pub["mass_topology_kg"], rel=1e-10)
            assert topo["frequencies_hz"][0] == pytest.approx(pub["f1_topology_Hz"], rel=ftol)
    else:
        assert "modal_initial_solid" not in saved


def test_cli_switches_erosion_check_and_modes(tmp_path):
    """--no-robust leaves the deck's erosion check, which reports the design at eta +- delta
    (TopologyOptimizer.cpp:396-420); --modes turns the modal analysis on."""
    from sparlab_amd.topopt import run
    d = D.load_deck(os.path.join(GOLD, "decks", "mbb_beam_robust.json"),
                    overrides=dict(robust=False, max_iterations=30, modes=2))
    print("deck mbb_beam_robust --no-robust --max-iterations 30 --modes 2", flush=True)
    s = run(d, str(tmp_path), log=lambda t: print(t, flush=True))
    r = s["optimization_result"]
    des = r["erosion_check"]["designs"]
    assert des[0]["eta"] == pytest.approx(0.6) and des[2]["eta"] == pytest.approx(0.4)
    assert des[0]["compliance_J"] > des[1]["compliance_J"] > des[2]["compliance_J"]
    assert des[0]["volume_fraction"] < des[1]["volume_fraction"] < des[2]["volume_fraction"]
    assert des[1]["compliance_J"] == r["compliance_J"]
    assert "robust" not in r
    modes = s["modal_initial_solid"]
    assert modes["num_modes"] == 2 and modes["converged"]


def test_deck_with_the_gauss_seidel_smoother_against_the_oracle(tmp_path):
    """A deck whose solver.linear.amg block selects the symmetric Gauss-Seidel
    smoother (and a small coarse size, so that the sweeps run on two levels):
    the driver's run against the oracle's OC loop with the same options."""
    from oracle.refcpu import Oracle
    from sparlab_amd.topopt import run
    doc = json.load(open(os.path.join(GOLD, "decks", "mbb_beam.json")))
    doc.setdefault("solver", {}).setdefault("linear", {})["amg"] = {"smoother": "gauss_seidel", "coarse_size": 100}
    d = D.parse_deck(doc, overrides=dict(nx=60, ny=20, max_iterations=12))
    print("deck mbb_beam 60 x 20, gauss_seidel, 12 iterations", flush=True)
    s = run(d, str(tmp_path), log=lambda t: print(t, flush=True))
    r = s["optimization_result"]
    assert r["linear_solver"]["multigrid_options"]["smoother"] == "gauss_seidel"
    assert r["linear_solver"]["multigrid"]["levels"] >= 3
    o = Oracle(d.model, tol=d.tolerance, warm_start=d.run["warm_start"], amg=dict(smoother=1, coarse_size=100))
    kw = {k: v for k, v in d.run.items() if k not in ("max_iterations", "warm_start")}
    xo, ho, no, fo = o.run_oc(d.initial_design(), d.run["max_iterations"], **kw)
    rows = list(csv.reader(open(tmp_path / "history.csv")))[1:]
    assert r["iterations"] == no
    np.testing.assert_allclose([float(x[2]) for x in rows], ho[:no, 2], rtol=1e-8)
    assert r["compliance_J"] == pytest.approx(fo, rel=1e-7)
