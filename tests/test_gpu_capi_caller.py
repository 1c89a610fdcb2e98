"""A compiled C program calls the C ABI as a C/C++ host would (tests/capi_caller.c):
the LinearSolver seam (slx_factorize_csr + slx_solve + slx_amg_stats_get) on the
sparlab_bench Hex8 box, checked against the published row of the reference
(docs/results/solver_scaling.csv:25, tests/golden/published.json), plus the error
statuses of out-of-order calls."""
import json
import os
import shutil
import subprocess

import pytest

from sparlab_amd import native as N

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "published.json")))


def test_c_caller_factorize_solve_and_stats(tmp_path):
    print("test_c_caller_factorize_solve_and_stats", flush=True)
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    libdir = os.path.dirname(N.LIB_PATH)
    exe = tmp_path / "capi_caller"
    subprocess.run(["gcc", "-std=c99", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "capi_caller.c"), "-o", str(exe), "-L", libdir, "-lsparlab_hip",
                    f"-Wl,-rpath,{libdir}", "-lm"], check=True)
    r = subprocess.run([str(exe), "16"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    out = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    assert "error" not in out, out.get("error")
    row = next(p for p in GOLDEN["hex8_bench_amg"] if p["nx"] == 16)
    assert int(out["free_dofs"]) == row["num_dofs"] - 3 * 9 * 5
    assert int(out["kff_nnz"]) == row["kff_nnz"]
    assert int(out["converged"]) == 1 and float(out["relative_residual"]) <= 1e-10
    c = float(out["compliance"])
    assert abs(c - row["compliance"]) <= 1e-11 * row["compliance"]
    assert abs(int(out["iterations"]) - row["iterations"]) <= 1
    assert int(out["levels"]) == row["levels"]
    assert float(out["operator_complexity"]) == pytest.approx(row["operator_complexity"], rel=1e-15)
    assert int(out["storage_nonzeros"]) == row["solver_nonzeros"]
    assert int(out["nonzeros0"]) == row["kff_nnz"]
    assert int(out["setup_phases_nonnegative"]) == 1
    # error statuses: SolverError before factorize; ModelError for the structured entry points once a
    # caller matrix replaced K_ff; the structured model works again after slx_model_finalize
    assert int(out["solve_before_factorize_status"]) == 2
    assert int(out["solve_before_factorize_message_nonempty"]) == 1
    assert int(out["solve_case_after_caller_matrix_status"]) == 4
    assert int(out["evaluate_after_caller_matrix_status"]) == 4
    assert int(out["evaluate_after_refinalize_status"]) == 0
    # uniform design 0.5 without a filter: c = c_solid / (Emin + (1 - Emin) 0.5^3)
    assert float(out["evaluate_compliance"]) == pytest.approx(row["compliance"] / (1e-9 + (1 - 1e-9) * 0.125),
                                                              rel=1e-9)
    assert int(out["done"]) == 1
