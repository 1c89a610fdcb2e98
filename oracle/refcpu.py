"""ctypes wrapper of the refcpu oracle library.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librefcpu.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def heaviside_project(r, beta, eta):
    return lib().ref_heaviside_project(float(r), float(beta), float(eta))


def heaviside_derivative(r, beta, eta):
    return lib().ref_heaviside_derivative(float(r), float(beta), float(eta))


def projection_stages(beta_start, beta_max, beta_factor, k):
    """(num_stages, beta_of_stage(k)) of the beta continuation."""
    b = C.c_double()
    n = lib().ref_projection_stages(beta_start, beta_max, beta_factor, k, C.byref(b))
    return n, b.value


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "-s"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    v = C.c_void_p
    sig = {
        "ref_last_error": (C.c_char_p, []),
        "ref_problem_new": (v, [C.c_int] * 4 + [C.c_double] * 6),
        "ref_problem_free": (None, [v]),
        "ref_problem_nnodes": (C.c_int, [v]),
        "ref_problem_nelem": (C.c_int, [v]),
        "ref_problem_ndofs": (C.c_int, [v]),
        "ref_problem_nfree": (C.c_int, [v]),
        "ref_problem_set_fixed": (None, [v, _up]),
        "ref_problem_set_loads": (None, [v, C.c_int, _dp, _dp]),
        "ref_problem_set_simp": (None, [v, C.c_double, C.c_double]),
        "ref_problem_set_solver": (None, [v, C.c_double, C.c_int, C.c_int]),
        "ref_problem_set_amg": (None, [v, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_int, C.c_int]),
        "ref_problem_set_design": (None, [v, C.c_double, _dp, _dp, _up]),
        "ref_problem_finalize": (C.c_int, [v, C.c_double]),
        "ref_mean_element_size": (C.c_double, [v]),
        "ref_get_ke0": (None, [v, _dp]),
        "ref_get_free": (None, [v, _ip]),
        "ref_get_volumes": (None, [v, _dp]),
        "ref_pattern_nnz": (C.c_longlong, [v]),
        "ref_get_pattern": (None, [v, _ip, _ip, C.c_void_p]),
        "ref_assemble_full": (None, [v, C.c_void_p, _dp]),
        "ref_assemble_kff": (C.c_longlong, [v, C.c_void_p]),
        "ref_get_kff": (None, [v, _ip, _ip, _dp]),
        "ref_set_kff": (None, [v, C.c_int, _ip, _ip, _dp]),
        "ref_amg_setup": (C.c_int, [v]),
        "ref_amg_nlevels": (C.c_int, [v]),
        "ref_amg_reused": (C.c_int, [v]),
        "ref_amg_operator_complexity": (C.c_double, [v]),
        "ref_amg_coarse_pivot_ratio": (C.c_double, [v]),
        "ref_amg_storage_nonzeros": (C.c_longlong, [v]),
        "ref_last_evaluate_seconds": (None, [v, _dp]),
        "ref_set_num_threads": (None, [C.c_int]),
        "ref_amg_level_info": (None, [v, C.c_int, C.c_int, _ip]),
        "ref_amg_level_get": (None, [v, C.c_int, C.c_int, _ip, _ip, _dp]),
        "ref_amg_lambda": (C.c_double, [v, C.c_int]),
        "ref_amg_nullspace_dim": (C.c_int, [v]),
        "ref_amg_nullspace": (None, [v, C.c_int, _dp]),
        "ref_amg_num_nodes": (C.c_int, [v, C.c_int]),
        "ref_amg_aggregates": (None, [v, C.c_int, _ip]),
        "ref_amg_node_ptr": (None, [v, C.c_int, _ip]),
        "ref_amg_apply": (C.c_int, [v, _dp, _dp]),
        "ref_pcg": (C.c_int, [v, _dp, _dp, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                              C.POINTER(C.c_int)]),
        "ref_filter_nnz": (C.c_longlong, [v]),
        "ref_filter_get": (None, [v, _ip, _ip, _dp, C.c_void_p]),
        "ref_filter_apply": (None, [v, C.c_int, _dp, _dp]),
        "ref_evaluate": (C.c_int, [v, _dp, C.c_int, _dp, C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
        "ref_oc": (C.c_int, [v, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_int)]),
        "ref_run_oc": (C.c_int, [v, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int), C.POINTER(C.c_double), _dp]),
        "ref_problem_set_robust": (None, [v, C.c_int, C.c_double, C.c_int]),
        "ref_problem_set_projection": (None, [v, C.c_double, C.c_double]),
        "ref_problem_set_projection_schedule": (None, [v, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                                       C.c_int, C.c_int]),
        "ref_heaviside_project": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "ref_heaviside_derivative": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "ref_projection_stages": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]),
        "ref_chunked_dot": (C.c_double, [_dp, _dp, C.c_int]),
        "ref_cpqr": (C.c_int, [_dp, C.c_int, C.c_int, _dp, _dp]),
        "ref_tridiag_max": (C.c_double, [_dp, _dp, C.c_int]),
        "ref_num_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().ref_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One structured SIMP problem in the oracle (mirrors the reference's
    FemModel + Assembler + AmgCgSolver + ComplianceObjective state)."""

    def __init__(self, model, tol=1e-8, max_iterations=1000, warm_start=True, amg=None):
        L = lib()
        self.model = model
        h = L.ref_problem_new(model.dim, model.nx, model.ny, max(model.nz, 1), model.lx, model.ly, model.lz,
                              model.E, model.nu, model.thickness)
        if not h:
            raise OracleError(1, L.ref_last_error().decode())
        self.h = C.c_void_p(h)
        L.ref_problem_set_fixed(self.h, np.ascontiguousarray(model.fixed, dtype=np.uint8))
        if model.loads is not None:
            L.ref_problem_set_loads(self.h, model.loads.shape[0], np.ascontiguousarray(model.loads, dtype=np.float64),
                                    np.ascontiguousarray(model.weights, dtype=np.float64))
        L.ref_problem_set_simp(self.h, model.penalty, model.emin_ratio)
        L.ref_problem_set_solver(self.h, tol, max_iterations, 1 if warm_start else 0)
        a = dict(theta=0.02, max_levels=12, coarse_size=1500, smoother=0, degree=3, cheb_ratio=30.0,
                 damping=4.0 / 3.0, lanczos=12, reuse=1)
        a.update(amg or {})
        L.ref_problem_set_amg(self.h, a["theta"], a["max_levels"], a["coarse_size"], a["smoother"], a["degree"],
                              a["cheb_ratio"], a["damping"], a["lanczos"], a["reuse"])
        L.ref_problem_set_design(self.h, model.volume_fraction, np.ascontiguousarray(model.lower, dtype=np.float64),
                                 np.ascontiguousarray(model.upper, dtype=np.float64),
                                 np.ascontiguousarray(model.passive, dtype=np.uint8))
        _check(L.ref_problem_finalize(self.h, model.filter_radius))
        self.ne = L.ref_problem_nelem(self.h)
        self.ndofs = L.ref_problem_ndofs(self.h)
        self.nfree = L.ref_problem_nfree(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_problem_free(self.h)
            self.h = None

    # ---- structure -------------------------------------------------------
    def ke0(self):
        nd = (8 if self.model.dim == 3 else 4) * self.model.dim
        k = np.empty(nd * nd)
        lib().ref_get_ke0(self.h, k)
        return k.reshape(nd, nd)

    def free(self):
        f = np.empty(self.nfree, dtype=np.int32)
        lib().ref_get_free(self.h, f)
        return f

    def volumes(self):
        v = np.empty(self.ne)
        lib().ref_get_volumes(self.h, v)
        return v

    def pattern(self, with_offsets=False):
        L = lib()
        nnz = L.ref_pattern_nnz(self.h)
        outer = np.empty(self.ndofs + 1, dtype=np.int32)
        inner = np.empty(nnz, dtype=np.int32)
        npe = 8 if self.model.dim == 3 else 4
        bo = np.empty(self.ne * npe * npe, dtype=np.int32) if with_offsets else None
        L.ref_get_pattern(self.h, outer, inner, _ptr(bo))
        return (outer, inner, bo) if with_offsets else (outer, inner)

    def assemble_full(self, scale=None):
        nnz = lib().ref_pattern_nnz(self.h)
        v = np.empty(nnz)
        s = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
        lib().ref_assemble_full(self.h, _ptr(s), v)
        return v

    def assemble_kff(self, scale=None):
        """Assemble K_ff (CSR arrays) and make it the current matrix."""
        s = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
        nnz = lib().ref_assemble_kff(self.h, _ptr(s))
        if nnz < 0:
            raise OracleError(4, lib().ref_last_error().decode())
        ptr = np.empty(self.nfree + 1, dtype=np.int32)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz)
        lib().ref_get_kff(self.h, ptr, col, val)
        return ptr, col, val

    def set_kff(self, ptr, col, val):
        lib().ref_set_kff(self.h, len(ptr) - 1, np.ascontiguousarray(ptr, dtype=np.int32),
                          np.ascontiguousarray(col, dtype=np.int32), np.ascontiguousarray(val, dtype=np.float64))

    # ---- AMG ----------------------------------------------------------------
    def amg_setup(self):
        _check(lib().ref_amg_setup(self.h))

    def amg_levels(self):
        return lib().ref_amg_nlevels(self.h)

    def amg_reused(self):
        return bool(lib().ref_amg_reused(self.h))

    def amg_operator_complexity(self):
        return lib().ref_amg_operator_complexity(self.h)

    def amg_storage_nonzeros(self):
        return lib().ref_amg_storage_nonzeros(self.h)

    def amg_matrix(self, level, which):
        """which: 'A', 'P', 'Ptent', 'R' -> (rows, cols, ptr, col, val)."""
        w = {"A": 0, "P": 1, "Ptent": 2, "R": 3}[which]
        info = np.empty(3, dtype=np.int32)
        lib().ref_amg_level_info(self.h, level, w, info)
        rows, cols, nnz = (int(t) for t in info)
        ptr = np.empty(rows + 1, dtype=np.int32)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz)
        lib().ref_amg_level_get(self.h, level, w, ptr, col, val)
        return rows, cols, ptr, col, val

    def amg_lambda(self, level):
        return lib().ref_amg_lambda(self.h, level)

    def amg_nullspace(self, level):
        m = lib().ref_amg_nullspace_dim(self.h)
        rows = self.amg_matrix(level, "A")[0]
        b = np.empty(rows * m)
        lib().ref_amg_nullspace(self.h, level, b)
        return b.reshape(rows, m)

    def amg_aggregates(self, level):
        nn = lib().ref_amg_num_nodes(self.h, level)
        a = np.empty(nn, dtype=np.int32)
        lib().ref_amg_aggregates(self.h, level, a)
        return a

    def amg_node_ptr(self, level):
        nn = lib().ref_amg_num_nodes(self.h, level)
        a = np.empty(nn + 1, dtype=np.int32)
        lib().ref_amg_node_ptr(self.h, level, a)
        return a

    def amg_apply(self, r):
        z = np.empty_like(r)
        _check(lib().ref_amg_apply(self.h, np.ascontiguousarray(r, dtype=np.float64), z))
        return z

    def pcg(self, b, x0=None, tol=1e-8, maxit=1000):
        x = np.zeros(len(b)) if x0 is None else np.array(x0, dtype=np.float64)
        it, rr, cv = C.c_int(), C.c_double(), C.c_int()
        _check(lib().ref_pcg(self.h, np.ascontiguousarray(b, dtype=np.float64), x, tol, maxit, C.byref(it),
                             C.byref(rr), C.byref(cv)))
        return x, it.value, rr.value, bool(cv.value)

    # ---- filter / objective / OC ----------------------------------------------
    def filter_matrix(self):
        nnz = lib().ref_filter_nnz(self.h)
        ptr = np.empty(self.ne + 1, dtype=np.int32)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz)
        rs = np.empty(self.ne)
        lib().ref_filter_get(self.h, ptr, col, val, _ptr(rs))
        return ptr, col, val, rs

    def filter(self, x, adjoint=False):
        y = np.empty(self.ne)
        lib().ref_filter_apply(self.h, 1 if adjoint else 0, np.ascontiguousarray(x, dtype=np.float64), y)
        return y

    def evaluate(self, x, need_gradients=True, want_u=False):
        ncases = self.model.loads.shape[0]
        sc = np.empty(3)
        it = C.c_int()
        dc = np.empty(self.ne) if need_gradients else None
        dv = np.empty(self.ne) if need_gradients else None
        u = np.empty(ncases * self.ndofs) if want_u else None
        en = np.empty(self.ne)
        ph = np.empty(self.ne)
        _check(lib().ref_evaluate(self.h, np.ascontiguousarray(x, dtype=np.float64), 1 if need_gradients else 0, sc,
                                  C.byref(it), _ptr(dc), _ptr(dv), _ptr(u), _ptr(en), _ptr(ph)))
        out = dict(compliance=sc[0], volume=sc[1], volume_fraction=sc[2], iterations=it.value, dc_dx=dc, dv_dx=dv,
                   energy=en, physical=ph)
        if want_u:
            out["u"] = u.reshape(ncases, self.ndofs)
        return out

    def last_evaluate_seconds(self):
        """Wall seconds of the last evaluate(): assembly, AMG setup, PCG, the whole call."""
        t = np.empty(4)
        lib().ref_last_evaluate_seconds(self.h, t)
        return dict(assemble=t[0], setup=t[1], pcg=t[2], total=t[3])

    def oc(self, x, dc, dv, move=0.2, damping=0.5):
        xn = np.empty(self.ne)
        sc = np.empty(4)
        nb = C.c_int()
        _check(lib().ref_oc(self.h, np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(dc),
                            np.ascontiguousarray(dv), move, damping, xn, sc, C.byref(nb)))
        return xn, dict(lambda_=sc[0], volume=sc[1], max_change=sc[2], converged=bool(sc[3]), bisections=nb.value)

    def set_projection(self, beta=0.0, eta=0.5):
        """Project the physical density in evaluate() (beta 0: off)."""
        lib().ref_problem_set_projection(self.h, beta, eta)

    def run_oc(self, x0, max_iterations, change_tol=1e-3, objective_tol=0.0, window=5, cont_steps=1,
               penalty_start=1.0, cont_iterations=25, move=0.2, damping=0.5, projection=None):
        """History rows have 16 columns: the 12 of history.csv, beta (0 without projection), and the
        eroded / dilated volume fractions and dilated target fraction of a robust run.
        projection: dict(eta, beta_start, beta_max, beta_factor, beta_interval, advance_on_convergence,
        robust, robust_delta, robust_volume_interval)."""
        pr = dict(eta=0.5, beta_start=1.0, beta_max=32.0, beta_factor=2.0, beta_interval=50,
                  advance_on_convergence=True, robust=False, robust_delta=0.1, robust_volume_interval=1)
        pr.update(projection or {})
        lib().ref_problem_set_projection_schedule(self.h, 1 if projection else 0, pr["eta"], pr["beta_start"],
                                                  pr["beta_max"], pr["beta_factor"], int(pr["beta_interval"]),
                                                  1 if pr["advance_on_convergence"] else 0)
        lib().ref_problem_set_robust(self.h, 1 if pr["robust"] else 0, pr["robust_delta"],
                                     int(pr["robust_volume_interval"]))
        x = np.array(x0, dtype=np.float64)
        hist = np.zeros(max(max_iterations, 1) * 16)
        rob = np.zeros(5)
        n = C.c_int()
        fc = C.c_double()
        stop = C.c_int()
        fb = C.c_double()
        _check(lib().ref_run_oc(self.h, max_iterations, change_tol, objective_tol, window, cont_steps, penalty_start,
                                cont_iterations, move, damping, x, hist, C.byref(n), C.byref(fc), C.byref(stop),
                                C.byref(fb), rob))
        self.last_stop_reason = ("iteration_cap", "design_change", "objective_stall")[stop.value]
        self.last_final_beta = fb.value
        self.last_robust = dict(compliance_eroded=rob[0], compliance_dilated=rob[1], volume_fraction_eroded=rob[2],
                                volume_fraction_dilated=rob[3], dilated_target_fraction=rob[4])
        return x, hist.reshape(-1, 16)[: n.value], n.value, fc.value


def chunked_dot(a, b):
    return lib().ref_chunked_dot(np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64),
                                 len(a))


def cpqr(block):
    block = np.ascontiguousarray(block, dtype=np.float64)
    rows, m = block.shape
    q = np.zeros(rows * m)
    r = np.zeros(m * m)
    rank = lib().ref_cpqr(block.ravel(), rows, m, q, r)
    return q[: rows * rank].reshape(rows, rank), r[: rank * m].reshape(rank, m)


def tridiag_max(diag, sub):
    return lib().ref_tridiag_max(np.ascontiguousarray(diag, dtype=np.float64),
                                 np.ascontiguousarray(sub if len(sub) else [0.0], dtype=np.float64), len(diag))


def set_num_threads(n):
    lib().ref_set_num_threads(int(n))


def num_threads():
    return lib().ref_num_threads()
