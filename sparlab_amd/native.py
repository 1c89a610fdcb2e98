"""ctypes binding of libsparlab_hip.so (include/sparlab_hip.h).

The Python classes mirror the reference's host interface for this path:
``HipAmgCgSolver`` follows ``LinearSolver`` / ``AmgCgSolver``
(include/sparlab/fem/LinearSolver.hpp:64-111), ``HipProblem`` bundles the
FemModel + Assembler + DensityFilter + ComplianceObjective state of a
structured SIMP problem (src/topopt/Sensitivity.cpp:110-237) and
``run_oc`` follows TopologyOptimizer::run_oc (src/topopt/TopologyOptimizer.cpp:556-724).
Status codes become the reference's exception classes.

The library is the only compute path: if it is missing or no GPU is
visible, the calls raise; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsparlab_hip.so")


class SparlabError(RuntimeError):
    """Base of the errors the native library reports."""


class ConfigError(SparlabError):
    pass


class SolverError(SparlabError):
    pass


class ConvergenceError(SparlabError):
    pass


class ModelError(SparlabError):
    pass


class DeviceError(SparlabError):
    pass


_ERRORS = {1: ConfigError, 2: SolverError, 3: ConvergenceError, 4: ModelError, 5: DeviceError}


class AmgOptions(C.Structure):
    _fields_ = [("strength_threshold", C.c_double), ("max_levels", C.c_int), ("coarse_size", C.c_int),
                ("smoother", C.c_int), ("smoother_degree", C.c_int), ("chebyshev_ratio", C.c_double),
                ("prolongator_damping", C.c_double), ("lanczos_steps", C.c_int),
                ("coarse_pivot_tolerance", C.c_double), ("reuse_aggregates", C.c_int), ("precision", C.c_int),
                ("coarse_reuse", C.c_int)]


class PcgResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("relative_residual", C.c_double), ("converged", C.c_int)]


class AmgStats(C.Structure):
    _fields_ = [("levels", C.c_int), ("reused_aggregates", C.c_int), ("operator_complexity", C.c_double),
                ("grid_complexity", C.c_double), ("coarse_pivot_ratio", C.c_double),
                ("near_null_space_dimension", C.c_int), ("unknowns", C.c_longlong * 16),
                ("nonzeros", C.c_longlong * 16), ("aggregates", C.c_longlong * 16), ("lambda_max", C.c_double * 16),
                ("setup_seconds", C.c_double), ("check_seconds", C.c_double), ("spectral_seconds", C.c_double),
                ("aggregation_seconds", C.c_double), ("smoothing_seconds", C.c_double),
                ("galerkin_seconds", C.c_double), ("coarse_seconds", C.c_double), ("node_product_levels", C.c_int),
                ("storage_nonzeros", C.c_longlong), ("tentative_seconds", C.c_double)]


class EvalOptions(C.Structure):
    _fields_ = [("penalty", C.c_double), ("emin_ratio", C.c_double), ("tolerance", C.c_double),
                ("max_iterations", C.c_int), ("warm_start", C.c_int), ("beta", C.c_double), ("eta", C.c_double)]


class Evaluation(C.Structure):
    _fields_ = [("compliance", C.c_double), ("volume", C.c_double), ("volume_fraction", C.c_double),
                ("linear_iterations", C.c_int), ("assemble_seconds", C.c_double), ("solve_seconds", C.c_double)]


class OcOptions(C.Structure):
    _fields_ = [("move_limit", C.c_double), ("damping", C.c_double), ("lambda_lower", C.c_double),
                ("lambda_upper", C.c_double), ("max_bisections", C.c_int), ("volume_tolerance", C.c_double)]


class OcStep(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("achieved_volume", C.c_double), ("max_change", C.c_double),
                ("bisections", C.c_int), ("volume_converged", C.c_int)]


class ProjectionOptions(C.Structure):
    _fields_ = [("enabled", C.c_int), ("eta", C.c_double), ("beta_start", C.c_double), ("beta_max", C.c_double),
                ("beta_factor", C.c_double), ("beta_interval", C.c_int), ("advance_on_convergence", C.c_int),
                ("robust", C.c_int), ("robust_delta", C.c_double), ("robust_volume_interval", C.c_int)]


class RunOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("change_tolerance", C.c_double), ("objective_tolerance", C.c_double),
                ("objective_window", C.c_int), ("continuation_steps", C.c_int), ("penalty_start", C.c_double),
                ("continuation_iterations", C.c_int), ("oc", OcOptions), ("eval", EvalOptions),
                ("projection", ProjectionOptions)]


class RunResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("stop_reason", C.c_int),
                ("compliance", C.c_double), ("volume", C.c_double), ("volume_fraction", C.c_double),
                ("grey_level", C.c_double), ("final_design_change", C.c_double),
                ("final_objective_change", C.c_double), ("linear_solves", C.c_longlong),
                ("linear_iterations", C.c_longlong), ("total_seconds", C.c_double), ("final_beta", C.c_double),
                ("compliance_eroded", C.c_double), ("compliance_dilated", C.c_double),
                ("volume_fraction_eroded", C.c_double), ("volume_fraction_dilated", C.c_double),
                ("dilated_target_fraction", C.c_double)]


# slx_stop_reason -> the reference's stop_reason strings (TopologyOptimizer.hpp:206,
# TopologyOptimizer.cpp:692)
STOP_REASONS = ("iteration_cap", "design_change", "objective_stall")

# columns of a run_oc history row, the order of the reference's history.csv
# (src/io/ResultWriter.cpp:1428-1470); the reference writes beta only for a projected run
HISTORY_COLUMNS = ("iteration", "penalty[-]", "compliance[J]", "volume[m3]", "volume_fraction[-]",
                   "max_design_change[-]", "lagrange_multiplier[-]", "grey_level[-]", "oc_bisections[-]",
                   "volume_converged[-]", "seconds[s]", "linear_iterations[-]", "beta[-]",
                   "eroded_volume_fraction[-]", "dilated_volume_fraction[-]", "dilated_target_fraction[-]")
H_ITER, H_PENALTY, H_COMPLIANCE, H_VOLUME, H_VF, H_CHANGE, H_LAMBDA, H_GREY, H_BISECT, H_VCONV, H_SECONDS, H_LINEAR, \
    H_BETA, H_ERODED_VF, H_DILATED_VF, H_DILATED_TARGET = range(len(HISTORY_COLUMNS))


class ModalOptions(C.Structure):
    _fields_ = [("num_modes", C.c_int), ("lumped", C.c_int), ("max_iterations", C.c_int), ("tolerance", C.c_double),
                ("residual_tolerance", C.c_double), ("rigid_body_ratio", C.c_double), ("shift_factor", C.c_double),
                ("density", C.c_double), ("linear_tolerance", C.c_double), ("linear_max_iterations", C.c_int),
                ("warm_start", C.c_int)]


class ModalResult(C.Structure):
    _fields_ = [("num_modes", C.c_int), ("subspace_size", C.c_int), ("iterations", C.c_int), ("converged", C.c_int),
                ("rigid_body_modes", C.c_int), ("residual_warnings", C.c_int), ("linear_iterations", C.c_longlong),
                ("final_change", C.c_double), ("total_mass", C.c_double), ("seconds", C.c_double)]


class WarmStartOptions(C.Structure):
    _fields_ = [("enable", C.c_int), ("train_solves", C.c_int), ("min_samples", C.c_int),
                ("steps_per_solve", C.c_int), ("buffer", C.c_int), ("learning_rate", C.c_double),
                ("seed", C.c_ulonglong)]


class WarmStartStats(C.Structure):
    _fields_ = [("samples", C.c_int), ("train_steps", C.c_int), ("first_loss", C.c_double),
                ("last_loss", C.c_double), ("predictions", C.c_int), ("rejected", C.c_int),
                ("step_sum", C.c_double), ("energy_ratio_sum", C.c_double), ("last_energy_ratio", C.c_double),
                ("net_ratio_sum", C.c_double), ("secant_ratio_sum", C.c_double),
                ("residual_ratio_sum", C.c_double), ("train_seconds", C.c_double),
                ("predict_seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

# every entry point of include/sparlab_hip.h with its ctypes signature
_VP = C.c_void_p
_SIGNATURES = {
    "slx_create": (C.c_int, [C.c_int, C.POINTER(_VP)]),
    "slx_destroy": (None, [_VP]),
    "slx_last_error": (C.c_char_p, [_VP]),
    "slx_amg_default_options": (None, [C.POINTER(AmgOptions)]),
    "slx_eval_default_options": (None, [C.POINTER(EvalOptions)]),
    "slx_oc_default_options": (None, [C.POINTER(OcOptions)]),
    "slx_model_structured": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 6),
    "slx_model_set_supports": (C.c_int, [_VP, _VP]),
    "slx_model_set_loads": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "slx_model_finalize": (C.c_int, [_VP]),
    "slx_model_sizes": (C.c_int, [_VP, _VP]),
    "slx_ke0_get": (C.c_int, [_VP, _VP]),
    "slx_free_dofs_get": (C.c_int, [_VP, _VP]),
    "slx_assemble": (C.c_int, [_VP, _VP]),
    "slx_kff_get": (C.c_int, [_VP, _VP, _VP, _VP]),
    "slx_kff_apply": (C.c_int, [_VP, _VP, _VP, C.POINTER(C.c_int)]),
    "slx_kff_apply_noalloc": (C.c_int, [_VP, _VP, _VP, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "slx_amg_setup": (C.c_int, [_VP, C.POINTER(AmgOptions)]),
    "slx_factorize_csr": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP, C.c_int, C.c_int, C.c_int, _VP, _VP,
                                    C.POINTER(AmgOptions)]),
    "slx_solve": (C.c_int, [_VP, _VP, _VP, C.c_double, C.c_int, C.POINTER(PcgResult)]),
    "slx_solve_case": (C.c_int, [_VP, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(PcgResult)]),
    "slx_solve_segmented": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP, C.c_double, C.c_int, _VP]),
    "slx_amg_apply": (C.c_int, [_VP, _VP, _VP]),
    "slx_amg_time_smoother": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "slx_amg_stats_get": (C.c_int, [_VP, C.POINTER(AmgStats)]),
    "slx_amg_level_info": (C.c_int, [_VP, C.c_int, C.c_int, _VP]),
    "slx_amg_level_get": (C.c_int, [_VP, C.c_int, C.c_int, _VP, _VP, _VP]),
    "slx_amg_aggregates_get": (C.c_int, [_VP, C.c_int, _VP]),
    "slx_solution_get": (C.c_int, [_VP, C.c_int, _VP]),
    "slx_design_set": (C.c_int, [_VP, C.c_double, _VP, _VP, _VP]),
    "slx_filter_build": (C.c_int, [_VP, C.c_double]),
    "slx_filter_apply": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "slx_evaluate": (C.c_int, [_VP, _VP, C.POINTER(EvalOptions), C.POINTER(Evaluation), _VP, _VP, _VP]),
    "slx_oc_update": (C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(OcOptions), _VP, C.POINTER(OcStep)]),
    "slx_run_oc": (C.c_int, [_VP, C.POINTER(RunOptions), _VP, _VP, C.POINTER(RunResult)]),
    "slx_load_case_compliance_get": (C.c_int, [_VP, _VP]),
    "slx_solve_adjoint": (C.c_int, [_VP, _VP, C.c_int, C.c_double, C.c_int, _VP, C.POINTER(PcgResult)]),
    "slx_modal_default_options": (None, [C.POINTER(ModalOptions)]),
    "slx_modal_solve": (C.c_int, [_VP, C.POINTER(ModalOptions), _VP, _VP, _VP, _VP, _VP, C.POINTER(ModalResult)]),
    "slx_warmstart_default_options": (None, [C.POINTER(WarmStartOptions)]),
    "slx_warmstart_configure": (C.c_int, [_VP, C.POINTER(WarmStartOptions)]),
    "slx_warmstart_stats_get": (C.c_int, [_VP, C.POINTER(WarmStartStats)]),
    "slx_net_configure": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]),
    "slx_net_parameter_count": (C.c_int, [_VP, C.POINTER(C.c_longlong)]),
    "slx_net_parameters_get": (C.c_int, [_VP, _VP]),
    "slx_net_parameters_set": (C.c_int, [_VP, _VP]),
    "slx_net_forward": (C.c_int, [_VP, _VP, _VP]),
    "slx_net_gradients": (C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(C.c_double), _VP]),
    "slx_net_train": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, C.c_double, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]),
    "slx_timing_enable": (C.c_int, [_VP, C.c_int]),
    "slx_timing_get": (C.c_int, [_VP, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "slx_synchronize": (C.c_int, [_VP]),
}


def load():
    """Load the native library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DeviceError(f"{LIB_PATH} is missing: build it with `python sparlab_amd/build.py` "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def amg_options(**kw) -> AmgOptions:
    o = AmgOptions()
    load().slx_amg_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def eval_options(**kw) -> EvalOptions:
    o = EvalOptions()
    load().slx_eval_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def oc_options(**kw) -> OcOptions:
    o = OcOptions()
    load().slx_oc_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Context:
    """One GPU stream and its device-resident problem state."""

    def __init__(self, device: int = 0):
        L = load()
        h = _VP()
        rc = L.slx_create(device, C.byref(h))
        if rc != 0:
            raise DeviceError(f"slx_create failed on device {device} (code {rc})")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().slx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            msg = load().slx_last_error(self.h).decode()
            raise _ERRORS.get(rc, SparlabError)(msg)

    def timing(self, on=True):
        self.check(load().slx_timing_enable(self.h, 1 if on else 0))

    def timer(self, name):
        ms, n = C.c_double(), C.c_longlong()
        self.check(load().slx_timing_get(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        self.check(load().slx_synchronize(self.h))


def warmstart_options(**kw) -> WarmStartOptions:
    o = WarmStartOptions()
    load().slx_warmstart_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class WarmNet:
    """The warm-start network on an X x Y x Z node grid (the slx_net_* entry points).

    Parameters are FP32, per layer weights [cout][27][cin] then biases [cout];
    inputs [V][nin], outputs and targets [V][nout], voxel = (z*Y + y)*X + x.
    """

    LAYERS = ((16, 16), (16, 32), (32, 64), (64, 32), (32, 16), (16, 16))

    def __init__(self, X, Y, Z, nin, nout, seed=1, ctx: Context | None = None):
        self.ctx = ctx or Context(0)
        self.shape = (X, Y, Z)
        self.V = X * Y * Z
        self.nin, self.nout = nin, nout
        self.ctx.check(load().slx_net_configure(self.ctx.h, X, Y, Z, nin, nout, seed))
        n = C.c_longlong()
        self.ctx.check(load().slx_net_parameter_count(self.ctx.h, C.byref(n)))
        self.nparams = n.value

    def parameters(self):
        p = np.empty(self.nparams, dtype=np.float32)
        self.ctx.check(load().slx_net_parameters_get(self.ctx.h, _p(p)))
        return p

    def set_parameters(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self.ctx.check(load().slx_net_parameters_set(self.ctx.h, _p(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.V, self.nin)
        y = np.empty((self.V, self.nout), dtype=np.float32)
        self.ctx.check(load().slx_net_forward(self.ctx.h, _p(x), _p(y)))
        return y

    def _xtw(self, x, t, w):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.V, self.nin)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(self.V, self.nout)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32).reshape(self.V)
        return x, t, w

    def gradients(self, x, t, w=None):
        x, t, w = self._xtw(x, t, w)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        self.ctx.check(load().slx_net_gradients(self.ctx.h, _p(x), _p(t), _p(w), C.byref(loss), _p(g)))
        return loss.value, g

    def train(self, x, t, steps, lr=1e-3, w=None):
        x, t, w = self._xtw(x, t, w)
        a, b = C.c_double(), C.c_double()
        self.ctx.check(load().slx_net_train(self.ctx.h, _p(x), _p(t), _p(w), steps, lr, C.byref(a), C.byref(b)))
        return a.value, b.value


class HipAmgCgSolver:
    """The LinearSolver seam (AmgCgSolver, src/fem/LinearSolver.cpp:219-290) for a
    caller-assembled CSR matrix and its DOF layout."""

    def __init__(self, tolerance=1e-12, max_iterations=0, ctx: Context | None = None, **amg):
        self.ctx = ctx or Context()
        self.tolerance = tolerance
        self.max_iterations = max_iterations
        self.options = amg_options(**amg)
        self.layout = None
        self._last = PcgResult()

    def name(self):
        return "AMG-CG(chebyshev)"

    def iterative(self):
        return True

    def set_layout(self, dim, coords, unknowns, dofs_per_node=0):
        self.layout = (dim, dofs_per_node, np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, dim),
                       np.ascontiguousarray(unknowns, dtype=np.int32))

    def factorize(self, ptr, col, val):
        if self.layout is None:
            raise SolverError(self.name() + " needs the DOF layout (node coordinates of the unknowns) to build its "
                              "rigid-body near-null space; the caller did not supply one")
        dim, ndpn, coords, unknowns = self.layout
        ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = _f64(val)
        self._n = len(ptr) - 1
        self.ctx.check(load().slx_factorize_csr(self.ctx.h, self._n, _p(ptr), _p(col), _p(val), dim, ndpn,
                                                coords.shape[0], _p(coords), _p(unknowns), C.byref(self.options)))
        self.ctx.owner = self  # the context now holds this solver's matrix

    def _require_matrix(self, n):
        if getattr(self, "_n", None) is None or getattr(self.ctx, "owner", None) is not self:
            raise SolverError(f"{self.name()}: factorize() this solver's matrix first (another solver sharing the "
                              "context replaced it)")
        if n != self._n:
            raise SolverError(f"{self.name()}: right-hand side of length {n} for a matrix of size {self._n}")

    def solve_from(self, b, x0=None):
        b = _f64(b)
        self._require_matrix(len(b))
        x = np.zeros_like(b) if x0 is None or len(x0) != len(b) else _f64(x0).copy()
        cap = self.max_iterations if self.max_iterations > 0 else 1000
        self.ctx.check(load().slx_solve(self.ctx.h, _p(b), _p(x), self.tolerance, cap, C.byref(self._last)))
        if not self._last.converged:
            raise ConvergenceError(f"{self.name()} did not converge: {self._last.iterations} iterations left a "
                                   f"relative residual of {self._last.relative_residual} against a tolerance of "
                                   f"{self.tolerance}")
        return x

    def solve(self, b):
        return self.solve_from(b, None)

    def solve_segmented(self, seg_ptr, b, x0=None):
        """Independent solves on the diagonal blocks [seg_ptr[s], seg_ptr[s+1]);
        returns x and per segment (iterations, relative residual, converged)."""
        seg_ptr = np.ascontiguousarray(seg_ptr, dtype=np.int32)
        b = _f64(b)
        self._require_matrix(len(b))
        x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
        res = (PcgResult * (len(seg_ptr) - 1))()
        cap = self.max_iterations if self.max_iterations > 0 else 1000
        self.ctx.check(load().slx_solve_segmented(self.ctx.h, len(seg_ptr) - 1, _p(seg_ptr), _p(b), _p(x),
                                                  self.tolerance, cap, res))
        return x, [(r.iterations, r.relative_residual, bool(r.converged)) for r in res]

    def last_iterations(self):
        return self._last.iterations

    def last_error(self):
        return self._last.relative_residual

    def amg_stats(self):
        """LinearSolver::amg_stats() (include/sparlab/fem/LinearSolver.hpp:100-104)."""
        self._require_matrix(self._n if getattr(self, "_n", None) is not None else -1)
        s = AmgStats()
        self.ctx.check(load().slx_amg_stats_get(self.ctx.h, C.byref(s)))
        return s

    def storage_nonzeros(self):
        """LinearSolver::storage_nonzeros() of AmgCgSolver: AmgStats::storage_nonzeros
        (src/fem/LinearSolver.cpp:258, src/fem/Multigrid.cpp:1089-1100)."""
        return int(self.amg_stats().storage_nonzeros)


class HipProblem:
    """A structured SIMP problem resident on one GPU."""

    def __init__(self, model, device=0, filter_radius=None, ctx: Context | None = None):
        L = load()
        self.ctx = ctx or Context(device)
        self.model = model
        h = self.ctx.h
        chk = self.ctx.check
        chk(L.slx_model_structured(h, model.dim, model.nx, model.ny, max(model.nz, 1), model.lx, model.ly,
                                   model.lz, model.E, model.nu, model.thickness))
        fixed = np.ascontiguousarray(model.fixed, dtype=np.uint8)
        chk(L.slx_model_set_supports(h, _p(fixed)))
        if model.loads is not None:
            loads = _f64(model.loads)
            w = _f64(model.weights)
            chk(L.slx_model_set_loads(h, loads.shape[0], _p(loads), _p(w)))
        chk(L.slx_model_finalize(h))
        sizes = np.zeros(6, dtype=np.int64)
        chk(L.slx_model_sizes(h, _p(sizes)))
        self.nnodes, self.ne, self.ndofs, self.nfree, self.kff_nnz, self.ncases = (int(v) for v in sizes)
        chk(L.slx_design_set(h, model.volume_fraction, _p(_f64(model.lower)), _p(_f64(model.upper)),
                             _p(np.ascontiguousarray(model.passive, dtype=np.uint8))))
        r = model.filter_radius if filter_radius is None else filter_radius
        chk(L.slx_filter_build(h, float(r)))
        self.ctx.owner = self  # the context now holds this problem

    def _own(self):
        if getattr(self.ctx, "owner", None) is not self:
            raise ModelError("this problem's context now holds another problem or matrix; give each problem its "
                             "own Context to use them side by side")

    # ---- structure -----------------------------------------------------------
    def ke0(self):
        self._own()
        nd = (8 if self.model.dim == 3 else 4) * self.model.dim
        k = np.empty(nd * nd)
        self.ctx.check(load().slx_ke0_get(self.ctx.h, _p(k)))
        return k.reshape(nd, nd)

    def free(self):
        self._own()
        f = np.empty(self.nfree, dtype=np.int32)
        self.ctx.check(load().slx_free_dofs_get(self.ctx.h, _p(f)))
        return f

    # ---- assembler / solver seams ------------------------------------------------
    def assemble(self, scale=None):
        self._own()
        s = None if scale is None else _f64(scale)
        self.ctx.check(load().slx_assemble(self.ctx.h, _p(s)))

    def kff(self):
        self._own()
        ptr = np.empty(self.nfree + 1, dtype=np.int32)
        col = np.empty(self.kff_nnz, dtype=np.int32)
        val = np.empty(self.kff_nnz)
        self.ctx.check(load().slx_kff_get(self.ctx.h, _p(ptr), _p(col), _p(val)))
        return ptr, col, val

    def kff_apply(self, x):
        """y = K_ff x with the solver's operator; returns (y, matrix_free)."""
        self._own()
        y = np.empty(self.nfree)
        mf = C.c_int()
        self.ctx.check(load().slx_kff_apply(self.ctx.h, _p(_f64(x)), _p(y), C.byref(mf)))
        return y, bool(mf.value)

    def kff_apply_noalloc(self, x):
        """y = K_ff x and x.K_ff x through the entry point that allocates no device
        memory; returns (y, xKx, matrix_free)."""
        self._own()
        y = np.empty(self.nfree)
        q = C.c_double()
        mf = C.c_int()
        self.ctx.check(load().slx_kff_apply_noalloc(self.ctx.h, _p(_f64(x)), _p(y), C.byref(q), C.byref(mf)))
        return y, q.value, bool(mf.value)

    def amg_setup(self, **amg):
        self._own()
        self.ctx.check(load().slx_amg_setup(self.ctx.h, C.byref(amg_options(**amg))))

    def amg_stats(self):
        self._own()
        s = AmgStats()
        self.ctx.check(load().slx_amg_stats_get(self.ctx.h, C.byref(s)))
        return s

    def amg_matrix(self, level, which):
        self._own()
        w = {"A": 0, "P": 1, "Ptent": 2, "R": 3}[which]
        info = np.zeros(3, dtype=np.int64)
        L = load()
        self.ctx.check(L.slx_amg_level_info(self.ctx.h, level, w, _p(info)))
        rows, cols, nnz = (int(v) for v in info)
        ptr = np.empty(rows + 1, dtype=np.int32)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz)
        self.ctx.check(L.slx_amg_level_get(self.ctx.h, level, w, _p(ptr), _p(col), _p(val)))
        return rows, cols, ptr, col, val

    def mf_tile(self):
        """Nodes per direction of the matrix-free operator's tile that the setup
        timed fastest, (0, 0, 0) when level 0 is assembled."""
        self._own()
        info = np.zeros(3, dtype=np.int64)
        self.ctx.check(load().slx_amg_level_info(self.ctx.h, 0, 9, _p(info)))
        return tuple(int(v) for v in info)

    def amg_aggregates(self, level, num_nodes):
        self._own()
        a = np.empty(num_nodes, dtype=np.int32)
        self.ctx.check(load().slx_amg_aggregates_get(self.ctx.h, level, _p(a)))
        return a

    def time_smoother(self, level=1, reps=50):
        """Milliseconds per Chebyshev step of a middle level (back-to-back launches)."""
        self._own()
        ms = C.c_double()
        self.ctx.check(load().slx_amg_time_smoother(self.ctx.h, level, reps, C.byref(ms)))
        return ms.value

    def amg_apply(self, r):
        self._own()
        r = _f64(r)
        z = np.empty_like(r)
        self.ctx.check(load().slx_amg_apply(self.ctx.h, _p(r), _p(z)))
        return z

    def solve(self, b, x0=None, tol=1e-8, maxit=1000):
        self._own()
        b = _f64(b)
        x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
        res = PcgResult()
        self.ctx.check(load().slx_solve(self.ctx.h, _p(b), _p(x), tol, maxit, C.byref(res)))
        return x, res.iterations, res.relative_residual, bool(res.converged)

    def solve_case(self, case=0, warm=True, tol=1e-8, maxit=1000):
        self._own()
        res = PcgResult()
        self.ctx.check(load().slx_solve_case(self.ctx.h, case, 1 if warm else 0, tol, maxit, C.byref(res)))
        return res.iterations, res.relative_residual, bool(res.converged)

    def solution(self, case=0):
        self._own()
        u = np.empty(self.nfree)
        self.ctx.check(load().slx_solution_get(self.ctx.h, case, _p(u)))
        return u

    # ---- topology optimisation ---------------------------------------------------
    def filter(self, x, adjoint=False):
        self._own()
        y = np.empty(self.ne)
        self.ctx.check(load().slx_filter_apply(self.ctx.h, 1 if adjoint else 0, _p(_f64(x)), _p(y)))
        return y

    def warm_start(self, enable=True, **kw):
        """Turn the learned warm start on (or off) and reset its network."""
        self._own()
        o = warmstart_options(enable=1 if enable else 0, **kw)
        self.ctx.check(load().slx_warmstart_configure(self.ctx.h, C.byref(o)))

    def warm_start_stats(self):
        self._own()
        st = WarmStartStats()
        self.ctx.check(load().slx_warmstart_stats_get(self.ctx.h, C.byref(st)))
        return st.as_dict()

    def evaluate(self, x, tol=1e-8, need_gradients=True, warm_start=True, penalty=None, max_iterations=1000,
                 beta=0.0, eta=0.5):
        """ComplianceObjective::evaluate; beta > 0 projects the filtered density at eta."""
        self._own()
        o = eval_options(tolerance=tol, warm_start=1 if warm_start else 0, max_iterations=max_iterations,
                         penalty=self.model.penalty if penalty is None else penalty,
                         emin_ratio=self.model.emin_ratio, beta=beta, eta=eta)
        ev = Evaluation()
        dc = np.empty(self.ne) if need_gradients else None
        dv = np.empty(self.ne) if need_gradients else None
        en = np.empty(self.ne)
        self.ctx.check(load().slx_evaluate(self.ctx.h, _p(_f64(x)), C.byref(o), C.byref(ev), _p(dc), _p(dv),
                                           _p(en)))
        return dict(compliance=ev.compliance, volume=ev.volume, volume_fraction=ev.volume_fraction,
                    iterations=ev.linear_iterations, dc_dx=dc, dv_dx=dv, energy=en,
                    assemble_seconds=ev.assemble_seconds, solve_seconds=ev.solve_seconds)

    def oc(self, x, dc, dv, move=0.2, damping=0.5):
        self._own()
        xn = np.empty(self.ne)
        st = OcStep()
        self.ctx.check(load().slx_oc_update(self.ctx.h, _p(_f64(x)), _p(_f64(dc)), _p(_f64(dv)),
                                            C.byref(oc_options(move_limit=move, damping=damping)), _p(xn),
                                            C.byref(st)))
        return xn, dict(lambda_=st.lambda_, volume=st.achieved_volume, max_change=st.max_change,
                        converged=bool(st.volume_converged), bisections=st.bisections)

    def run_oc(self, x0, max_iterations, change_tol=1e-3, objective_tol=0.0, window=5, cont_steps=1,
               penalty_start=1.0, cont_iterations=25, move=0.2, damping=0.5, tol=1e-8, projection=None,
               warm_start=True):
        """TopologyOptimizer::run_oc: returns (x, history, iterations, final compliance);
        the full result of the last call is in ``last_run``. projection: None, or a dict of
        ProjectionOptions (eta, beta_start, beta_max, beta_factor, beta_interval,
        advance_on_convergence, robust, robust_delta, robust_volume_interval) for the Heaviside
        projection with beta continuation. warm_start: each solve starts from the previous
        solution (linear.warm_start, Config.cpp:1086), else from zero."""
        self._own()
        o = RunOptions()
        o.max_iterations = max_iterations
        o.change_tolerance = change_tol
        o.objective_tolerance = objective_tol
        o.objective_window = window
        o.continuation_steps = cont_steps
        o.penalty_start = penalty_start
        o.continuation_iterations = cont_iterations
        o.oc = oc_options(move_limit=move, damping=damping)
        o.eval = eval_options(tolerance=tol, penalty=self.model.penalty, emin_ratio=self.model.emin_ratio,
                              warm_start=1 if warm_start else 0)
        pr = dict(eta=0.5, beta_start=1.0, beta_max=32.0, beta_factor=2.0, beta_interval=50,
                  advance_on_convergence=True, robust=False, robust_delta=0.1, robust_volume_interval=1)
        pr.update({k: v for k, v in (projection or {}).items() if k != "erosion_check"})
        o.projection = ProjectionOptions(1 if projection else 0, pr["eta"], pr["beta_start"], pr["beta_max"],
                                         pr["beta_factor"], int(pr["beta_interval"]),
                                         1 if pr["advance_on_convergence"] else 0, 1 if pr["robust"] else 0,
                                         pr["robust_delta"], int(pr["robust_volume_interval"]))
        x = _f64(x0).copy()
        hist = np.zeros(max(max_iterations, 1) * len(HISTORY_COLUMNS))
        r = RunResult()
        self.ctx.check(load().slx_run_oc(self.ctx.h, C.byref(o), _p(x), _p(hist), C.byref(r)))
        h = hist.reshape(-1, len(HISTORY_COLUMNS))[: r.iterations]
        self.last_run = {name: getattr(r, name) for name, _ in RunResult._fields_}
        self.last_run["stop_reason"] = STOP_REASONS[r.stop_reason]
        return x, h, r.iterations, r.compliance

    def modal(self, density, num_modes=6, mass_type="consistent", max_iterations=200, tolerance=1e-10,
              residual_tolerance=1e-6, rigid_body_ratio=1e-10, shift_factor=0.0, linear_tolerance=1e-8,
              linear_max_iterations=1000, stiffness_scale=None, mass_scale=None, shapes=False, warm_start=True):
        """solve_modal (src/fem/ModalAnalysis.cpp:120-418) on the device: the lowest eigenpairs of
        (K_ff, M_ff) by subspace iteration with warm-started AMG-PCG solves. Returns the fields of the
        reference's modal_json (src/io/ResultWriter.cpp:114-136) plus ``mode_shapes`` (ndofs x m,
        M-orthonormal) when ``shapes``. Re-assembles and re-factorizes the context's K_ff."""
        self._own()
        if mass_type not in ("consistent", "lumped"):
            raise ConfigError(f"unknown modal.mass_type '{mass_type}' (consistent or lumped)")
        o = ModalOptions()
        load().slx_modal_default_options(C.byref(o))
        o.num_modes = int(num_modes)
        o.lumped = 1 if mass_type == "lumped" else 0
        o.max_iterations = int(max_iterations)
        o.tolerance = tolerance
        o.residual_tolerance = residual_tolerance
        o.rigid_body_ratio = rigid_body_ratio
        o.shift_factor = shift_factor
        o.density = float(density)
        o.linear_tolerance = linear_tolerance
        o.linear_max_iterations = int(linear_max_iterations)
        o.warm_start = 1 if warm_start else 0
        m = max(1, min(int(num_modes), self.nfree))
        clamp = None
        if m < int(num_modes):  # ModalAnalysis.cpp:153-160
            clamp = (f"requested {int(num_modes)} modes but the model has only {self.nfree} free DOFs; "
                     f"returning {m}")
            warnings.warn(clamp, RuntimeWarning, stacklevel=2)
        lam = np.zeros(m)
        res = np.zeros(m)
        phi = np.zeros((m, self.ndofs)) if shapes else None
        ks = None if stiffness_scale is None else _f64(stiffness_scale)
        ms = None if mass_scale is None else _f64(mass_scale)
        for name, v in (("stiffness", ks), ("mass", ms)):
            if v is not None and v.shape != (self.ne,):
                raise ModelError(f"{name} scale vector has {v.size} entries for {self.ne} elements")
        r = ModalResult()
        self.ctx.check(load().slx_modal_solve(self.ctx.h, C.byref(o), _p(ks), _p(ms), _p(lam), _p(res), _p(phi),
                                              C.byref(r)))
        omega = np.sqrt(np.maximum(lam, 0.0))
        out = {
            "num_modes": r.num_modes,
            "eigenvalues_1_per_s2": lam.tolist(),
            "angular_frequencies_rad_per_s": omega.tolist(),
            "frequencies_hz": (omega / (2.0 * np.pi)).tolist(),
            "eigenpair_residuals": res.tolist(),
            "total_mass_kg": r.total_mass,
            "subspace_size": r.subspace_size,
            "iterations": r.iterations,
            "converged": bool(r.converged),
            "final_relative_change": r.final_change,
            "linear_solver": "amg_cg",
            "linear_iterations": int(r.linear_iterations),
            "warnings": [clamp] if clamp else [],
            "seconds": r.seconds,
        }
        if r.rigid_body_modes:
            out["warnings"].append(f"{r.rigid_body_modes} mode(s) indistinguishable from zero at this "
                                   "stiffness/mass scale: the model retains a rigid-body or mechanism mode")
        if r.residual_warnings:
            out["warnings"].append(f"{r.residual_warnings} mode(s) above the eigenpair residual tolerance "
                                   f"{residual_tolerance}")
        if shapes:
            out["mode_shapes"] = phi.T.copy()
        return out

    def solve_adjoint(self, rhs, slot=-1, tol=1e-8, max_iterations=1000):
        """ComplianceObjective::solve_adjoint: K_ff lambda = rhs (full-length vectors) on the
        operator of the last evaluation, warm-started from ``slot``'s previous solution when
        slot >= 0. Returns (lambda, PCG iterations)."""
        self._own()
        b = _f64(rhs)
        if b.shape != (self.ndofs,):
            raise ConfigError(f"the adjoint right-hand side has {b.size} entries for {self.ndofs} DOFs")
        lam = np.empty(self.ndofs)
        r = PcgResult()
        self.ctx.check(load().slx_solve_adjoint(self.ctx.h, _p(b), int(slot), tol, int(max_iterations), _p(lam),
                                                C.byref(r)))
        return lam, r.iterations

    def load_case_compliance(self):
        """f_l . u_l of each load case at the last evaluation."""
        out = np.zeros(self.model.loads.shape[0])
        self.ctx.check(load().slx_load_case_compliance_get(self.ctx.h, _p(out)))
        return out
