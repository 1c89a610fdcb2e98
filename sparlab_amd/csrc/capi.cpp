// C ABI of the sparlab_hip library (include/sparlab_hip.h): context, model,
// assembly, linear solver seam, density filter, compliance objective, OC
// update and the OC loop, all on one HIP stream.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>

#include "slx_kernels.hpp"
#include "slx_mf.hpp"
#include "slx_modal.hpp"
#include "slx_model.hpp"
#include "slx_net.hpp"
#include "slx_solver.hpp"

namespace slx {

void fail(int code, const std::string& msg) { throw Error(code, msg); }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::ostringstream os;
    os << "HIP error " << hipGetErrorString(e) << " in " << what;
    throw Error(SLX_DEVICE_ERROR, os.str());
  }
}

namespace {

using Clock = std::chrono::steady_clock;

template <class F>
slx_status guarded(slx_context* ctx, F&& f) {
  Context* c = reinterpret_cast<Context*>(ctx);
  if (!c) return SLX_CONFIG_ERROR;
  try {
    if (c->device >= 0) SLX_HIP(hipSetDevice(c->device));
    f(*c);
    c->last_error.clear();
    return SLX_OK;
  } catch (const Error& e) {
    c->last_error = e.msg;
    return (slx_status)e.code;
  } catch (const std::exception& e) {
    c->last_error = e.what();
    return SLX_DEVICE_ERROR;
  }
}

void require_model(const Context& c) {
  if (c.k_source == Context::KSource::caller)
    fail(SLX_MODEL_ERROR,
         "the context holds a caller-assembled matrix (slx_factorize_csr); call slx_model_finalize to use the "
         "structured model again");
  if (!c.finalized) fail(SLX_MODEL_ERROR, "the model is not finalised; call slx_model_finalize first");
}

void require_amg(const Context& c) {
  if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "AMG-CG: solve called before factorize");
}

void sync(Context& c) { SLX_HIP(hipStreamSynchronize(c.stream)); }

double read_scalar(Context& c, const double* dev) {
  double* hb = c.host_buffer(16);
  SLX_HIP(hipMemcpyAsync(hb + 14, dev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  sync(c);
  return hb[14];
}

void set_layout_from_model(Context& c) {
  c.layout_dim = c.dim;
  c.layout_ndpn = c.dim;
  c.layout_nnodes = (int)c.nnodes;
  c.layout_coords = c.coords;
  c.layout_unknowns = c.free_dofs;
  ++c.layout_version;
}

void upload_filter(Context& c) {
  c.dH.upload(c.H, c.stream);
  c.dHt.upload(c.Ht, c.stream);
  c.htvol_valid = false;
}

// dV/dx = H^T vol (DensityFilter::pull_back of the element volumes), computed
// once per filter
const double* htvol(Context& c) {
  if (!c.htvol_valid) {
    c.d_htvol.alloc((size_t)c.nelem);
    k::spmv_serial(c.dHt, c.d_vol.p, c.d_htvol.p, c.stream);
    c.htvol_valid = true;
  }
  return c.d_htvol.p;
}

void ensure_topopt_buffers(Context& c) {
  const size_t ne = (size_t)c.nelem;
  for (DBuf<double>* b : {&c.d_x, &c.d_phys, &c.d_fac, &c.d_dfac, &c.d_dc, &c.d_energy, &c.d_tmp, &c.d_tmp2,
                          &c.d_dcx, &c.d_dvx, &c.d_dproj})
    b->alloc(ne);
  if (c.d_vol.n != ne) c.d_vol.upload(c.volumes, c.stream);
  if (c.lower.size() != ne) {
    c.lower.assign(ne, 0.0);
    c.upper.assign(ne, 1.0);
    c.passive.assign(ne, 0);
    c.design_dirty = true;
  }
  if (c.design_dirty || c.d_lower.n != ne) {  // the bounds change only through slx_design_set
    c.d_lower.upload(c.lower, c.stream);
    c.d_upper.upload(c.upper, c.stream);
    c.d_passive.upload(c.passive, c.stream);
    c.unit_bounds = true;
    for (size_t e = 0; e < ne && c.unit_bounds; ++e) c.unit_bounds = c.lower[e] >= 0.0 && c.upper[e] <= 1.0;
    c.design_dirty = false;
  }
  if (c.H.rows != (int)ne) {
    identity_filter(c);
    upload_filter(c);
  }
  const size_t need = std::max<size_t>(4096, (size_t)k::filter_volume_blocks(c.dH) + 8);
  if (c.partials.n < need) c.partials.alloc(need);
  if (!c.scal.p) c.scal.alloc(16);
}

// values = false keeps K_ff's values (they then lag the factors: k_stale) and
// refreshes only the matrix-free operator's factors
void assemble(Context& c, const double* d_scale, bool values = true) {
  if (values) {
    Timed t(c, "assemble");
    k::assemble_structured(c.dim, c.K, c.d_free.p, d_scale, c.d_ke0.p, c.nx, c.ny, c.nz, c.stream, c.d_dofxyz.p,
                           c.d_nodemap.p);
    c.k_stale = false;
  }
  // the same factors drive the matrix-free products of Hex8 models
  c.mf = c.dim == 3 && matrix_free_enabled();
  if (!c.mf) return;
  const long long ne = c.nelem;
  c.mf_scale.alloc((size_t)ne);
  c.mf_scale32.alloc((size_t)ne);
  k::copy(ne, d_scale, c.mf_scale.p, c.stream);
  k::to_float(ne, d_scale, c.mf_scale32.p, c.stream);
  if (c.mf_ke32.n != c.d_ke0.n) {
    c.mf_ke32.alloc(c.d_ke0.n);
    k::to_float((long long)c.d_ke0.n, c.d_ke0.p, c.mf_ke32.p, c.stream);
  }
}

// K's pattern stays on the host beside its device copy: the multigrid setup
// reads its structure there, while the values stay on the device
void keep_host_pattern(Context& c, HCsr&& pat) {
  c.k_host = std::move(pat);
  std::vector<double>().swap(c.k_host.val);
  c.k_host_version = c.k_pattern_version;
}

// one load case from its last solution (or zero) to the tolerance
slx_pcg_result solve_case(Context& c, int l, bool warm, double tol, int maxit) {
  if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "AMG-CG: solve called before factorize");
  const size_t nf = c.free_dofs.size();
  double* x = c.d_u.p + (size_t)l * nf;
  if (!warm || !c.have_u[(size_t)l]) k::fill((long long)nf, 0.0, x, c.stream);
  slx_pcg_result r = pcg(c, c.d_loads_free.p + (size_t)l * nf, x, tol, maxit > 0 ? maxit : 1000);
  c.have_u[(size_t)l] = 1;
  return r;
}

void require_converged(Context& c, const slx_pcg_result& r, double tol) {
  if (r.converged) return;
  std::ostringstream os;
  os << "AMG-CG(" << (c.amg_opt.smoother == 1 ? "gauss_seidel" : "chebyshev") << ") did not converge: " << r.iterations << " iterations left a relative residual of "
     << r.relative_residual << " against a tolerance of " << tol << " (" << c.levels.size()
     << " levels). Raise solver.linear.max_iterations, lower solver.linear.amg.strength_threshold, raise "
        "amg.smoother_degree, or use \"simplicial_ldlt\"";
  fail(SLX_CONVERGENCE_ERROR, os.str());
}

// ComplianceObjective::evaluate on the device-resident design d_x
// (src/topopt/Sensitivity.cpp:110-237); fills d_phys, d_fac, d_dfac, d_energy,
// d_dc (dc / d physical), d_dcx and d_dvx (pulled back through the filter).
slx_evaluation evaluate_device(Context& c, const slx_eval_options& o, bool need_gradients) {
  require_model(c);
  if (c.ncases == 0) fail(SLX_MODEL_ERROR, "the model has no load case");
  const long long ne = c.nelem;
  const size_t nf = c.free_dofs.size();
  hipStream_t s = c.stream;
  slx_evaluation out{};
  const auto t0 = Clock::now();
  if (c.eval_dots.n < (size_t)c.ncases + 1) c.eval_dots.alloc((size_t)c.ncases + 1);
  k::spmv_serial(c.dH, c.d_x.p, c.d_tmp.p, s);
  // the projection of the clamped filtered density (Sensitivity.cpp:120-129)
  const bool proj = o.beta > 0.0;
  if (proj) {
    if (!(o.eta > 0.0 && o.eta < 1.0) || !std::isfinite(o.beta))
      fail(SLX_CONFIG_ERROR, "projection needs beta > 0 and eta in (0, 1)");
    k::project((int)ne, c.d_tmp.p, o.beta, o.eta, c.d_tmp.p, c.d_dproj.p, s);
  }
  c.proj_beta = proj ? o.beta : 0.0;
  c.proj_eta = o.eta;
  k::simp((int)ne, c.d_tmp.p, o.penalty, o.emin_ratio, c.d_phys.p, c.d_fac.p, c.d_dfac.p, s);
  // coarse_reuse: keep the coarse operators of the last full setup while the
  // iteration count stays within 20 % (and 2) of the count right after it
  const int reuse = c.amg_opt.coarse_reuse;
  const bool lag = reuse > 0 && c.dim == 3 && matrix_free_enabled() && c.mf && c.amg_ready &&
                   c.k_source == Context::KSource::evaluation && c.levels.size() > 1 && c.coarse_lag < reuse &&
                   c.lag_base_its > 0 && c.lag_last_its <= c.lag_base_its + std::max(2, c.lag_base_its / 5);
  if (lag) {
    assemble(c, c.d_fac.p, false);
    c.k_stale = true;
    amg_refresh_level0(c);
  } else {
    assemble(c, c.d_fac.p);
    amg_setup(c, true);
  }
  c.k_source = Context::KSource::evaluation;
  out.assemble_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
  const auto t1 = Clock::now();
  std::vector<double> comp((size_t)c.ncases);
  int most_its = 0;
  // several load cases (Sensitivity.cpp:169-176 solves them one by one): one
  // block PCG drives them through each operator and V-cycle pass together,
  // each case's result equal to its own solve; the learned warm start, which
  // predicts one case at a time, keeps the case-by-case loop
  const bool block = c.ncases > 1 && block_cases_enabled() && !(c.ws && c.ws->opt.enable);
  std::vector<slx_pcg_result> block_res;
  if (block) {
    if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "AMG-CG: solve called before factorize");
    for (int l = 0; l < c.ncases; ++l)
      if (!o.warm_start || !c.have_u[(size_t)l]) k::fill((long long)nf, 0.0, c.d_u.p + (size_t)l * nf, s);
    pcg_block(c, c.ncases, c.d_loads_free.p, c.d_u.p, o.tolerance, o.max_iterations > 0 ? o.max_iterations : 1000,
              block_res);
    for (int l = 0; l < c.ncases; ++l) c.have_u[(size_t)l] = 1;
  }
  for (int l = 0; l < c.ncases; ++l) {
    slx_pcg_result r;
    if (block) {
      r = block_res[(size_t)l];
    } else {
      if (o.warm_start) warm_before_solve(c, l);
      r = solve_case(c, l, o.warm_start != 0, o.tolerance, o.max_iterations);
    }
    require_converged(c, r, o.tolerance);
    if (!block && o.warm_start) warm_after_solve(c, l);
    out.linear_iterations += r.iterations;
    most_its = std::max(most_its, r.iterations);
    // f_l . u_l, read with the volume at the end (no wait here)
    k::dot((long long)nf, c.d_loads_free.p + (size_t)l * nf, c.d_u.p + (size_t)l * nf, c.partials.p,
           c.eval_dots.p + l, s);
  }
  if (lag) {
    ++c.coarse_lag;
    c.lag_last_its = most_its;
  } else {
    c.coarse_lag = 0;
    c.lag_base_its = c.lag_last_its = most_its;
  }
  if (c.d_weights.n != c.weights.size()) c.d_weights.upload(c.weights, s);
  k::sensitivity(c.dim, (int)ne, c.nx, c.ny, c.d_ke0.p, c.d_reduced.p, c.d_u.p, (long long)nf, c.ncases, c.d_weights.p, c.d_fac.p,
                 c.d_dfac.p, c.d_energy.p, c.d_dc.p, s);
  k::dot(ne, c.d_phys.p, c.d_vol.p, c.partials.p, c.eval_dots.p + c.ncases, s);
  if (need_gradients) {
    if (proj) {
      // through the projection, then the density filter (Sensitivity.cpp:210-216)
      k::mul((int)ne, c.d_dc.p, c.d_dproj.p, c.d_tmp.p, s);
      k::spmv_serial(c.dHt, c.d_tmp.p, c.d_dcx.p, s);
      k::mul((int)ne, c.d_vol.p, c.d_dproj.p, c.d_tmp.p, s);
      k::spmv_serial(c.dHt, c.d_tmp.p, c.d_dvx.p, s);
    } else {
      k::spmv_serial(c.dHt, c.d_dc.p, c.d_dcx.p, s);
      k::copy(ne, htvol(c), c.d_dvx.p, s);
    }
  }
  warm_after_evaluate(c);
  // the load cases' f.u and the volume: one read
  double* hb = c.host_buffer((size_t)c.ncases + 1);
  SLX_HIP(hipMemcpyAsync(hb, c.eval_dots.p, ((size_t)c.ncases + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  sync(c);
  for (int l = 0; l < c.ncases; ++l) comp[(size_t)l] = hb[l];
  out.volume = hb[c.ncases];
  for (int l = 0; l < c.ncases; ++l) out.compliance += c.weights[(size_t)l] * comp[(size_t)l];
  c.case_compliance = comp;
  out.volume_fraction = out.volume / c.total_volume;
  out.solve_seconds = std::chrono::duration<double>(Clock::now() - t1).count();
  if (!std::isfinite(out.compliance)) fail(SLX_SOLVER_ERROR, "the compliance objective evaluated to a non-finite value");
  return out;
}

// optimality_criteria_update on device vectors d_x, d_dcx, d_dvx; the new
// design lands in d_tmp2 (src/topopt/OptimalityCriteria.cpp:40-149).
// target_override > 0 replaces the volume target; vol_eta >= 0 projects the
// trial volumes at that threshold; step_toward: when the move-limited box cannot
// reach the target, step to its nearest corner (OptimalityCriteria.cpp:82-97)
slx_oc_step oc_device(Context& c, const slx_oc_options& o, double target_override = 0.0, double vol_eta = -1.0,
                      bool step_toward = false) {
  const int ne = (int)c.nelem;
  hipStream_t s = c.stream;
  if (!(o.move_limit > 0.0 && o.move_limit <= 1.0)) fail(SLX_CONFIG_ERROR, "optimizer.move_limit must lie in (0, 1]");
  if (!(o.damping > 0.0 && o.damping <= 1.0)) fail(SLX_CONFIG_ERROR, "optimizer.damping must lie in (0, 1]");
  // both gradients finite: their sums of squares (read with the bracket's volumes)
  k::dot(ne, c.d_dcx.p, c.d_dcx.p, c.partials.p, c.scal.p + 11, s);
  k::dot(ne, c.d_dvx.p, c.d_dvx.p, c.partials.p, c.scal.p + 12, s);
  auto check_gradients = [&](const double* h) {
    if (!std::isfinite(h[0]) || !std::isfinite(h[1]))
      fail(SLX_SOLVER_ERROR, "optimality criteria received a non-finite gradient");
  };
  const double target = target_override > 0.0 ? target_override : c.volume_fraction * c.total_volume;
  const double trial_eta = vol_eta >= 0.0 ? vol_eta : c.proj_eta;
  // designs within [0, 1] filter into [0, 1] (rows of H are convex weights), so the
  // clamp of the filtered volume is inactive and V(x) = (H^T vol).x: no filter
  // product per bisection trial
  // (the projection makes V nonlinear in x: then each trial filters and projects)
  const bool bounded = c.unit_bounds && c.proj_beta == 0.0;
  const double* ocw = bounded ? htvol(c) : nullptr;
  if (bounded) c.oc_state.alloc(16);
  auto volume_at = [&](double lambda) {
    if (bounded) {
      k::oc_trial(ne, c.d_x.p, c.d_dcx.p, c.d_dvx.p, c.d_lower.p, c.d_upper.p, c.d_passive.p, o.move_limit,
                  o.damping, ocw, c.oc_state.p, 1, lambda, target, o.volume_tolerance, o.max_bisections,
                  c.d_tmp2.p, c.partials.p, s);
      return read_scalar(c, c.oc_state.p + 2);
    }
    k::oc_candidate(ne, c.d_x.p, c.d_dcx.p, c.d_dvx.p, c.d_lower.p, c.d_upper.p, c.d_passive.p, lambda,
                    o.move_limit, o.damping, c.d_tmp2.p, s);
    if (c.proj_beta > 0.0) {
      // the volume of the physical density the candidate would produce (TopologyOptimizer.cpp:573-580)
      k::spmv_serial(c.dH, c.d_tmp2.p, c.d_tmp.p, s);
      k::project(ne, c.d_tmp.p, c.proj_beta, trial_eta, c.d_tmp.p, nullptr, s);
      k::dot(ne, c.d_tmp.p, c.d_vol.p, c.partials.p, c.scal.p + 14, s);
    } else {
      k::filter_volume(c.dH, c.d_tmp2.p, c.d_vol.p, c.partials.p, c.scal.p + 14, s);
    }
    return read_scalar(c, c.scal.p + 14);
  };
  double lo = o.lambda_lower, hi = o.lambda_upper;
  double vlo, vhi;
  if (bounded) {
    // the volumes at both ends of the bracket and the gradient check in one read:
    // V(lo) moves to state[8] before the second trial
    k::oc_trial(ne, c.d_x.p, c.d_dcx.p, c.d_dvx.p, c.d_lower.p, c.d_upper.p, c.d_passive.p, o.move_limit,
                o.damping, ocw, c.oc_state.p, 1, lo, target, o.volume_tolerance, o.max_bisections, c.d_tmp2.p,
                c.partials.p, s);
    k::copy(1, c.oc_state.p + 2, c.oc_state.p + 8, s);
    k::oc_trial(ne, c.d_x.p, c.d_dcx.p, c.d_dvx.p, c.d_lower.p, c.d_upper.p, c.d_passive.p, o.move_limit,
                o.damping, ocw, c.oc_state.p, 1, hi, target, o.volume_tolerance, o.max_bisections, c.d_tmp2.p,
                c.partials.p, s);
    double* hb = c.host_buffer(16);
    SLX_HIP(hipMemcpyAsync(hb, c.scal.p + 11, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    SLX_HIP(hipMemcpyAsync(hb + 2, c.oc_state.p + 2, sizeof(double), hipMemcpyDeviceToHost, s));
    SLX_HIP(hipMemcpyAsync(hb + 3, c.oc_state.p + 8, sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    check_gradients(hb);
    vhi = hb[2];
    vlo = hb[3];
  } else {
    double* hb = c.host_buffer(16);
    SLX_HIP(hipMemcpyAsync(hb, c.scal.p + 11, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    check_gradients(hb);
    vlo = volume_at(lo);
    vhi = volume_at(hi);
  }
  int widen = 0;
  while (vlo < target && widen < 60) {
    lo *= 1.0e-3;
    vlo = volume_at(lo);
    ++widen;
  }
  while (vhi > target && widen < 120) {
    hi *= 1.0e3;
    vhi = volume_at(hi);
    ++widen;
  }
  if ((vlo < target || vhi > target) && step_toward) {
    const bool too_light = vlo < target;
    slx_oc_step step{};
    step.lambda = too_light ? lo : hi;
    step.achieved_volume = volume_at(step.lambda);  // leaves the candidate in d_tmp2
    step.volume_converged = 0;
    k::max_abs_diff(ne, c.d_tmp2.p, c.d_x.p, c.partials.p, c.scal.p + 13, s);
    step.max_change = read_scalar(c, c.scal.p + 13);
    return step;
  }
  if (vlo < target || vhi > target) {
    std::ostringstream os;
    os << "the optimality-criteria bisection cannot bracket the volume target " << target
       << " m^3: the move-limited box yields volumes in [" << vhi << ", " << vlo << "] m^3";
    fail(SLX_CONVERGENCE_ERROR, os.str());
  }
  slx_oc_step step{};
  if (bounded) {
    // the bisection on the device, checked from the host every few trials
    // state {lo, hi, 0, ...} set on the device (no host copy to wait for)
    k::fill(8, 0.0, c.oc_state.p, s);
    k::fill(1, lo, c.oc_state.p, s);
    k::fill(1, hi, c.oc_state.p + 1, s);
    k::oc_trial(ne, c.d_x.p, c.d_dcx.p, c.d_dvx.p, c.d_lower.p, c.d_upper.p, c.d_passive.p, o.move_limit, o.damping,
                ocw, c.oc_state.p, 1, lo, target, o.volume_tolerance, o.max_bisections, c.d_tmp2.p,
                c.partials.p, s);
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (o.max_bisections > 0) {
      // the first batch as long as the last update's bisection (trials after the
      // stop return at once), then 8 at a time; each batch ends with the largest
      // design change of the candidate it leaves, which is the step's once the
      // bisection has stopped
      int batch = std::max(8, (c.oc_last_bisections + 7) / 8 * 8);
      double change = 0.0;
      do {
        for (int k = 0; k < batch; ++k)
          k::oc_trial(ne, c.d_x.p, c.d_dcx.p, c.d_dvx.p, c.d_lower.p, c.d_upper.p, c.d_passive.p, o.move_limit,
                      o.damping, ocw, c.oc_state.p, 0, 0.0, target, o.volume_tolerance, o.max_bisections,
                      c.d_tmp2.p, c.partials.p, s);
        k::max_abs_diff(ne, c.d_tmp2.p, c.d_x.p, c.partials.p, c.scal.p + 13, s);
        double* hb = c.host_buffer(16);
        SLX_HIP(hipMemcpyAsync(hb, c.oc_state.p, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
        SLX_HIP(hipMemcpyAsync(hb + 8, c.scal.p + 13, sizeof(double), hipMemcpyDeviceToHost, s));
        sync(c);
        std::copy(hb, hb + 8, st);
        change = hb[8];
        batch = 8;
      } while (st[5] == 0.0);
      c.oc_last_bisections = (int)st[4];
      step.max_change = change;
    } else {
      c.oc_state.download(st, 8, s);
      sync(c);
      k::max_abs_diff(ne, c.d_tmp2.p, c.d_x.p, c.partials.p, c.scal.p + 13, s);
      step.max_change = read_scalar(c, c.scal.p + 13);
    }
    step.achieved_volume = st[2];
    step.lambda = st[3];
    step.bisections = (int)st[4];
    step.volume_converged = st[6] != 0.0;
    return step;
  }
  double volume = volume_at(lo);
  for (int it = 0; it < o.max_bisections; ++it) {
    const double mid = std::sqrt(lo * hi);
    volume = volume_at(mid);
    step.bisections = it + 1;
    step.lambda = mid;
    if (std::abs(volume - target) <= o.volume_tolerance * target) {
      step.volume_converged = 1;
      break;
    }
    if (volume > target) lo = mid; else hi = mid;
    if (hi / lo - 1.0 < 1.0e-15) {
      step.volume_converged = std::abs(volume - target) <= 1.0e-6 * std::max(target, 1.0e-30);
      break;
    }
  }
  step.achieved_volume = volume;
  k::max_abs_diff(ne, c.d_tmp2.p, c.d_x.p, c.partials.p, c.scal.p + 13, s);
  step.max_change = read_scalar(c, c.scal.p + 13);
  return step;
}

}  // namespace
}  // namespace slx

using namespace slx;

extern "C" {

slx_status slx_create(int device, slx_context** out) {
  if (!out) return SLX_CONFIG_ERROR;
  *out = nullptr;
  Context* c = new Context();
  c->device = device;
  try {
    SLX_HIP(hipSetDevice(device));
    (void)noalloc_scratch(device);  // the scratch array's address on this device, once per device
    SLX_HIP(hipStreamCreate(&c->stream));
    slx_amg_default_options(&c->amg_opt);  // evaluate and run_oc use these until slx_amg_setup
    c->scal.alloc(16);
    c->partials.alloc(4096);
  } catch (const Error&) {
    delete c;
    return SLX_DEVICE_ERROR;
  }
  *out = reinterpret_cast<slx_context*>(c);
  return SLX_OK;
}

void slx_destroy(slx_context* ctx) {
  Context* c = reinterpret_cast<Context*>(ctx);
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& pe : c->pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  c->pinned = nullptr;
  invalidate_graphs(*c);
  hipStream_t s = c->stream;
  delete c;
  if (s) (void)hipStreamDestroy(s);
}

const char* slx_last_error(const slx_context* ctx) {
  const Context* c = reinterpret_cast<const Context*>(ctx);
  return c ? c->last_error.c_str() : "null context";
}

void slx_amg_default_options(slx_amg_options* o) {
  o->strength_threshold = 0.02;
  o->max_levels = 12;
  o->coarse_size = 1500;
  o->smoother = 0;
  o->smoother_degree = 3;
  o->chebyshev_ratio = 30.0;
  o->prolongator_damping = 4.0 / 3.0;
  o->lanczos_steps = 12;
  o->coarse_pivot_tolerance = 1.0e-13;
  o->reuse_aggregates = 1;
  o->precision = 0;
  o->coarse_reuse = 0;
}

void slx_eval_default_options(slx_eval_options* o) {
  o->penalty = 3.0;
  o->emin_ratio = 1.0e-9;
  o->tolerance = 1.0e-8;
  o->max_iterations = 1000;
  o->warm_start = 1;
  o->beta = 0.0;
  o->eta = 0.5;
}

void slx_oc_default_options(slx_oc_options* o) {
  o->move_limit = 0.2;
  o->damping = 0.5;
  o->lambda_lower = 1.0e-12;
  o->lambda_upper = 1.0e12;
  o->max_bisections = 200;
  o->volume_tolerance = 1.0e-10;
}

slx_status slx_model_structured(slx_context* ctx, int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                                double E, double nu, double thickness) {
  return guarded(ctx, [&](Context& c) {
    if (dim != 2 && dim != 3) fail(SLX_CONFIG_ERROR, "dim must be 2 (Q4) or 3 (Hex8)");
    if (nx < 1 || ny < 1 || (dim == 3 && nz < 1)) fail(SLX_CONFIG_ERROR, "structured mesh needs positive element counts");
    if (!(lx > 0) || !(ly > 0) || (dim == 3 && !(lz > 0))) fail(SLX_CONFIG_ERROR, "structured mesh needs positive lengths");
    if (dim == 3 && thickness != 1.0) fail(SLX_CONFIG_ERROR, "a solid (Hex8) element has no thickness; leave it at 1");
    const long long nn = (long long)(nx + 1) * (ny + 1) * (dim == 3 ? nz + 1 : 1);
    if (nn * dim > 2147483647LL / 27) fail(SLX_CONFIG_ERROR, "the grid exceeds the int32 index range of K_ff");
    c.dim = dim;
    c.nx = nx;
    c.ny = ny;
    c.nz = dim == 3 ? nz : 0;
    c.lx = lx;
    c.ly = ly;
    c.lz = lz;
    c.E = E;
    c.nu = nu;
    c.thickness = thickness;
    build_structured_grid(c);
    c.fixed.assign((size_t)c.ndofs, 0);
    c.finalized = false;
    c.amg_ready = false;
    c.ncases = 0;
    c.loads.clear();
    c.H = HCsr();
    c.Ht = HCsr();
  });
}

slx_status slx_model_set_supports(slx_context* ctx, const unsigned char* fixed) {
  return guarded(ctx, [&](Context& c) {
    if (c.ndofs == 0) fail(SLX_MODEL_ERROR, "define the mesh first");
    for (long long d = 0; d < c.ndofs; ++d) c.fixed[(size_t)d] = fixed[d] ? 1 : 0;
    c.finalized = false;
    c.amg_ready = false;
  });
}

slx_status slx_model_set_loads(slx_context* ctx, int ncases, const double* loads, const double* weights) {
  return guarded(ctx, [&](Context& c) {
    if (ncases < 1) fail(SLX_CONFIG_ERROR, "at least one load case is needed");
    double sum = 0.0;
    for (int l = 0; l < ncases; ++l) sum += weights[l];
    if (!(sum > 0.0)) fail(SLX_CONFIG_ERROR, "load-case weights must sum to a positive number");
    c.ncases = ncases;
    c.loads.assign(loads, loads + (size_t)ncases * c.ndofs);
    c.weights.resize((size_t)ncases);
    for (int l = 0; l < ncases; ++l) c.weights[(size_t)l] = weights[l] / sum;
    c.d_weights.upload(c.weights, c.stream);
    if (c.finalized) {
      const size_t nf = c.free_dofs.size();
      std::vector<double> lf((size_t)ncases * nf);
      for (int l = 0; l < ncases; ++l)
        for (size_t i = 0; i < nf; ++i) lf[(size_t)l * nf + i] = c.loads[(size_t)l * c.ndofs + c.free_dofs[i]];
      c.d_loads_free.upload(lf, c.stream);
      c.d_u.alloc((size_t)ncases * nf);
      c.have_u.assign((size_t)ncases, 0);
      sync(c);
    }
  });
}

slx_status slx_model_finalize(slx_context* ctx) {
  return guarded(ctx, [&](Context& c) {
    if (c.ndofs == 0) fail(SLX_MODEL_ERROR, "define the mesh first");
    c.mf = false;
    // DofManager::rebuild (src/fem/DofManager.cpp:58-71)
    c.free_dofs.clear();
    c.reduced.assign((size_t)c.ndofs, -1);
    for (long long d = 0; d < c.ndofs; ++d) {
      if (c.fixed[(size_t)d]) continue;
      c.reduced[(size_t)d] = (int)c.free_dofs.size();
      c.free_dofs.push_back((int)d);
    }
    if (c.free_dofs.empty()) fail(SLX_MODEL_ERROR, "every DOF is fixed");
    const std::vector<double> xe = element0_coords(c);
    c.ke0 = c.dim == 3 ? hex8_ke0(xe.data(), c.E, c.nu) : quad4_ke0(xe.data(), c.E, c.nu, c.thickness);
    c.volumes.assign((size_t)c.nelem, element_volume(c));
    c.total_volume = 0.0;
    for (double v : c.volumes) c.total_volume += v;
    c.htvol_valid = false;
    HCsr pat = structured_kff_pattern(c);
    c.K.upload_pattern(pat, c.stream);
    ++c.k_pattern_version;
    keep_host_pattern(c, std::move(pat));
    c.d_free.upload(c.free_dofs, c.stream);
    c.d_reduced.upload(c.reduced, c.stream);
    if (c.dim == 3) {
      std::vector<int> map;
      if (!kmf::node_map(c.reduced, map)) fail(SLX_MODEL_ERROR, "the model has more than 2^28 free DOFs");
      c.d_nodemap.upload(map, c.stream);
    }
    {
      const long long nnx = c.nx + 1, nny = c.ny + 1;
      const bool fits = c.dim == 3 ? (nnx <= 1024 && nny <= 1024 && c.nz + 1 <= 1024) : (nnx <= 32768 && nny <= 32768);
      c.d_dofxyz.release();
      if (fits) {
        std::vector<unsigned> xyz(c.free_dofs.size());
        for (size_t f = 0; f < xyz.size(); ++f) {
          const long long g = c.free_dofs[f], node = g / c.dim, comp = g % c.dim;
          const unsigned x = (unsigned)(node % nnx), y = (unsigned)((node / nnx) % nny);
          xyz[f] = c.dim == 3 ? (x | y << 10 | (unsigned)(node / (nnx * nny)) << 20 | (unsigned)comp << 30)
                              : (x | y << 15 | (unsigned)comp << 30);
        }
        c.d_dofxyz.upload(xyz, c.stream);
      }
    }
    c.d_ke0.upload(c.ke0, c.stream);
    c.d_vol.upload(c.volumes, c.stream);
    set_layout_from_model(c);
    ensure_workspace(c);
    c.finalized = true;
    c.amg_ready = false;
    c.k_source = Context::KSource::none;
    if (c.ncases > 0) {
      const size_t nf = c.free_dofs.size();
      std::vector<double> lf((size_t)c.ncases * nf);
      for (int l = 0; l < c.ncases; ++l)
        for (size_t i = 0; i < nf; ++i) lf[(size_t)l * nf + i] = c.loads[(size_t)l * c.ndofs + c.free_dofs[i]];
      c.d_loads_free.upload(lf, c.stream);
      c.d_u.alloc((size_t)c.ncases * nf);
      c.have_u.assign((size_t)c.ncases, 0);
    }
    sync(c);
  });
}

slx_status slx_model_sizes(const slx_context* ctx, long long* out) {
  const Context* c = reinterpret_cast<const Context*>(ctx);
  if (!c) return SLX_CONFIG_ERROR;
  out[0] = c->nnodes;
  out[1] = c->nelem;
  out[2] = c->ndofs;
  out[3] = (long long)c->free_dofs.size();
  out[4] = c->K.nnz;
  out[5] = c->ncases;
  return SLX_OK;
}

slx_status slx_ke0_get(const slx_context* ctx, double* ke) {
  const Context* c = reinterpret_cast<const Context*>(ctx);
  if (!c || c->ke0.empty()) return SLX_MODEL_ERROR;
  std::memcpy(ke, c->ke0.data(), c->ke0.size() * sizeof(double));
  return SLX_OK;
}

slx_status slx_free_dofs_get(const slx_context* ctx, int* free_dofs) {
  const Context* c = reinterpret_cast<const Context*>(ctx);
  if (!c || !c->finalized) return SLX_MODEL_ERROR;
  std::memcpy(free_dofs, c->free_dofs.data(), c->free_dofs.size() * sizeof(int));
  return SLX_OK;
}

slx_status slx_assemble(slx_context* ctx, const double* scale) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    const size_t ne = (size_t)c.nelem;
    std::vector<double> ones;
    const double* h = scale;
    if (!h) {
      ones.assign(ne, 1.0);
      h = ones.data();
    }
    double mn = std::numeric_limits<double>::max();
    for (size_t e = 0; e < ne; ++e) {
      if (!std::isfinite(h[e])) fail(SLX_MODEL_ERROR, "stiffness scale vector contains NaN or Inf");
      mn = std::min(mn, h[e]);
    }
    if (mn < 0.0) fail(SLX_MODEL_ERROR, "stiffness scale vector has a negative entry");
    c.d_scale.upload(h, ne, c.stream);
    assemble(c, c.d_scale.p);
    c.k_source = Context::KSource::assembled;
    c.flush_timers();
    sync(c);
  });
}

slx_status slx_kff_apply(slx_context* ctx, const double* x, double* y, int* matrix_free) {
  return guarded(ctx, [&](Context& c) {
    if (c.K.rows == 0) fail(SLX_SOLVER_ERROR, "K_ff is not assembled");
    const size_t n = (size_t)c.K.rows;
    DBuf<double> dx, dy;
    dx.upload(x, n, c.stream);
    dy.alloc(n);
    op_apply(c, dx.p, dy.p);
    dy.download(y, n, c.stream);
    sync(c);
    if (matrix_free) *matrix_free = c.mf ? 1 : 0;
  });
}

slx_status slx_kff_get(const slx_context* ctx, int* ptr, int* col, double* val) {
  return guarded(const_cast<slx_context*>(ctx), [&](Context& c) {
    require_model(c);
    if (c.k_stale) {  // values of the factors the operator uses
      k::assemble_structured(c.dim, c.K, c.d_free.p, c.mf_scale.p, c.d_ke0.p, c.nx, c.ny, c.nz, c.stream,
                             c.d_dofxyz.p, c.d_nodemap.p);
      c.k_stale = false;
    }
    if (ptr) c.K.ptr.download(ptr, (size_t)c.K.rows + 1, c.stream);
    if (col) c.K.col.download(col, (size_t)c.K.nnz, c.stream);
    if (val) c.K.val.download(val, (size_t)c.K.nnz, c.stream);
    sync(c);
  });
}

slx_status slx_amg_setup(slx_context* ctx, const slx_amg_options* opt) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    slx_amg_options o;
    slx_amg_default_options(&o);
    if (opt) o = *opt;
    if (!(o.strength_threshold >= 0.0 && o.strength_threshold < 1.0)) fail(SLX_CONFIG_ERROR, "multigrid strength_threshold must lie in [0, 1)");
    if (o.max_levels < 1 || o.max_levels > 16) fail(SLX_CONFIG_ERROR, "multigrid max_levels must lie in [1, 16]");
    if (o.coarse_size < 1) fail(SLX_CONFIG_ERROR, "multigrid coarse_size must be positive");
    if (o.smoother != 0 && o.smoother != 1)
      fail(SLX_CONFIG_ERROR, "multigrid smoother must be 0 (chebyshev) or 1 (gauss_seidel)");
    if (o.smoother_degree < 1) fail(SLX_CONFIG_ERROR, "multigrid smoother_degree must be at least 1");
    if (o.smoother == 1 && o.coarse_reuse > 0)
      fail(SLX_CONFIG_ERROR, "multigrid coarse_reuse needs the Chebyshev smoother (the Gauss-Seidel sweeps read the "
                             "assembled level-0 operator, which a reused evaluation does not assemble)");
    if (!(o.chebyshev_ratio > 1.0)) fail(SLX_CONFIG_ERROR, "multigrid chebyshev_ratio must exceed 1");
    if (!(o.prolongator_damping > 0.0 && o.prolongator_damping < 2.0)) fail(SLX_CONFIG_ERROR, "multigrid prolongator_damping must lie in (0, 2)");
    if (o.lanczos_steps < 2) fail(SLX_CONFIG_ERROR, "multigrid lanczos_steps must be at least 2");
    if (o.coarse_reuse < 0) fail(SLX_CONFIG_ERROR, "multigrid coarse_reuse must be 0 or positive");
    if (o.precision < 0 || o.precision > 2)
      fail(SLX_CONFIG_ERROR, "multigrid precision must be 0 (FP64), 1 (FP32 cycle) or 2 (FP32 cycle, bfloat16 coarse operators)");
    if (c.amg_ready && (o.precision != c.amg_opt.precision || o.max_levels != c.amg_opt.max_levels ||
                        o.coarse_size != c.amg_opt.coarse_size || o.strength_threshold != c.amg_opt.strength_threshold))
      c.amg_ready = false;
    c.amg_opt = o;
    amg_setup(c);
  });
}

slx_status slx_factorize_csr(slx_context* ctx, int n, const int* ptr, const int* col, const double* val, int dim,
                             int ndpn, int nnodes, const double* coords, const int* unknowns,
                             const slx_amg_options* opt) {
  return guarded(ctx, [&](Context& c) {
    if (n <= 0) fail(SLX_SOLVER_ERROR, "multigrid: the system matrix must be square and non-empty");
    if (!ptr || !col || !val || !unknowns) fail(SLX_CONFIG_ERROR, "factorize: null matrix or layout array");
    if (dim != 2 && dim != 3) fail(SLX_CONFIG_ERROR, "factorize: the layout dimension must be 2 or 3");
    if (nnodes < 1 || !coords) fail(SLX_CONFIG_ERROR, "factorize: the layout needs its node coordinates");
    // the host checks every index the device and the aggregation will follow
    if (ptr[0] != 0) fail(SLX_CONFIG_ERROR, "factorize: CSR row pointers must start at 0");
    for (int i = 0; i < n; ++i)
      if (ptr[i + 1] < ptr[i]) fail(SLX_CONFIG_ERROR, "factorize: CSR row pointers must not decrease");
    HCsr h;
    h.rows = h.cols = n;
    h.ptr.assign(ptr, ptr + n + 1);
    h.col.assign(col, col + ptr[n]);
    h.val.assign(val, val + ptr[n]);
    unsigned long long hash = 1469598103934665603ULL;  // FNV-1a over the pattern
    auto mix = [&](int v) { hash = (hash ^ (unsigned long long)(unsigned)v) * 1099511628211ULL; };
    for (int i = 0; i < n; ++i) {
      const int k0 = h.ptr[(size_t)i], k1 = h.ptr[(size_t)i + 1];
      mix(k1);
      if (k1 > k0 && (h.col[(size_t)k0] < 0 || h.col[(size_t)k1 - 1] >= n))
        fail(SLX_CONFIG_ERROR, "factorize: a CSR column index lies outside the matrix");
      for (int k2 = k0; k2 < k1; ++k2) {
        mix(h.col[(size_t)k2]);
        if (k2 > k0 && h.col[(size_t)k2] <= h.col[(size_t)k2 - 1])
          fail(SLX_SOLVER_ERROR, "CSR rows must be sorted by column without duplicates");
      }
    }
    const int per_node = ndpn > 0 ? ndpn : dim;
    for (int i = 0; i < n; ++i)
      if (unknowns[i] < 0 || (long long)unknowns[i] >= (long long)nnodes * per_node)
        fail(SLX_CONFIG_ERROR, "factorize: a layout unknown lies outside the layout's nodes");
    const bool same_unknowns =
        c.layout_unknowns.size() == (size_t)n && std::equal(unknowns, unknowns + n, c.layout_unknowns.begin());
    // only the values change when the pattern is the one K already holds
    const bool same_pattern = c.K.rows == n && c.K.nnz == h.nnz() && same_unknowns && c.k_pattern_hash == hash &&
                              c.k_source == Context::KSource::caller;
    if (same_pattern) {
      c.K.upload_values(h.val.data(), c.stream);
    } else {
      c.K.upload(h, c.stream);
      ++c.k_pattern_version;
      keep_host_pattern(c, std::move(h));
    }
    c.k_pattern_hash = hash;
    c.mf = false;  // a caller's matrix: products use it as given
    // the structured model's K_ff pattern, solutions and adjoint slots no longer
    // describe K: the structured entry points refuse until slx_model_finalize
    c.k_source = Context::KSource::caller;
    c.finalized = false;
    c.have_u.assign(c.have_u.size(), 0);
    c.adjoint_slots.clear();
    c.layout_dim = dim;
    c.layout_ndpn = ndpn > 0 ? ndpn : dim;
    c.layout_nnodes = nnodes;
    c.layout_coords.assign(coords, coords + (size_t)nnodes * dim);
    if (!same_unknowns) {
      c.layout_unknowns.assign(unknowns, unknowns + n);
      ++c.layout_version;
    }
    slx_amg_options o;
    slx_amg_default_options(&o);
    if (opt) o = *opt;
    c.amg_opt = o;
    amg_setup(c);
  });
}

slx_status slx_solve(slx_context* ctx, const double* b, double* x, double tol, int max_iterations,
                     slx_pcg_result* res) {
  return guarded(ctx, [&](Context& c) {
    if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "AMG-CG: solve called before factorize");
    const size_t n = (size_t)c.K.rows;
    c.bw.upload(b, n, c.stream);
    DBuf<double> xd;
    xd.upload(x, n, c.stream);
    slx_pcg_result r = pcg(c, c.bw.p, xd.p, tol, max_iterations > 0 ? max_iterations : 1000);
    xd.download(x, n, c.stream);
    sync(c);
    if (res) *res = r;
  });
}

slx_status slx_solve_segmented(slx_context* ctx, int nseg, const int* seg_ptr, const double* b, double* x, double tol,
                               int max_iterations, slx_pcg_result* results) {
  return guarded(ctx, [&](Context& c) {
    if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "AMG-CG: solve called before factorize");
    const int n = c.K.rows;
    if (nseg < 1 || !seg_ptr || seg_ptr[0] != 0 || seg_ptr[nseg] != n)
      fail(SLX_CONFIG_ERROR, "segmented solve: seg_ptr must run from 0 to the matrix size");
    for (int g = 0; g < nseg; ++g)
      if (seg_ptr[g + 1] <= seg_ptr[g]) fail(SLX_CONFIG_ERROR, "segmented solve: segments must be non-empty");
    const std::vector<int> sp(seg_ptr, seg_ptr + nseg + 1);
    c.bw.upload(b, (size_t)n, c.stream);
    DBuf<double> xd;
    xd.upload(x, (size_t)n, c.stream);
    std::vector<slx_pcg_result> r;
    seg_pcg(c, sp, c.bw.p, xd.p, tol, max_iterations > 0 ? max_iterations : 1000, r);
    xd.download(x, (size_t)n, c.stream);
    sync(c);
    if (results) std::copy(r.begin(), r.end(), results);
  });
}

slx_status slx_solve_case(slx_context* ctx, int load_case, int warm_start, double tol, int max_iterations,
                          slx_pcg_result* res) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    if (load_case < 0 || load_case >= c.ncases) fail(SLX_CONFIG_ERROR, "load case index out of range");
    slx_pcg_result r = solve_case(c, load_case, warm_start != 0, tol, max_iterations);
    if (res) *res = r;
  });
}

slx_status slx_amg_time_smoother(slx_context* ctx, int level, int reps, double* ms_per_step) {
  return guarded(ctx, [&](Context& c) {
    require_amg(c);
    if (level < 0 || level >= (int)c.levels.size()) fail(SLX_CONFIG_ERROR, "no such multigrid level");
    if (reps < 1) fail(SLX_CONFIG_ERROR, "the smoother timing needs at least one repetition");
    finish_coarse(c);
    *ms_per_step = time_smoother(c, level, reps);
  });
}

slx_status slx_amg_apply(slx_context* ctx, const double* r, double* z) {
  return guarded(ctx, [&](Context& c) {
    require_amg(c);
    const size_t n = (size_t)c.K.rows;
    c.bw.upload(r, n, c.stream);
    vcycle(c, c.bw.p, c.xw.p);
    c.xw.download(z, n, c.stream);
    sync(c);
  });
}

slx_status slx_amg_stats_get(const slx_context* ctx, slx_amg_stats* out) {
  const Context* c = reinterpret_cast<const Context*>(ctx);
  if (!c || !c->amg_ready) return SLX_SOLVER_ERROR;
  std::memset(out, 0, sizeof(*out));
  out->levels = (int)c->levels.size();
  out->reused_aggregates = c->reused ? 1 : 0;
  out->coarse_pivot_ratio = c->coarse_pivot_ratio;
  out->near_null_space_dimension = c->nullspace_dim;
  out->setup_seconds = c->setup_seconds;
  out->check_seconds = c->phase_seconds[0];
  out->spectral_seconds = c->phase_seconds[1];
  out->aggregation_seconds = c->phase_seconds[2];
  out->smoothing_seconds = c->phase_seconds[3];
  out->galerkin_seconds = c->phase_seconds[4];
  out->coarse_seconds = c->phase_seconds[5];
  out->tentative_seconds = c->phase_seconds[6];
  for (const Level& lv : c->levels) out->node_product_levels += lv.node_ok ? 1 : 0;
  {
    // Multigrid.cpp:1089-1100
    double stored = 0.0;
    for (size_t l = 0; l < c->levels.size(); ++l) {
      const Level& lv = c->levels[l];
      if (l > 0) stored += (double)lv.A.nnz;
      if (!lv.has_next) continue;
      stored += (double)(lv.P.nnz + lv.R.nnz + lv.AP.nnz + lv.RAW.nnz + lv.Ptent.nnz);
    }
    const double nc = (double)c->levels.back().n;
    if (c->levels.back().n <= 4000) stored += nc * nc;
    out->storage_nonzeros = (long long)std::min(stored, 2.0e9);
  }
  double nnz_sum = 0, rows_sum = 0;
  for (size_t l = 0; l < c->levels.size() && l < 16; ++l) {
    const Level& lv = c->levels[l];
    out->unknowns[l] = lv.n;
    out->nonzeros[l] = l == 0 ? c->K.nnz : lv.A.nnz;
    out->aggregates[l] = l + 1 < c->levels.size() ? c->levels[l + 1].num_nodes : 0;
    out->lambda_max[l] = lv.lambda_max;
    nnz_sum += (double)out->nonzeros[l];
    rows_sum += (double)out->unknowns[l];
  }
  out->operator_complexity = nnz_sum / (double)out->nonzeros[0];
  out->grid_complexity = rows_sum / (double)out->unknowns[0];
  return SLX_OK;
}

slx_status slx_amg_level_info(const slx_context* ctx, int level, int which, long long* info) {
  return guarded(const_cast<slx_context*>(ctx), [&](Context& c) {
    if (!c.amg_ready || level < 0 || level >= (int)c.levels.size()) fail(SLX_SOLVER_ERROR, "no such multigrid level");
    Level& lv = c.levels[(size_t)level];
    if (which == 8) {  // the level's operator: nodes, shared column-list length (0: none), lanes per row
      const DCsr& a = level == 0 ? c.K : lv.A;
      info[0] = (long long)lv.node_ptr.size() - 1;
      info[1] = a.cbase.p ? (long long)a.ncol.n : 0;
      info[2] = a.lanes;
      return;
    }
    if (which == 9) {  // level 0's matrix-free tile (nodes per direction) for the V-cycle's precision; 0s: assembled
      info[0] = info[1] = info[2] = 0;
      if (level == 0 && c.mf) {
        int d[3];
        kmf::tile_shape(c.mf_tile[c.amg_opt.precision >= 1 ? 1 : 0], d);
        info[0] = d[0], info[1] = d[1], info[2] = d[2];
      }
      return;
    }
    if (which == 6 || which == 7) {  // row plans of A P and R (A P): C nodes, runs, entries
      const RowPlan& pl = which == 6 ? lv.plan_ap : lv.plan_raw;
      info[0] = pl.nnodes;
      info[1] = pl.runs;
      info[2] = pl.entries;
      return;
    }
    const DCsr* m = which == 0   ? (level == 0 ? &c.K : &lv.A)
                    : which == 1 ? &lv.P
                    : which == 2 ? &lv.Ptent
                    : which == 3 ? &lv.R
                    : which == 4 ? &lv.AP
                                 : &lv.RAW;
    info[0] = m->rows;
    info[1] = m->cols;
    info[2] = m->nnz;
  });
}

slx_status slx_amg_level_get(const slx_context* ctx, int level, int which, int* ptr, int* col, double* val) {
  return guarded(const_cast<slx_context*>(ctx), [&](Context& c) {
    if (!c.amg_ready || level < 0 || level >= (int)c.levels.size()) fail(SLX_SOLVER_ERROR, "no such multigrid level");
    Level& lv = c.levels[(size_t)level];
    const DCsr* m = which == 0 ? (level == 0 ? &c.K : &lv.A) : which == 1 ? &lv.P : which == 2 ? &lv.Ptent : &lv.R;
    m->ptr.download(ptr, (size_t)m->rows + 1, c.stream);
    m->col.download(col, (size_t)m->nnz, c.stream);
    m->val.download(val, (size_t)m->nnz, c.stream);
    sync(c);
  });
}

slx_status slx_amg_aggregates_get(const slx_context* ctx, int level, int* aggregates) {
  return guarded(const_cast<slx_context*>(ctx), [&](Context& c) {
    if (!c.amg_ready || level < 0 || level >= (int)c.levels.size()) fail(SLX_SOLVER_ERROR, "no such multigrid level");
    const std::vector<int>& a = c.levels[(size_t)level].aggregate;
    std::memcpy(aggregates, a.data(), a.size() * sizeof(int));
  });
}

slx_status slx_solution_get(const slx_context* ctx, int load_case, double* u) {
  return guarded(const_cast<slx_context*>(ctx), [&](Context& c) {
    if (load_case < 0 || load_case >= c.ncases) fail(SLX_CONFIG_ERROR, "load case index out of range");
    const size_t nf = c.free_dofs.size();
    SLX_HIP(hipMemcpyAsync(u, c.d_u.p + (size_t)load_case * nf, sizeof(double) * nf, hipMemcpyDeviceToHost, c.stream));
    sync(c);
  });
}

slx_status slx_design_set(slx_context* ctx, double volume_fraction, const double* lower, const double* upper,
                          const unsigned char* passive) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    if (!(volume_fraction > 0.0 && volume_fraction <= 1.0)) fail(SLX_CONFIG_ERROR, "volume fraction must lie in (0, 1]");
    const size_t ne = (size_t)c.nelem;
    c.volume_fraction = volume_fraction;
    c.lower.assign(lower, lower + ne);
    c.upper.assign(upper, upper + ne);
    c.passive.assign(passive, passive + ne);
    c.design_dirty = true;
    ensure_topopt_buffers(c);
    sync(c);
  });
}

slx_status slx_filter_build(slx_context* ctx, double radius) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    if (radius > 0.0) build_filter(c, radius); else identity_filter(c);
    upload_filter(c);
    ensure_topopt_buffers(c);
    sync(c);
  });
}

slx_status slx_filter_apply(slx_context* ctx, int adjoint, const double* in, double* out) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    ensure_topopt_buffers(c);
    const size_t ne = (size_t)c.nelem;
    c.d_tmp.upload(in, ne, c.stream);
    k::spmv_serial(adjoint ? c.dHt : c.dH, c.d_tmp.p, c.d_tmp2.p, c.stream);
    c.d_tmp2.download(out, ne, c.stream);
    sync(c);
  });
}

slx_status slx_evaluate(slx_context* ctx, const double* x, const slx_eval_options* opt, slx_evaluation* out,
                        double* dc_dx, double* dv_dx, double* energy) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    ensure_topopt_buffers(c);
    slx_eval_options o;
    slx_eval_default_options(&o);
    if (opt) o = *opt;
    const size_t ne = (size_t)c.nelem;
    for (size_t e = 0; e < ne; ++e)
      if (!std::isfinite(x[e])) fail(SLX_CONFIG_ERROR, "design vector contains NaN or Inf");
    c.d_x.upload(x, ne, c.stream);
    const slx_evaluation ev = evaluate_device(c, o, dc_dx != nullptr || dv_dx != nullptr);
    if (dc_dx) c.d_dcx.download(dc_dx, ne, c.stream);
    if (dv_dx) c.d_dvx.download(dv_dx, ne, c.stream);
    if (energy) c.d_energy.download(energy, ne, c.stream);
    sync(c);
    if (out) *out = ev;
  });
}

slx_status slx_oc_update(slx_context* ctx, const double* x, const double* dc_dx, const double* dv_dx,
                         const slx_oc_options* opt, double* x_new, slx_oc_step* step) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    ensure_topopt_buffers(c);
    slx_oc_options o;
    slx_oc_default_options(&o);
    if (opt) o = *opt;
    const size_t ne = (size_t)c.nelem;
    c.d_x.upload(x, ne, c.stream);
    c.d_dcx.upload(dc_dx, ne, c.stream);
    c.d_dvx.upload(dv_dx, ne, c.stream);
    const slx_oc_step st = oc_device(c, o);
    c.d_tmp2.download(x_new, ne, c.stream);
    sync(c);
    if (step) *step = st;
  });
}

slx_status slx_run_oc(slx_context* ctx, const slx_run_options* opt, double* x, double* history,
                      slx_run_result* result) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    ensure_topopt_buffers(c);
    if (!opt) fail(SLX_CONFIG_ERROR, "run options are required");
    const slx_run_options& o = *opt;
    if (o.max_iterations < 0) fail(SLX_CONFIG_ERROR, "max_iterations must not be negative");
    const auto t_total = Clock::now();
    const int ne = (int)c.nelem;
    c.d_x.upload(x, (size_t)ne, c.stream);
    k::clamp(ne, c.d_x.p, c.d_lower.p, c.d_upper.p, c.stream);
    std::vector<double> chist;
    // beta continuation (TopologyOptimizer.cpp:17-46; ProjectionOptions::validate, Projection.cpp:10-56)
    const slx_projection_options& pr = o.projection;
    if (pr.enabled) {
      std::ostringstream os;
      if (!(pr.eta > 0.0 && pr.eta < 1.0))
        os << "topology.projection.eta must lie in (0, 1), got " << pr.eta;
      else if (!(pr.beta_start > 0.0) || !std::isfinite(pr.beta_start))
        os << "topology.projection.beta_start must be positive, got " << pr.beta_start;
      else if (!(pr.beta_max >= pr.beta_start) || !std::isfinite(pr.beta_max))
        os << "topology.projection.beta_max (" << pr.beta_max << ") must be at least beta_start (" << pr.beta_start
           << ")";
      else if (pr.beta_max > 1024.0)
        os << "topology.projection.beta_max = " << pr.beta_max << " is beyond any useful sharpness";
      else if (!(pr.beta_factor > 1.0))
        os << "topology.projection.beta_factor must exceed 1, got " << pr.beta_factor;
      else if (pr.beta_interval < 1)
        os << "topology.projection.beta_interval must be at least 1, got " << pr.beta_interval;
      else if (pr.robust && !(pr.robust_delta > 0.0 && pr.eta - pr.robust_delta > 0.0 && pr.eta + pr.robust_delta < 1.0))
        os << "topology.projection.robust_delta = " << pr.robust_delta
           << " must be positive and keep eta +- delta inside (0, 1) (eta = " << pr.eta << ")";
      else if (pr.robust && pr.robust_volume_interval < 1)
        os << "topology.projection.robust_volume_interval must be at least 1, got " << pr.robust_volume_interval;
      if (!os.str().empty()) fail(SLX_CONFIG_ERROR, os.str());
    }
    auto beta_of_stage = [&](int k) {
      double b = pr.beta_start;
      for (int i = 0; i < k && b < pr.beta_max; ++i) b *= pr.beta_factor;
      return std::min(b, pr.beta_max);
    };
    int stages = 1;
    if (pr.enabled) {
      double b = pr.beta_start;
      while (b < pr.beta_max && stages < 64) {
        b *= pr.beta_factor;
        ++stages;
      }
    }
    int stage = 0, stage_start = 1;
    double beta = pr.enabled ? beta_of_stage(0) : 0.0;
    // robust formulation (TopologyOptimizer.cpp:60-80,566-640): objective on the
    // eroded design, volume on the dilated one against a rescaled target
    const bool robust = pr.enabled && pr.robust;
    const double objective_eta = robust ? pr.eta + pr.robust_delta : pr.eta;
    const double dilated_eta = pr.eta - pr.robust_delta;
    const double domain = c.total_volume;
    double dilated_target = c.volume_fraction * domain;
    DBuf<double> rfilt, rtmp, rdp, rs;
    if (robust) {
      rfilt.alloc((size_t)ne);
      rtmp.alloc((size_t)ne);
      rdp.alloc((size_t)ne);
      rs.alloc(2);
    }
    // gray_level (src/topopt/TopologyOptimizer.cpp:106-113): 4/n sum rho (1 - rho)
    auto grey = [&]() {
      k::grey_sum(ne, c.d_phys.p, c.partials.p, c.scal.p + 10, c.stream);
      return 4.0 * read_scalar(c, c.scal.p + 10) / double(std::max(ne, 1));
    };
    slx_run_result res{};
    res.stop_reason = SLX_STOP_ITERATION_CAP;
    res.final_objective_change = std::numeric_limits<double>::infinity();
    int iteration;
    for (iteration = 1; iteration <= o.max_iterations; ++iteration) {
      const auto t_iter = Clock::now();
      // penalty continuation (src/topopt/TopologyOptimizer.cpp:245-253)
      slx_eval_options eo = o.eval;
      if (o.continuation_steps > 1) {
        const int stage = std::min((iteration - 1) / std::max(o.continuation_iterations, 1), o.continuation_steps - 1);
        const double t = double(stage) / double(o.continuation_steps - 1);
        eo.penalty = o.penalty_start + t * (o.eval.penalty - o.penalty_start);
      }
      eo.beta = beta;
      eo.eta = objective_eta;
      const slx_evaluation ev = evaluate_device(c, eo, true);
      res.linear_solves += c.ncases;
      res.linear_iterations += ev.linear_iterations;
      double v_dil = 0.0, v_int = 0.0;
      if (robust) {
        // robust_volumes: the dilated volume and its gradient (replacing dv/dx), the blueprint volume
        hipStream_t s = c.stream;
        k::spmv_serial(c.dH, c.d_x.p, rfilt.p, s);
        k::project(ne, rfilt.p, beta, dilated_eta, rtmp.p, rdp.p, s);
        k::dot(ne, rtmp.p, c.d_vol.p, c.partials.p, rs.p, s);
        k::mul(ne, c.d_vol.p, rdp.p, rtmp.p, s);
        k::spmv_serial(c.dHt, rtmp.p, c.d_dvx.p, s);
        k::project(ne, rfilt.p, beta, pr.eta, rtmp.p, nullptr, s);
        k::dot(ne, rtmp.p, c.d_vol.p, c.partials.p, rs.p + 1, s);
        double* hb = c.host_buffer(16);
        SLX_HIP(hipMemcpyAsync(hb, rs.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        sync(c);
        v_dil = hb[0];
        v_int = hb[1];
        if ((iteration - 1) % pr.robust_volume_interval == 0 && v_int > 0.0)
          dilated_target = c.volume_fraction * domain * v_dil / v_int;
      }
      const slx_oc_step st = robust ? oc_device(c, o.oc, dilated_target, dilated_eta, true) : oc_device(c, o.oc);
      const double gray = grey();
      if (history) {
        double* row = history + (size_t)(iteration - 1) * SLX_HISTORY_COLUMNS;
        row[0] = iteration;
        row[1] = eo.penalty;
        row[2] = ev.compliance;
        row[3] = ev.volume;
        row[4] = ev.volume_fraction;
        row[5] = st.max_change;
        row[6] = st.lambda;
        row[7] = gray;
        row[8] = st.bisections;
        row[9] = st.volume_converged ? 1.0 : 0.0;
        row[10] = std::chrono::duration<double>(Clock::now() - t_iter).count();
        row[11] = ev.linear_iterations;
        row[12] = beta;
        row[13] = row[14] = row[15] = 0.0;
        if (robust) {
          row[3] = v_int;
          row[4] = v_int / domain;
          row[13] = ev.volume_fraction;
          row[14] = v_dil / domain;
          row[15] = dilated_target / domain;
        }
      }
      chist.push_back(ev.compliance);
      std::swap(c.d_x, c.d_tmp2);
      k::clamp(ne, c.d_x.p, c.d_lower.p, c.d_upper.p, c.stream);
      // convergence (src/topopt/TopologyOptimizer.cpp:672-701): only the final
      // penalty stage may stop the loop
      const bool at_final = o.continuation_steps <= 1 || iteration >= o.continuation_iterations * (o.continuation_steps - 1);
      const bool change_ok = st.max_change <= o.change_tolerance;
      double spread = std::numeric_limits<double>::infinity();
      const size_t w = (size_t)std::max(o.objective_window, 1);
      if (chist.size() > w) {
        auto mm = std::minmax_element(chist.end() - (long)(w + 1), chist.end());
        spread = (*mm.second - *mm.first) / std::max(std::abs(chist.back()), 1.0e-30);
      }
      const bool objective_ok = o.objective_tolerance > 0.0 && spread <= o.objective_tolerance;
      res.final_design_change = st.max_change;
      res.final_objective_change = spread;
      const bool final_stage = stage + 1 >= stages;
      if (at_final && final_stage && (change_ok || objective_ok)) {
        res.converged = 1;
        res.stop_reason = change_ok ? SLX_STOP_DESIGN_CHANGE : SLX_STOP_OBJECTIVE_STALL;
        break;
      }
      if (!final_stage && (iteration - stage_start + 1 >= pr.beta_interval ||
                           (pr.advance_on_convergence && st.max_change <= o.change_tolerance))) {
        ++stage;  // the objective changes with beta: its window restarts
        stage_start = iteration + 1;
        beta = beta_of_stage(stage);
        chist.clear();
      }
    }
    res.iterations = std::min(iteration, o.max_iterations);
    // the final evaluation at the returned design (src/topopt/TopologyOptimizer.cpp:263-283)
    slx_eval_options eo = o.eval;
    eo.beta = beta;
    eo.eta = objective_eta;
    slx_evaluation fin = evaluate_device(c, eo, true);
    res.final_beta = beta;
    if (robust) {
      // the blueprint is the result, the eroded and dilated designs are recorded
      // (TopologyOptimizer.cpp:347-390)
      res.compliance_eroded = fin.compliance;
      res.volume_fraction_eroded = fin.volume_fraction;
      res.linear_solves += c.ncases;
      res.linear_iterations += fin.linear_iterations;
      eo.eta = dilated_eta;
      const slx_evaluation dil = evaluate_device(c, eo, false);
      res.compliance_dilated = dil.compliance;
      res.volume_fraction_dilated = dil.volume_fraction;
      res.linear_solves += c.ncases;
      res.linear_iterations += dil.linear_iterations;
      eo.eta = pr.eta;
      fin = evaluate_device(c, eo, false);
      res.dilated_target_fraction = dilated_target / domain;
    }
    res.linear_solves += c.ncases;
    res.linear_iterations += fin.linear_iterations;
    res.compliance = fin.compliance;
    res.volume = fin.volume;
    res.volume_fraction = fin.volume_fraction;
    res.grey_level = grey();
    c.d_x.download(x, (size_t)ne, c.stream);
    sync(c);
    res.total_seconds = std::chrono::duration<double>(Clock::now() - t_total).count();
    if (result) *result = res;
  });
}

slx_status slx_load_case_compliance_get(const slx_context* ctx, double* compliance) {
  return guarded(const_cast<slx_context*>(ctx), [&](Context& c) {
    if (c.case_compliance.empty()) fail(SLX_CONFIG_ERROR, "no evaluation has run yet");
    std::copy(c.case_compliance.begin(), c.case_compliance.end(), compliance);
  });
}

slx_status slx_solve_adjoint(slx_context* ctx, const double* rhs, int slot, double tol, int max_iterations,
                             double* lambda, slx_pcg_result* result) {
  return guarded(ctx, [&](Context& c) {
    // ComplianceObjective::solve_adjoint (src/topopt/Sensitivity.cpp:239-257) through
    // StaticAnalysis::solve_homogeneous: K_ff lambda_f = rhs_f, zero on the fixed DOFs
    require_model(c);
    if (!c.amg_ready || c.K.rows == 0 || c.k_source != Context::KSource::evaluation)
      fail(SLX_MODEL_ERROR,
           "solve_adjoint needs the operator of the last objective evaluation; evaluate the design again (the "
           "context's K_ff was replaced by slx_assemble, a modal analysis or a caller matrix)");
    if (!(tol > 0.0)) fail(SLX_CONFIG_ERROR, "the adjoint tolerance must be positive");
    const size_t nf = c.free_dofs.size();
    std::vector<double> bf(nf);
    for (size_t i = 0; i < nf; ++i) bf[i] = rhs[c.free_dofs[i]];
    DBuf<double> b, x;
    b.upload(bf, c.stream);
    x.alloc(nf);
    if (slot >= 0 && (size_t)slot >= c.adjoint_slots.size()) c.adjoint_slots.resize((size_t)slot + 1);
    if (slot >= 0 && c.adjoint_slots[(size_t)slot].n == nf)
      k::copy((long long)nf, c.adjoint_slots[(size_t)slot].p, x.p, c.stream);  // warm start from the slot
    else
      k::fill((long long)nf, 0.0, x.p, c.stream);
    const slx_pcg_result r = pcg(c, b.p, x.p, tol, max_iterations > 0 ? max_iterations : 1000);
    ++c.adjoint_solves;
    c.adjoint_iterations += r.iterations;
    require_converged(c, r, tol);
    if (slot >= 0) {
      DBuf<double>& keep = c.adjoint_slots[(size_t)slot];
      keep.alloc(nf);
      k::copy((long long)nf, x.p, keep.p, c.stream);
    }
    std::vector<double> xf(nf);
    x.download(xf.data(), nf, c.stream);
    sync(c);
    std::fill(lambda, lambda + c.ndofs, 0.0);
    for (size_t i = 0; i < nf; ++i) lambda[c.free_dofs[i]] = xf[i];
    if (result) *result = r;
  });
}

void slx_modal_default_options(slx_modal_options* o) {
  // ModalAnalysisOptions (include/sparlab/fem/ModalAnalysis.hpp:44-68), Config.cpp:1120-1136
  o->num_modes = 6;
  o->lumped = 0;
  o->max_iterations = 200;
  o->tolerance = 1.0e-10;
  o->residual_tolerance = 1.0e-6;
  o->rigid_body_ratio = 1.0e-10;
  o->shift_factor = 0.0;
  o->density = 0.0;
  o->linear_tolerance = 1.0e-8;
  o->linear_max_iterations = 1000;
  o->warm_start = 1;
}

slx_status slx_modal_solve(slx_context* ctx, const slx_modal_options* opt, const double* stiffness_scale,
                           const double* mass_scale, double* eigenvalues, double* residuals, double* shapes,
                           slx_modal_result* result) {
  return guarded(ctx, [&](Context& c) {
    require_model(c);
    slx_modal_options o;
    slx_modal_default_options(&o);
    if (opt) o = *opt;
    if (o.num_modes < 1) fail(SLX_CONFIG_ERROR, "modal analysis requires num_modes >= 1");
    if (!(o.density > 0.0))
      fail(SLX_MODEL_ERROR,
           "mass assembly requires a positive density of every material. Set its 'density' to run a modal or "
           "dynamic analysis");
    if (o.shift_factor != 0.0)
      fail(SLX_CONFIG_ERROR, "modal analysis on the device supports shift_factor = 0 only");
    if (o.max_iterations < 1) fail(SLX_CONFIG_ERROR, "modal.max_iterations must be at least 1");
    const size_t ne = (size_t)c.nelem;
    auto check_scale = [&](const double* h, const char* what) {
      if (!h) return;
      for (size_t e = 0; e < ne; ++e)
        if (!std::isfinite(h[e]) || h[e] < 0.0)
          fail(SLX_MODEL_ERROR, std::string(what) + " scale vector has a negative or non-finite entry");
    };
    check_scale(stiffness_scale, "stiffness");
    check_scale(mass_scale, "mass");
    double total_mass = 0.0;
    for (size_t e = 0; e < ne; ++e) total_mass += (mass_scale ? mass_scale[e] : 1.0) * o.density * c.volumes[e];
    std::vector<double> ones;
    const double* h = stiffness_scale;
    if (!h) {
      ones.assign(ne, 1.0);
      h = ones.data();
    }
    c.d_scale.upload(h, ne, c.stream);
    assemble(c, c.d_scale.p);
    c.k_source = Context::KSource::modal;
    amg_setup(c);
    DBuf<double> dm;
    if (mass_scale) dm.upload(mass_scale, ne, c.stream);
    slx_modal_result r{};
    modal_solve(c, o, mass_scale ? dm.p : nullptr, total_mass, eigenvalues, residuals, shapes, r);
    c.have_u.assign(c.have_u.size(), 0);  // the stored solutions belong to another operator
    if (result) *result = r;
  });
}

slx_status slx_timing_enable(slx_context* ctx, int on) {
  return guarded(ctx, [&](Context& c) {
    c.flush_timers();
    c.timing = on != 0;
    c.timers.clear();
  });
}

slx_status slx_timing_get(slx_context* ctx, const char* kernel, double* total_ms, long long* launches) {
  return guarded(ctx, [&](Context& c) {
    c.flush_timers();
    auto it = c.timers.find(kernel);
    *total_ms = it == c.timers.end() ? 0.0 : it->second.total_ms;
    *launches = it == c.timers.end() ? 0 : it->second.launches;
  });
}

slx_status slx_synchronize(slx_context* ctx) {
  return guarded(ctx, [&](Context& c) { sync(c); });
}

}  // extern "C"
