// C entry points of the refcpu oracle, for the tests (ctypes) and for the CPU
// baseline leg of bench.py. TEST INFRASTRUCTURE ONLY: the product library never
// links this file.
#include <cstdint>
#include <cstring>
#include <string>

#include "refcpu.hpp"

#if defined(_OPENMP)
#include <omp.h>
#endif

using namespace refcpu;

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    g_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 9;
  }
}
Problem* P(void* h) { return static_cast<Problem*>(h); }
}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

void* ref_problem_new(int dim, int nx, int ny, int nz, double lx, double ly, double lz, double E, double nu,
                      double t) {
  Problem* p = nullptr;
  const int rc = guarded([&] {
    p = new Problem();
    p->mesh = dim == 3 ? structured_hex(nx, ny, nz, lx, ly, lz) : structured_quad(nx, ny, lx, ly);
    p->E = E;
    p->nu = nu;
    p->thickness = t;
    p->ndpn = dim;
    p->dofs.fixed.assign((size_t)p->mesh.nn * dim, 0);
  });
  if (rc != 0) {
    delete p;
    return nullptr;
  }
  return p;
}
void ref_problem_free(void* h) { delete P(h); }
int ref_problem_nnodes(void* h) { return P(h)->mesh.nn; }
int ref_problem_nelem(void* h) { return P(h)->mesh.ne; }
int ref_problem_ndofs(void* h) { return P(h)->mesh.nn * P(h)->mesh.dim; }
int ref_problem_nfree(void* h) { return (int)P(h)->dofs.free.size(); }

void ref_problem_set_fixed(void* h, const uint8_t* mask) {
  Problem* p = P(h);
  for (size_t i = 0; i < p->dofs.fixed.size(); ++i) p->dofs.fixed[i] = mask[i] ? 1 : 0;
}
void ref_problem_set_loads(void* h, int ncases, const double* f, const double* weights) {
  Problem* p = P(h);
  const size_t nd = (size_t)p->mesh.nn * p->mesh.dim;
  p->loads.assign((size_t)ncases, std::vector<double>(nd));
  double sum = 0.0;
  for (int l = 0; l < ncases; ++l) sum += weights[l];
  p->weights.resize((size_t)ncases);
  for (int l = 0; l < ncases; ++l) {
    std::memcpy(p->loads[(size_t)l].data(), f + (size_t)l * nd, nd * sizeof(double));
    p->weights[(size_t)l] = weights[l] / sum;
  }
}
void ref_problem_set_simp(void* h, double penalty, double emin) {
  P(h)->simp.penalty = penalty;
  P(h)->simp.emin_ratio = emin;
}
void ref_problem_set_solver(void* h, double tol, int maxit, int warm) {
  P(h)->tolerance = tol;
  P(h)->max_iterations = maxit;
  P(h)->warm_start = warm != 0;
}
void ref_problem_set_amg(void* h, double theta, int max_levels, int coarse_size, int smoother, int degree,
                         double cheb_ratio, double damping, int lanczos, int reuse) {
  AmgOptions& o = P(h)->amg_options;
  o.strength_threshold = theta;
  o.max_levels = max_levels;
  o.coarse_size = coarse_size;
  o.smoother = smoother;
  o.smoother_degree = degree;
  o.chebyshev_ratio = cheb_ratio;
  o.prolongator_damping = damping;
  o.lanczos_steps = lanczos;
  o.reuse_aggregates = reuse != 0;
}
void ref_problem_set_design(void* h, double volume_fraction, const double* lower, const double* upper,
                            const uint8_t* passive) {
  Problem* p = P(h);
  const size_t ne = (size_t)p->mesh.ne;
  p->volume_fraction = volume_fraction;
  p->lower.assign(lower, lower + ne);
  p->upper.assign(upper, upper + ne);
  p->passive.assign(ne, 0);
  for (size_t e = 0; e < ne; ++e) p->passive[e] = passive[e] ? 1 : 0;
}
int ref_problem_finalize(void* h, double filter_radius) {
  return guarded([&] {
    Problem* p = P(h);
    p->finalize();
    p->filter_on = filter_radius > 0.0;
    if (p->filter_on) p->filter.build(p->mesh, filter_radius);
  });
}
double ref_mean_element_size(void* h) { return mean_element_size(P(h)->mesh); }
void ref_get_ke0(void* h, double* out) { std::memcpy(out, P(h)->ke0.data(), P(h)->ke0.size() * sizeof(double)); }
void ref_get_free(void* h, int* out) { std::memcpy(out, P(h)->dofs.free.data(), P(h)->dofs.free.size() * sizeof(int)); }
void ref_get_volumes(void* h, double* out) { std::memcpy(out, P(h)->volumes.data(), P(h)->volumes.size() * sizeof(double)); }
long long ref_pattern_nnz(void* h) { return (long long)P(h)->pattern.inner.size(); }
void ref_get_pattern(void* h, int* outer, int* inner, int* block_offset) {
  const Pattern& pt = P(h)->pattern;
  std::memcpy(outer, pt.outer.data(), pt.outer.size() * sizeof(int));
  std::memcpy(inner, pt.inner.data(), pt.inner.size() * sizeof(int));
  if (block_offset) std::memcpy(block_offset, pt.block_offset.data(), pt.block_offset.size() * sizeof(int));
}
void ref_assemble_full(void* h, const double* scale, double* values) {
  Problem* p = P(h);
  std::vector<double> v;
  scatter(p->pattern, p->mesh, p->ndpn, p->ke0.data(), scale, v);
  std::memcpy(values, v.data(), v.size() * sizeof(double));
}

// Assemble K_ff for `scale` (NULL = ones) and keep it as the current matrix.
long long ref_assemble_kff(void* h, const double* scale) {
  long long nnz = -1;
  guarded([&] {
    Problem* p = P(h);
    p->kff = p->assemble_kff(scale);
    nnz = p->kff.nnz();
  });
  return nnz;
}
void ref_get_kff(void* h, int* ptr, int* col, double* val) {
  const Csr& k = P(h)->kff;
  std::memcpy(ptr, k.ptr.data(), k.ptr.size() * sizeof(int));
  std::memcpy(col, k.col.data(), k.col.size() * sizeof(int));
  std::memcpy(val, k.val.data(), k.val.size() * sizeof(double));
}
// Replace the current matrix by caller-provided CSR arrays (same pattern rules).
void ref_set_kff(void* h, int n, const int* ptr, const int* col, const double* val) {
  Csr& k = P(h)->kff;
  k.rows = k.cols = n;
  k.ptr.assign(ptr, ptr + n + 1);
  k.col.assign(col, col + ptr[n]);
  k.val.assign(val, val + ptr[n]);
}

int ref_amg_setup(void* h) {
  return guarded([&] {
    Problem* p = P(h);
    Layout layout;
    layout.dim = p->mesh.dim;
    layout.ndpn = p->ndpn;
    layout.coords = p->mesh.coords.data();
    layout.nn = p->mesh.nn;
    layout.unknowns = p->dofs.free.data();
    layout.n = (Index)p->dofs.free.size();
    p->amg.setup(p->kff, layout);
  });
}
int ref_amg_nlevels(void* h) { return (int)P(h)->amg.levels().size(); }
int ref_amg_reused(void* h) { return P(h)->amg.stats().reused_aggregates ? 1 : 0; }
double ref_amg_operator_complexity(void* h) { return P(h)->amg.stats().operator_complexity; }
double ref_amg_coarse_pivot_ratio(void* h) { return P(h)->amg.stats().coarse_pivot_ratio; }
long long ref_amg_storage_nonzeros(void* h) { return P(h)->amg.stats().storage_nonzeros; }
// which: 0 = A_l, 1 = P, 2 = P-hat, 3 = R. Returns rows, cols, nnz through info.
void ref_amg_level_info(void* h, int l, int which, int* info) {
  const Level& lv = P(h)->amg.levels()[(size_t)l];
  const Csr* m = which == 0 ? &lv.a : which == 1 ? &lv.p : which == 2 ? &lv.p_tent : &lv.r;
  info[0] = m->rows;
  info[1] = m->cols;
  info[2] = m->nnz();
}
void ref_amg_level_get(void* h, int l, int which, int* ptr, int* col, double* val) {
  const Level& lv = P(h)->amg.levels()[(size_t)l];
  const Csr* m = which == 0 ? &lv.a : which == 1 ? &lv.p : which == 2 ? &lv.p_tent : &lv.r;
  std::memcpy(ptr, m->ptr.data(), m->ptr.size() * sizeof(int));
  std::memcpy(col, m->col.data(), m->col.size() * sizeof(int));
  std::memcpy(val, m->val.data(), m->val.size() * sizeof(double));
}
double ref_amg_lambda(void* h, int l) { return P(h)->amg.levels()[(size_t)l].lambda_max; }
int ref_amg_nullspace_dim(void* h) { return P(h)->amg.nullspace_dim(); }
void ref_amg_nullspace(void* h, int l, double* out) {
  const Level& lv = P(h)->amg.levels()[(size_t)l];
  std::memcpy(out, lv.nullspace.data(), lv.nullspace.size() * sizeof(double));
}
int ref_amg_num_nodes(void* h, int l) { return P(h)->amg.levels()[(size_t)l].num_nodes; }
void ref_amg_aggregates(void* h, int l, int* out) {
  const Level& lv = P(h)->amg.levels()[(size_t)l];
  std::memcpy(out, lv.aggregate.data(), lv.aggregate.size() * sizeof(int));
}
void ref_amg_node_ptr(void* h, int l, int* out) {
  const Level& lv = P(h)->amg.levels()[(size_t)l];
  std::memcpy(out, lv.node_ptr.data(), lv.node_ptr.size() * sizeof(int));
}
int ref_amg_apply(void* h, const double* r, double* z) {
  return guarded([&] { P(h)->amg.apply(r, z); });
}
int ref_pcg(void* h, const double* b, double* x, double tol, int maxit, int* iterations, double* relres,
            int* converged) {
  return guarded([&] {
    Problem* p = P(h);
    const PcgResult r = pcg_solve(p->kff, b, x, p->amg, tol, maxit);
    *iterations = r.iterations;
    *relres = r.relative_residual;
    *converged = r.converged ? 1 : 0;
  });
}

long long ref_filter_nnz(void* h) { return P(h)->filter.h.nnz(); }
void ref_filter_get(void* h, int* ptr, int* col, double* val, double* row_sums) {
  const Filter& f = P(h)->filter;
  std::memcpy(ptr, f.h.ptr.data(), f.h.ptr.size() * sizeof(int));
  std::memcpy(col, f.h.col.data(), f.h.col.size() * sizeof(int));
  std::memcpy(val, f.h.val.data(), f.h.val.size() * sizeof(double));
  if (row_sums) std::memcpy(row_sums, f.row_sums.data(), f.row_sums.size() * sizeof(double));
}
void ref_filter_apply(void* h, int adjoint, const double* in, double* out) {
  if (adjoint) P(h)->filter.pull_back(in, out); else P(h)->filter.to_physical(in, out);
}

int ref_evaluate(void* h, const double* x, int need_gradients, double* out_scalars, int* iterations, double* dc_dx,
                 double* dv_dx, double* u_out, double* energy_out, double* physical_out) {
  return guarded([&] {
    Problem* p = P(h);
    std::vector<double> xv(x, x + p->mesh.ne);
    Evaluation ev = evaluate(*p, xv, need_gradients != 0);
    out_scalars[0] = ev.compliance;
    out_scalars[1] = ev.volume;
    out_scalars[2] = ev.volume_fraction;
    *iterations = ev.linear_iterations;
    if (need_gradients && dc_dx) std::memcpy(dc_dx, ev.dc_dx.data(), ev.dc_dx.size() * sizeof(double));
    if (need_gradients && dv_dx) std::memcpy(dv_dx, ev.dv_dx.data(), ev.dv_dx.size() * sizeof(double));
    if (u_out)
      for (size_t l = 0; l < ev.u.size(); ++l)
        std::memcpy(u_out + l * ev.u[l].size(), ev.u[l].data(), ev.u[l].size() * sizeof(double));
    if (energy_out) std::memcpy(energy_out, ev.energy.data(), ev.energy.size() * sizeof(double));
    if (physical_out) std::memcpy(physical_out, ev.physical.data(), ev.physical.size() * sizeof(double));
  });
}

int ref_oc(void* h, const double* x, const double* dc, const double* dv, double move, double damping, double* xnew,
           double* out_scalars, int* bisections) {
  return guarded([&] {
    Problem* p = P(h);
    const size_t ne = (size_t)p->mesh.ne;
    OcOptions o;
    o.move_limit = move;
    o.damping = damping;
    OcStep s = oc_update(*p, std::vector<double>(x, x + ne), std::vector<double>(dc, dc + ne),
                         std::vector<double>(dv, dv + ne), o);
    std::memcpy(xnew, s.x.data(), ne * sizeof(double));
    out_scalars[0] = s.lambda;
    out_scalars[1] = s.achieved_volume;
    out_scalars[2] = s.max_change;
    out_scalars[3] = s.volume_converged ? 1.0 : 0.0;
    *bisections = s.bisections;
  });
}

// projection of the physical density for ref_evaluate (beta 0: off)
void ref_problem_set_projection(void* h, double beta, double eta) {
  P(h)->beta = beta;
  P(h)->eta = eta;
}

// run_oc's beta continuation (ProjectionOptions, non-robust)
void ref_problem_set_projection_schedule(void* h, int enabled, double eta, double beta_start, double beta_max,
                                         double beta_factor, int beta_interval, int advance_on_convergence) {
  Projection& pr = P(h)->projection;
  pr.enabled = enabled != 0;
  pr.eta = eta;
  pr.beta_start = beta_start;
  pr.beta_max = beta_max;
  pr.beta_factor = beta_factor;
  pr.beta_interval = beta_interval;
  pr.advance_on_convergence = advance_on_convergence != 0;
}

// the robust formulation of the projection schedule
void ref_problem_set_robust(void* h, int robust, double delta, int volume_interval) {
  Projection& pr = P(h)->projection;
  pr.robust = robust != 0;
  pr.robust_delta = delta;
  pr.robust_volume_interval = volume_interval;
}

double ref_heaviside_project(double r, double beta, double eta) { return heaviside_project(r, beta, eta); }
double ref_heaviside_derivative(double r, double beta, double eta) { return heaviside_derivative(r, beta, eta); }
int ref_projection_stages(double beta_start, double beta_max, double beta_factor, int k, double* beta) {
  Projection pr;
  pr.beta_start = beta_start;
  pr.beta_max = beta_max;
  pr.beta_factor = beta_factor;
  *beta = pr.beta_of_stage(k);
  return pr.num_stages();
}

// history: max_iterations x 16, the columns of history.csv (src/io/ResultWriter.cpp:1428-1470), then beta,
// then the robust columns (eroded and dilated volume fractions, dilated target fraction);
// robust[5]: eroded and dilated compliance and volume fraction, the last dilated target fraction;
// seconds are not timed here
int ref_run_oc(void* h, int max_iterations, double change_tol, double objective_tol, int window, int cont_steps,
               double penalty_start, int cont_iterations, double move, double damping, double* x_inout,
               double* history, int* performed, double* final_compliance, int* stop_reason, double* final_beta,
               double* robust) {
  return guarded([&] {
    Problem* p = P(h);
    RunOptions o;
    o.max_iterations = max_iterations;
    o.change_tolerance = change_tol;
    o.objective_tolerance = objective_tol;
    o.objective_window = window;
    o.continuation_steps = cont_steps;
    o.penalty_start = penalty_start;
    o.continuation_iterations = cont_iterations;
    o.oc.move_limit = move;
    o.oc.damping = damping;
    RunResult r = run_oc(*p, o, std::vector<double>(x_inout, x_inout + p->mesh.ne));
    std::memcpy(x_inout, r.x.data(), r.x.size() * sizeof(double));
    for (size_t i = 0; i < r.history.size(); ++i) {
      const IterationRecord& it = r.history[i];
      double* row = history + i * 16;  // the columns of history.csv
      row[0] = it.iteration;
      row[1] = it.penalty;
      row[2] = it.compliance;
      row[3] = it.volume;
      row[4] = it.volume_fraction;
      row[5] = it.max_change;
      row[6] = it.lambda;
      row[7] = it.gray;
      row[8] = it.bisections;
      row[9] = it.volume_converged ? 1.0 : 0.0;
      row[10] = 0.0;
      row[11] = it.linear_iterations;
      row[12] = it.beta;
      row[13] = it.eroded_vf;
      row[14] = it.dilated_vf;
      row[15] = it.dilated_target_fraction;
    }
    robust[0] = r.compliance_eroded;
    robust[1] = r.compliance_dilated;
    robust[2] = r.vf_eroded;
    robust[3] = r.vf_dilated;
    robust[4] = r.dilated_target_fraction;
    *performed = r.iterations;
    *final_beta = r.final_beta;
    *final_compliance = r.compliance;
    *stop_reason = r.stop_reason;
  });
}

// Stand-alone kernels for the tests.
double ref_chunked_dot(const double* a, const double* b, int n) { return chunked_dot(a, b, n); }
int ref_cpqr(const double* block, int rows, int m, double* q, double* r) {
  std::vector<double> qv, rv;
  const int rank = cpqr_block(std::vector<double>(block, block + (size_t)rows * m), rows, m, qv, rv);
  std::memcpy(q, qv.data(), qv.size() * sizeof(double));
  std::memcpy(r, rv.data(), rv.size() * sizeof(double));
  return rank;
}
double ref_tridiag_max(const double* diag, const double* sub, int k) {
  return tridiagonal_max_eigenvalue(std::vector<double>(diag, diag + k), std::vector<double>(sub, sub + (k > 0 ? k - 1 : 0)));
}
// wall seconds of the last ref_evaluate: assembly, AMG setup, PCG, whole call
void ref_last_evaluate_seconds(void* h, double* out) {
  out[0] = P(h)->seconds_assemble;
  out[1] = P(h)->seconds_setup;
  out[2] = P(h)->seconds_pcg;
  out[3] = P(h)->seconds_total;
}
void ref_set_num_threads(int n) {
#if defined(_OPENMP)
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int ref_num_threads() {
#if defined(_OPENMP)
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
