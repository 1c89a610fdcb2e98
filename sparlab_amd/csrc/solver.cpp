// V-cycle and preconditioned CG on the device (src/fem/Multigrid.cpp:783-885,
// 1113-1121, 1152-1237).
//
// The CG loop runs on the host and reads back one pair of scalars per
// iteration (p^T A p for the breakdown test and r^T r for the stopping test);
// every vector operation, every reduction and the whole V-cycle stay on the
// device stream. Step lengths alpha and beta are formed on the device from
// reduced scalars, so the host never waits inside an iteration except for
// that one read.
//
// precision = 1 runs the V-cycle in FP32 (operators, smoother, transfers and
// the coarse inverse stored and applied in FP32, accumulated in FP32 lanes
// with FP64 row sums) under the FP64 outer CG; the FP32 cycle is not exactly
// symmetric, so the outer method becomes flexible CG (Polak-Ribiere beta).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <cmath>
#include <limits>
#include <mutex>
#include <sstream>

#include "slx_kernels.hpp"
#include "slx_mf.hpp"
#include "slx_solver.hpp"

namespace slx {

namespace {
// device scalar slots
constexpr int S_PQ = 0, S_RR = 1, S_RZ = 2, S_RZN = 3, S_RZX = 4, S_TMP = 5;

DCsr& op(Context& c, size_t l) { return l == 0 ? c.K : c.levels[l].A; }

bool use_mf(const Context& c, size_t l) { return l == 0 && c.mf; }

template <typename T>
kmf::MfArgs<T> mf_base(const Context& c);
template <>
kmf::MfArgs<double> mf_base<double>(const Context& c) {
  kmf::MfArgs<double> a{};
  a.X = c.nx + 1, a.Y = c.ny + 1, a.Z = c.nz + 1;
  a.tile = c.mf_tile[0];
  a.nodemap = c.d_nodemap.p;
  a.scale = c.mf_scale.p;
  a.ke = c.d_ke0.p;
  return a;
}
template <>
kmf::MfArgs<float> mf_base<float>(const Context& c) {
  kmf::MfArgs<float> a{};
  a.X = c.nx + 1, a.Y = c.ny + 1, a.Z = c.nz + 1;
  a.tile = c.mf_tile[1];
  a.nodemap = c.d_nodemap.p;
  a.scale = c.mf_scale32.p;
  a.ke = c.mf_ke32.p;
  return a;
}

// r = b - A_l x at level l
template <typename T>
void level_residual(Context& c, size_t l, const T* x, const T* b, T* r) {
  if (use_mf(c, l)) {
    kmf::MfArgs<T> a = mf_base<T>(c);
    a.x = x;
    a.b = b;
    a.y = r;
    kmf::apply(a, 2, c.stream);
  } else if constexpr (sizeof(T) == 8) {
    k::residual(op(c, l), x, b, r, c.stream);
  } else {
    k::residual32(op(c, l), x, b, r, c.stream);
  }
}

// Level-0 transfers with the matrix-free operator: the smoothed prolongator is
// P = (I - omega D^-1 K) P-hat (Multigrid.cpp:686-701) and R = P^T =
// P-hat^T (I - omega K D^-1), so both apply as one operator pass and one
// product with the tentative prolongator (through its aggregate blocks when
// build_tent_blocks found them) instead of reading the assembled P and R.
// scratch is a free level-0 vector.
k::TentBlocks tent_blocks(const Level& lv) {
  k::TentBlocks t;
  t.n = lv.n;
  t.naggs = lv.tent_naggs;
  t.c0len = lv.tent_c0len.p;
  t.agg_ptr = lv.tent_agg_ptr.p;
  t.agg_rows = lv.tent_agg_rows.p;
  t.agg_c0len = lv.tent_agg_c0len.p;
  return t;
}

// Restriction into the next level's right-hand side bc; with cdinv set, the
// next level's zero-guess smoother start follows (cd = cx = cdinv bc / theta).
template <typename T>
void restrict_mf(Context& c, const T* r, T* scratch, T* bc, const T* cdinv = nullptr, double theta = 1.0,
                 T* cd = nullptr, T* cx = nullptr);
template <typename T>
void prolong_mf(Context& c, const T* xc, T* scratch, T* x);
template <>
void restrict_mf<double>(Context& c, const double* r, double* scratch, double* bc, const double* cdinv, double theta,
                         double* cd, double* cx) {
  Level& lv = c.levels[0];
  kmf::MfArgs<double> a = mf_base<double>(c);
  a.x = r;
  a.in_scale = lv.dinv.p;
  a.c1 = lv.omega;
  a.y = scratch;
  kmf::apply(a, 5, c.stream);
  const double* sc = scratch;
  if (lv.tent_ok) k::tent_restrict(tent_blocks(lv), lv.tent_ella.p, sc, bc, c.stream, cdinv, theta, cd, cx);
  else if (cdinv) k::spmv_start(lv.Rt, sc, bc, cdinv, theta, cd, cx, c.stream);
  else k::spmv(lv.Rt, sc, bc, c.stream);
}
template <>
void restrict_mf<float>(Context& c, const float* r, float* scratch, float* bc, const float* cdinv, double theta,
                        float* cd, float* cx) {
  Level& lv = c.levels[0];
  kmf::MfArgs<float> a = mf_base<float>(c);
  a.x = r;
  a.in_scale = lv.dinv32.p;
  a.c1 = lv.omega;
  a.y = scratch;
  kmf::apply(a, 5, c.stream);
  const float* sc = scratch;
  if (lv.tent_ok) k::tent_restrict(tent_blocks(lv), lv.tent_ella32.p, sc, bc, c.stream, cdinv, theta, cd, cx);
  else if (cdinv) k::spmv_start32(lv.Rt, sc, bc, cdinv, theta, cd, cx, c.stream);
  else k::spmv32(lv.Rt, sc, bc, c.stream);
}
template <>
void prolong_mf<double>(Context& c, const double* xc, double* scratch, double* x) {
  Level& lv = c.levels[0];
  if (lv.tent_ok) k::tent_prolong(tent_blocks(lv), lv.tent_ell.p, xc, scratch, c.stream);
  else k::spmv(lv.Ptent, xc, scratch, c.stream);
  kmf::MfArgs<double> a = mf_base<double>(c);
  a.x = scratch;
  a.dinv = lv.dinv.p;
  a.c1 = lv.omega;
  a.xs = x;
  kmf::apply(a, 6, c.stream);
}
template <>
void prolong_mf<float>(Context& c, const float* xc, float* scratch, float* x) {
  Level& lv = c.levels[0];
  if (lv.tent_ok) k::tent_prolong(tent_blocks(lv), lv.tent_ell32.p, xc, scratch, c.stream);
  else k::spmv32(lv.Ptent, xc, scratch, c.stream);
  kmf::MfArgs<float> a = mf_base<float>(c);
  a.x = scratch;
  a.dinv = lv.dinv32.p;
  a.c1 = lv.omega;
  a.xs = x;
  kmf::apply(a, 6, c.stream);
}

struct Cheb {
  double theta, delta, sigma;
};
Cheb cheb_params(const Context& c, const Level& lv) {
  // [lambda_max / ratio, 1.1 lambda_max] of D^-1 A (Multigrid.cpp:809-813)
  const double upper = 1.1 * lv.lambda_max;
  const double lower = upper / c.amg_opt.chebyshev_ratio;
  Cheb p;
  p.theta = 0.5 * (upper + lower);
  p.delta = 0.5 * (upper - lower);
  p.sigma = p.theta / p.delta;
  return p;
}

// ---- the V-cycle ------------------------------------------------------------
// One implementation serves one right-hand side and a block of nv load cases
// (case-major level vectors, nv x n_l): every assembled-operator pass
// (smoother, residual, transfers, coarse solve) reads its matrix once for all
// cases, each case summed in the single-case order, so each case's result
// equals its own cycle bit for bit. The matrix-free level-0 operator, which
// reads no matrix, runs once per case.
//
// The smoother's first step is fused into the pass before it: the restriction
// that forms a level's right-hand side also forms its zero-guess start
// (d = x = D^-1 b / theta), and the post-smoothing residual also forms the
// start from the prolongated iterate (d = D^-1 r / theta into a second
// iterate x2 = x + d, since other rows still read x). The post-smoothed result
// of a level therefore lands in its second iterate.
template <typename T>
struct CycleVecs {
  T *x, *x2, *b, *res, *d, *d2;
  const T* dinv;
};
template <typename T>
CycleVecs<T> level_vecs(Level& lv, bool block);
template <>
CycleVecs<double> level_vecs<double>(Level& lv, bool block) {
  if (block) return {lv.vx.p, lv.vx2.p, lv.vb.p, lv.vres.p, lv.vd.p, lv.vd2.p, lv.dinv.p};
  return {lv.x.p, lv.x2.p, lv.b.p, lv.res.p, lv.d.p, lv.d2.p, lv.dinv.p};
}
template <>
CycleVecs<float> level_vecs<float>(Level& lv, bool block) {
  if (block) return {lv.vx32.p, lv.vx232.p, lv.vb32.p, lv.vres32.p, lv.vd32.p, lv.vd232.p, lv.dinv32.p};
  return {lv.x32.p, lv.x232.p, lv.b32.p, lv.res32.p, lv.d32.p, lv.d232.p, lv.dinv32.p};
}

// the assembled-operator kernels of either storage, for nv cases
void csr_residual(const DCsr& A, const double* x, const double* b, double* r, hipStream_t s, int nv) {
  k::residual(A, x, b, r, s, nv);
}
void csr_residual(const DCsr& A, const float* x, const float* b, float* r, hipStream_t s, int nv) {
  k::residual32(A, x, b, r, s, nv);
}
void csr_residual_first(const DCsr& A, const double* x, const double* b, double* r, const double* dinv, double th,
                        double* d, double* x2, hipStream_t s, int nv) {
  k::residual_first(A, x, b, r, dinv, th, d, x2, s, nv);
}
void csr_residual_first(const DCsr& A, const float* x, const float* b, float* r, const float* dinv, double th,
                        float* d, float* x2, hipStream_t s, int nv) {
  k::residual_first32(A, x, b, r, dinv, th, d, x2, s, nv);
}
void csr_apply(const DCsr& A, const double* x, double* y, hipStream_t s, int nv) { k::spmv(A, x, y, s, nv); }
void csr_apply(const DCsr& A, const float* x, float* y, hipStream_t s, int nv) { k::spmv32(A, x, y, s, nv); }
void csr_apply_start(const DCsr& A, const double* x, double* y, const double* dinv, double th, double* d, double* x2,
                     hipStream_t s, int nv) {
  k::spmv_start(A, x, y, dinv, th, d, x2, s, nv);
}
void csr_apply_start(const DCsr& A, const float* x, float* y, const float* dinv, double th, float* d, float* x2,
                     hipStream_t s, int nv) {
  k::spmv_start32(A, x, y, dinv, th, d, x2, s, nv);
}
void csr_apply_add(const DCsr& A, const double* x, double* y, hipStream_t s, int nv) { k::spmv_add(A, x, y, s, nv); }
void csr_apply_add(const DCsr& A, const float* x, float* y, hipStream_t s, int nv) { k::spmv_add32(A, x, y, s, nv); }
void csr_cheb(const DCsr& A, const double* dinv, const double* ri, double* ro, const double* di, double* dout,
              double* x, double c1, double c2, hipStream_t s, int nv) {
  k::cheb_step(A, dinv, ri, ro, di, dout, x, c1, c2, s, nv);
}
void csr_cheb(const DCsr& A, const float* dinv, const float* ri, float* ro, const float* di, float* dout, float* x,
              double c1, double c2, hipStream_t s, int nv) {
  k::cheb_step32(A, dinv, ri, ro, di, dout, x, c1, c2, s, nv);
}
void first_zero(int n, const double* dinv, const double* b, double th, double* x, double* d, hipStream_t s, int nv) {
  k::cheb_first_zero(n, dinv, b, th, x, d, s, nv);
}
void first_zero(int n, const float* dinv, const float* b, double th, float* x, float* d, hipStream_t s, int nv) {
  k::cheb_first_zero32(n, dinv, b, th, x, d, s, nv);
}
void coarse_solve(Context& c, int n, const double* b, double* x, int nv) {
  k::dense_gemv(n, c.coarse_inv.p, b, x, c.stream, nv);
}
void coarse_solve(Context& c, int n, const float* b, float* x, int nv) {
  k::dense_gemv32(n, c.coarse_inv32.p, b, x, c.stream, nv);
}

// r = b - A_l x for nv cases
template <typename T>
void block_residual(Context& c, size_t l, const T* x, const T* b, T* r, int nv) {
  if (use_mf(c, l)) {
    const long long n = c.levels[l].n;
    for (int v = 0; v < nv; ++v) level_residual<T>(c, l, x + v * n, b + v * n, r + v * n);
  } else {
    csr_residual(op(c, l), x, b, r, c.stream, nv);
  }
}

// Chebyshev steps 1..degree-1 from x and d (Multigrid.cpp:833-853); r_first is
// the residual the first step starts from. d and d2 swap roles per step.
template <typename T>
void cheb_steps(Context& c, size_t l, const T* r_first, T* x, T* d, T* d2, T* res, const T* dinv, int nv) {
  Level& lv = c.levels[l];
  DCsr& A = op(c, l);
  const Cheb cp = cheb_params(c, lv);
  const long long n = lv.n;
  double rho = 1.0 / cp.sigma;
  hipStream_t s = c.stream;
  for (int k = 1; k < c.amg_opt.smoother_degree; ++k) {
    const double rho_new = 1.0 / (2.0 * cp.sigma - rho);
    const double c1 = rho_new * rho;
    const double c2 = 2.0 * rho_new / cp.delta;
    const T* rin = k == 1 ? r_first : res;
    if (use_mf(c, l)) {
      for (int v = 0; v < nv; ++v) {
        kmf::MfArgs<T> a = mf_base<T>(c);
        a.x = d + v * n, a.b = rin + v * n, a.y = res + v * n, a.dinv = dinv;
        a.d_out = d2 + v * n, a.xs = x + v * n, a.c1 = c1, a.c2 = c2;
        kmf::apply(a, 3, s);
      }
    } else {
      csr_cheb(A, dinv, rin, res, d, d2, x, c1, c2, s, nv);
    }
    std::swap(d, d2);
    rho = rho_new;
  }
}

// Symmetric Gauss-Seidel V-cycle from level l (Multigrid.cpp:788-806,857-885):
// x = 0, forward sweeps, residual, restriction, the coarse correction added to
// x, backward sweeps. The result stays in x, which is returned.
void gs_sweeps(Context& c, size_t l, bool forward, const double* b, double* x, int nv) {
  Level& lv = c.levels[l];
  const DBuf<int>& rows = forward ? lv.gs_frows : lv.gs_brows;
  const std::vector<int>& lp = forward ? lv.gs_fptr : lv.gs_bptr;
  for (int k = 0; k < c.amg_opt.smoother_degree; ++k) k::gs_sweep(op(c, l), rows.p, lp, lv.dinv.p, b, x, c.stream, nv);
}
void gs_sweeps(Context& c, size_t l, bool forward, const float* b, float* x, int nv) {
  Level& lv = c.levels[l];
  const DBuf<int>& rows = forward ? lv.gs_frows : lv.gs_brows;
  const std::vector<int>& lp = forward ? lv.gs_fptr : lv.gs_bptr;
  for (int k = 0; k < c.amg_opt.smoother_degree; ++k)
    k::gs_sweep32(op(c, l), rows.p, lp, lv.dinv32.p, b, x, c.stream, nv);
}
void zero(long long n, double* x, hipStream_t s) { k::fill(n, 0.0, x, s); }
void zero(long long n, float* x, hipStream_t s) { k::fill32(n, 0.0f, x, s); }

template <typename T>
T* cycle_gs(Context& c, size_t l, const T* b, T* x, T* x_out, int nv, bool block) {
  Level& lv = c.levels[l];
  if (l + 1 == c.levels.size()) {
    coarse_solve(c, lv.n, b, x_out, nv);
    return x_out;
  }
  if (lv.gs_fptr.empty()) fail(SLX_SOLVER_ERROR, "multigrid: internal error (Gauss-Seidel schedule missing)");
  const CycleVecs<T> v = level_vecs<T>(lv, block);
  const long long n = lv.n;
  zero(n * nv, x, c.stream);
  gs_sweeps(c, l, true, b, x, nv);
  block_residual<T>(c, l, x, b, v.res, nv);
  Level& nl = c.levels[l + 1];
  const CycleVecs<T> w = level_vecs<T>(nl, block);
  const long long nc = nl.n;
  if (use_mf(c, l)) {
    for (int q = 0; q < nv; ++q) restrict_mf<T>(c, v.res + q * n, v.d2 + q * n, w.b + q * nc);
  } else {
    csr_apply(lv.R, v.res, w.b, c.stream, nv);
  }
  const T* xc = cycle_gs<T>(c, l + 1, w.b, w.x, w.x2, nv, block);
  if (use_mf(c, l)) {
    for (int q = 0; q < nv; ++q) prolong_mf<T>(c, xc + q * nc, v.d2 + q * n, x + q * n);
  } else {
    csr_apply_add(lv.P, xc, x, c.stream, nv);
  }
  gs_sweeps(c, l, false, b, x, nv);
  return x;
}

// V-cycle from level l for right-hand side b (Multigrid.cpp:857-885). x is the
// pre-smoothing iterate and x_out receives the result (the second iterate of
// the fused post-smoothing start); started: x and the level's d already hold
// the zero-guess smoother start. Returns x_out.
template <typename T>
T* cycle(Context& c, size_t l, const T* b, T* x, T* x_out, int nv, bool block, bool started) {
  Level& lv = c.levels[l];
  hipStream_t s = c.stream;
  if (l + 1 == c.levels.size()) {
    coarse_solve(c, lv.n, b, x_out, nv);
    return x_out;
  }
  const CycleVecs<T> v = level_vecs<T>(lv, block);
  const long long n = lv.n;
  const bool mf = use_mf(c, l);
  const Cheb cp = cheb_params(c, lv);
  // pre-smoothing from a zero guess: x = d = D^-1 b / theta, r starts as b
  if (!started) first_zero(lv.n, v.dinv, b, cp.theta, x, v.d, s, nv);
  cheb_steps<T>(c, l, b, x, v.d, v.d2, v.res, v.dinv, nv);
  block_residual<T>(c, l, x, b, v.res, nv);
  // the smoother's directions are free between pre- and post-smoothing: d2
  // serves as the transfers' scratch
  T* const scratch = v.d2;
  Level& nl = c.levels[l + 1];
  const CycleVecs<T> w = level_vecs<T>(nl, block);
  const long long nc = nl.n;
  // a level above the coarsest starts its smoother in the restriction's pass
  const bool start_next = l + 2 < c.levels.size();
  const double th_next = start_next ? cheb_params(c, nl).theta : 1.0;
  if (mf) {
    for (int q = 0; q < nv; ++q) {
      if (start_next) restrict_mf<T>(c, v.res + q * n, scratch + q * n, w.b + q * nc, w.dinv, th_next, w.d + q * nc,
                                     w.x + q * nc);
      else restrict_mf<T>(c, v.res + q * n, scratch + q * n, w.b + q * nc);
    }
  } else if (start_next) {
    csr_apply_start(lv.R, v.res, w.b, w.dinv, th_next, w.d, w.x, s, nv);
  } else {
    csr_apply(lv.R, v.res, w.b, s, nv);
  }
  const T* xc = cycle<T>(c, l + 1, w.b, w.x, w.x2, nv, block, start_next);
  if (mf) {
    for (int q = 0; q < nv; ++q) prolong_mf<T>(c, xc + q * nc, scratch + q * n, x + q * n);
  } else {
    csr_apply_add(lv.P, xc, x, s, nv);
  }
  // post-smoothing from the current x: its residual and the smoother's start
  // (d = D^-1 r / theta, x_out = x + d) in one pass
  if (mf) {
    for (int q = 0; q < nv; ++q) {
      kmf::MfArgs<T> a = mf_base<T>(c);
      a.x = x + q * n, a.b = b + q * n, a.y = v.res + q * n, a.dinv = v.dinv;
      a.d_out = v.d + q * n, a.xs = x_out + q * n, a.c1 = cp.theta;
      kmf::apply(a, 7, s);
    }
  } else {
    csr_residual_first(op(c, l), x, b, v.res, v.dinv, cp.theta, v.d, x_out, s, nv);
  }
  cheb_steps<T>(c, l, v.res, x_out, v.d, v.d2, v.res, v.dinv, nv);
  return x_out;
}

// z = M^-1 r for nv cases (block: the level block vectors). The FP64 cycle
// writes z (returns null); the FP32 cycle returns level 0's FP32 result.
const float* vcycle_cases(Context& c, const double* r, double* z, int nv, bool block) {
  Level& top = c.levels[0];
  const int n = top.n;
  if (c.amg_opt.smoother == 1) {
    if (c.amg_opt.precision >= 1) {
      const CycleVecs<float> v = level_vecs<float>(top, block);
      k::to_float((long long)n * nv, r, v.b, c.stream);
      return cycle_gs<float>(c, 0, v.b, v.x, v.x2, nv, block);
    }
    const CycleVecs<double> v = level_vecs<double>(top, block);
    const double* zr = cycle_gs<double>(c, 0, r, v.x, z, nv, block);
    if (zr != z) k::copy((long long)n * nv, zr, z, c.stream);
    return nullptr;
  }
  if (c.amg_opt.precision >= 1) {
    const CycleVecs<float> v = level_vecs<float>(top, block);
    if (c.levels.size() > 1) {
      // b = (float) r and the smoother's zero-guess start in one pass
      k::cheb_first_zero_from64(n, v.dinv, r, cheb_params(c, top).theta, v.b, v.x, v.d, c.stream, nv);
      return cycle<float>(c, 0, v.b, v.x, v.x2, nv, block, true);
    }
    k::to_float((long long)n * nv, r, v.b, c.stream);
    return cycle<float>(c, 0, v.b, v.x, v.x2, nv, block, false);
  }
  const CycleVecs<double> v = level_vecs<double>(top, block);
  cycle<double>(c, 0, r, v.x, z, nv, block, false);
  return nullptr;
}

const float* block_vcycle_raw(Context& c, const double* r, double* z, int nv) {
  return vcycle_cases(c, r, z, nv, true);
}

}  // namespace

// ---------------------------------------------------------------------------
Timed::Timed(Context& ctx, const char* n) : c(ctx), name(n) {
  if (!c.timing || !name || c.capturing) return;
  a = c.take_event();
  b = c.take_event();
  SLX_HIP(hipEventRecord(a, c.stream));
}
Timed::~Timed() {
  if (!a) return;
  (void)hipEventRecord(b, c.stream);
  c.pending.push_back({name, {a, b}});
}

namespace {
constexpr size_t kStageChunk = size_t(32) << 20;
std::mutex g_stage_mutex;
char* g_stage[2] = {nullptr, nullptr};

void stage_init() {
  if (!g_stage[0])
    for (char*& b : g_stage) SLX_HIP(hipHostMalloc((void**)&b, kStageChunk, hipHostMallocDefault));
}

// host memcpy of one chunk, split over the host threads
void host_copy(char* dst, const char* src, size_t bytes) {
  const size_t slice = size_t(1) << 20;
  const long long nslices = (long long)((bytes + slice - 1) / slice);
#pragma omp parallel for schedule(static) if (nslices > 1)
  for (long long i = 0; i < nslices; ++i) {
    const size_t off = (size_t)i * slice;
    std::memcpy(dst + off, src + off, std::min(slice, bytes - off));
  }
}
}  // namespace

void stage_copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_stage_mutex);
  stage_init();
  char* d = static_cast<char*>(dst);
  const char* h = static_cast<const char*>(src);
  size_t i = 0;
  for (size_t off = 0; off < bytes; off += kStageChunk, ++i) {
    const size_t len = std::min(kStageChunk, bytes - off);
    if (i >= 2) SLX_HIP(hipStreamSynchronize(s));  // the copy that last read this half is done
    host_copy(g_stage[i % 2], h + off, len);
    SLX_HIP(hipMemcpyAsync(d + off, g_stage[i % 2], len, hipMemcpyHostToDevice, s));
  }
  SLX_HIP(hipStreamSynchronize(s));
}

void stage_copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_stage_mutex);
  stage_init();
  char* h = static_cast<char*>(dst);
  const char* d = static_cast<const char*>(src);
  const size_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
  if (nchunks == 0) return;
  SLX_HIP(hipMemcpyAsync(g_stage[0], d, std::min(kStageChunk, bytes), hipMemcpyDeviceToHost, s));
  for (size_t i = 0; i < nchunks; ++i) {
    SLX_HIP(hipStreamSynchronize(s));  // chunk i has arrived
    const size_t off = i * kStageChunk, len = std::min(kStageChunk, bytes - off);
    if (i + 1 < nchunks) {
      const size_t noff = off + kStageChunk;
      SLX_HIP(hipMemcpyAsync(g_stage[(i + 1) % 2], d + noff, std::min(kStageChunk, bytes - noff),
                             hipMemcpyDeviceToHost, s));
    }
    host_copy(h + off, g_stage[i % 2], len);
  }
  SLX_HIP(hipStreamSynchronize(s));
}

double* Context::host_buffer(size_t n) {
  if (n > pinned_n) {
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_n = 0;
    SLX_HIP(hipHostMalloc((void**)&pinned, std::max<size_t>(n, 512) * sizeof(double), hipHostMallocDefault));
    pinned_n = std::max<size_t>(n, 512);
  }
  return pinned;
}

hipEvent_t Context::take_event() {
  if (event_pool.empty()) {
    hipEvent_t e;
    SLX_HIP(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = event_pool.back();
  event_pool.pop_back();
  return e;
}

void Context::flush_timers() {
  if (pending.empty()) return;
  SLX_HIP(hipStreamSynchronize(stream));
  for (auto& pe : pending) {
    float ms = 0.0f;
    SLX_HIP(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
    TimerSlot& t = timers[pe.first];
    t.total_ms += ms;
    t.launches += 1;
    event_pool.push_back(pe.second.first);
    event_pool.push_back(pe.second.second);
  }
  pending.clear();
}

bool matrix_free_enabled() {
  const char* e = std::getenv("SPARLAB_NO_MATRIX_FREE");
  return !(e && e[0] == '1');
}

void op_apply(Context& c, const double* x, double* y) {
  if (c.mf) {
    kmf::MfArgs<double> a = mf_base<double>(c);
    a.x = x;
    a.y = y;
    kmf::apply(a, 0, c.stream);
  } else {
    k::spmv(c.K, x, y, c.stream);
  }
}

void op_residual(Context& c, const double* x, const double* b, double* r) { level_residual<double>(c, 0, x, b, r); }

int op_dot_blocks(const Context& c) {
  return c.mf ? kmf::blocks(c.nx + 1, c.ny + 1, c.nz + 1) : k::spmv_dot_blocks(c.K);
}

void op_apply_dot(Context& c, const double* x, double* y, double* partials) {
  if (c.mf) {
    kmf::MfArgs<double> a = mf_base<double>(c);
    a.x = x;
    a.y = y;
    a.partials = partials;
    kmf::apply(a, 1, c.stream);
  } else {
    k::spmv_dot(c.K, x, y, partials, c.stream);
  }
}

int op_lanczos(Context& c, const double* sq, const double* v, const double* vprev, const double* beta, double* w,
               double* partials) {
  if (c.mf) {
    kmf::MfArgs<double> a = mf_base<double>(c);
    a.x = v;
    a.y = w;
    a.in_scale = sq;
    a.vprev = vprev;
    a.beta = beta;
    a.partials = partials;
    kmf::apply(a, 4, c.stream);
    return op_dot_blocks(c);
  }
  k::lanczos_apply(c.K, sq, v, vprev, beta, w, partials, c.stream);
  return k::lanczos_blocks(c.K.rows);
}

void invalidate_graphs(Context& c) {
  for (hipGraphExec_t* g : {&c.g_update, &c.g_precond, &c.g_vupdate, &c.g_vprecond}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
}

void ensure_workspace(Context& c) {
  const size_t n = (size_t)c.K.rows;
  if (c.r.n != n) invalidate_graphs(c);
  c.xpcg.alloc(n);
  c.zold.alloc(n);
  c.r.alloc(n);
  c.z.alloc(n);
  c.p.alloc(n);
  c.q.alloc(n);
  c.xw.alloc(n);
  c.bw.alloc(n);
  const size_t need = std::max<size_t>(4096, (size_t)std::max(k::spmv_dot_blocks(c.K), op_dot_blocks(c)) + 8);
  if (c.partials.n < need) {
    invalidate_graphs(c);
    c.partials.alloc(need);
  }
  if (!c.scal.p) c.scal.alloc(16);
}

bool device_all_finite(Context& c, long long n, const double* v) {
  // the sum of squares is finite exactly when every entry is (for entries
  // below 1e150 in magnitude, far above any stiffness)
  const double s = device_dot(c, n, v, v);
  return std::isfinite(s);
}

void prepare_mixed(Context& c) {
  hipStream_t s = c.stream;
  for (size_t l = 0; l < c.levels.size(); ++l) {
    Level& lv = c.levels[l];
    DCsr& A = op(c, l);
    // (the Gauss-Seidel sweeps read level 0's assembled values even where its
    // other products are matrix-free)
    if (!use_mf(c, l) || (c.amg_opt.smoother == 1 && lv.has_next)) {
      A.val32.alloc((size_t)A.nnz);
      k::to_float(A.nnz, A.val.p, A.val32.p, s);
    } else {
      A.val32.release();
    }
    if (c.amg_opt.precision == 2 && l > 0) {  // coarse operators stored in bfloat16
      A.val16.alloc((size_t)A.nnz);
      k::to_bf16(A.nnz, A.val.p, A.val16.p, s);
    } else {
      A.val16.release();
    }
    if (lv.has_next && use_mf(c, l) && lv.tent_ok) {
      // the transfers read P-hat's aggregate blocks (FP32 copy made with them)
    } else if (lv.has_next && use_mf(c, l)) {
      lv.Ptent.val32.alloc((size_t)lv.Ptent.nnz);
      k::to_float(lv.Ptent.nnz, lv.Ptent.val.p, lv.Ptent.val32.p, s);
      lv.Rt.val32.alloc((size_t)lv.Rt.nnz);
      k::to_float(lv.Rt.nnz, lv.Rt.val.p, lv.Rt.val32.p, s);
    } else if (lv.has_next) {
      lv.P.val32.alloc((size_t)lv.P.nnz);
      k::to_float(lv.P.nnz, lv.P.val.p, lv.P.val32.p, s);
      lv.R.val32.alloc((size_t)lv.R.nnz);
      k::to_float(lv.R.nnz, lv.R.val.p, lv.R.val32.p, s);
    }
    lv.dinv32.alloc((size_t)lv.n);
    k::to_float(lv.n, lv.dinv.p, lv.dinv32.p, s);
    lv.x32.alloc((size_t)lv.n);
    lv.x232.alloc((size_t)lv.n);
    lv.b32.alloc((size_t)lv.n);
    lv.res32.alloc((size_t)lv.n);
    lv.d32.alloc((size_t)lv.n);
    lv.d232.alloc((size_t)lv.n);
  }
}

// One V-cycle on r. FP64 cycle: the result goes to z, returns null. FP32
// cycle: the result stays in level 0's FP32 iterate, which is returned.
const float* vcycle_raw(Context& c, const double* r, double* z) {
  if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "multigrid: apply() called before setup()");
  finish_coarse(c);
  return vcycle_cases(c, r, z, 1, false);
}

void vcycle(Context& c, const double* r, double* z) {
  const float* zc = vcycle_raw(c, r, z);
  if (zc) k::to_double((long long)c.levels[0].n, zc, z, c.stream);
}

namespace {

double smoother_ms(Context& c, int level, int reps) {
  Level& lv = c.levels[(size_t)level];
  DCsr& A = op(c, (size_t)level);
  const bool f32 = c.amg_opt.precision >= 1;
  hipStream_t s = c.stream;
  // finite inputs; c1 = c2 = 0 keeps d and x unchanged, every load and store still runs
  if (f32) {
    k::to_float(lv.n, lv.dinv.p, lv.d32.p, s);
    k::to_float(lv.n, lv.dinv.p, lv.res32.p, s);
  } else {
    k::copy(lv.n, lv.dinv.p, lv.d.p, s);
    k::copy(lv.n, lv.dinv.p, lv.res.p, s);
  }
  const bool mf = use_mf(c, (size_t)level);
  auto step = [&] {
    if (mf && f32) {
      kmf::MfArgs<float> a = mf_base<float>(c);
      a.x = lv.d32.p, a.b = lv.res32.p, a.y = lv.b32.p, a.dinv = lv.dinv32.p, a.d_out = lv.d232.p, a.xs = lv.x32.p;
      kmf::apply(a, 3, s);
    } else if (mf) {
      kmf::MfArgs<double> a = mf_base<double>(c);
      a.x = lv.d.p, a.b = lv.res.p, a.y = lv.b.p, a.dinv = lv.dinv.p, a.d_out = lv.d2.p, a.xs = lv.x.p;
      kmf::apply(a, 3, s);
    } else if (f32) {
      k::cheb_step32(A, lv.dinv32.p, lv.res32.p, lv.b32.p, lv.d32.p, lv.d232.p, lv.x32.p, 0.0, 0.0, s);
    } else {
      k::cheb_step(A, lv.dinv.p, lv.res.p, lv.b.p, lv.d.p, lv.d2.p, lv.x.p, 0.0, 0.0, s);
    }
  };
  step();  // warm
  hipEvent_t a = c.take_event(), b = c.take_event();
  SLX_HIP(hipEventRecord(a, s));
  for (int r = 0; r < reps; ++r) step();
  SLX_HIP(hipEventRecord(b, s));
  SLX_HIP(hipEventSynchronize(b));
  float ms = 0.0f;
  SLX_HIP(hipEventElapsedTime(&ms, a, b));
  c.event_pool.push_back(a);
  c.event_pool.push_back(b);
  return ms / reps;
}

}  // namespace

double time_smoother(Context& c, int level, int reps) {
  if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "multigrid: no hierarchy to time");
  if (level < 0 || level + 1 >= (int)c.levels.size() || reps < 1)
    fail(SLX_CONFIG_ERROR, "smoother timing: the level must lie above the coarsest");
  return smoother_ms(c, level, reps);
}

void tune_tile(Context& c) {
  c.mf_tile[0] = c.mf_tile[1] = 0;
  if (!c.mf || c.levels.size() < 2) return;
  const int slot = c.amg_opt.precision >= 1 ? 1 : 0;
  const char* f = std::getenv("SPARLAB_MF_TILE");
  if (f && f[0] >= '0' && f[0] < '0' + kmf::kTileShapes && f[1] == 0) {
    c.mf_tile[slot] = f[0] - '0';
    return;
  }
  const char* e = std::getenv("SPARLAB_NO_AUTOTUNE");
  if (e && e[0] == '1') return;
  float best = std::numeric_limits<float>::max();
  int pick = 0;
  for (int t = 0; t < kmf::kTileShapes; ++t) {
    c.mf_tile[slot] = t;
    // the best of three batches: one noisy batch must not pick a shape
    float ms = std::numeric_limits<float>::max();
    for (int batch = 0; batch < 3; ++batch) ms = std::min(ms, (float)smoother_ms(c, 0, 5));
    if (ms < best) {
      best = ms;
      pick = t;
    }
  }
  c.mf_tile[slot] = pick;
}

bool block_cases_enabled() {
  const char* e = std::getenv("SPARLAB_NO_BLOCK_CASES");
  return !(e && e[0] == '1');
}

// Replaying the captured iteration saved no time on MI355X at any size measured
// (from 48 x 24 x 24 Hex8 to 128 x 64 x 64, single solves and segmented batches),
// and every setup pays for a new capture, so the CG launches eagerly unless
// SPARLAB_GRAPHS=1 asks for graphs (SPARLAB_NO_GRAPHS=1 still forces eager).
bool graphs_enabled() {
  const char* no = std::getenv("SPARLAB_NO_GRAPHS");
  if (no && no[0] == '1') return false;
  const char* e = std::getenv("SPARLAB_GRAPHS");
  return e && e[0] == '1';
}

hipGraphExec_t capture_graph(Context& c, const std::function<void()>& body) {
  hipGraph_t graph = nullptr;
  c.capturing = true;
  SLX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(c.stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    c.capturing = false;
    throw;
  }
  SLX_HIP(hipStreamEndCapture(c.stream, &graph));
  c.capturing = false;
  hipGraphExec_t exec = nullptr;
  SLX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  SLX_HIP(hipGraphDestroy(graph));
  return exec;
}

slx_pcg_result pcg(Context& c, const double* b, double* x_out, double tol, int max_iterations) {
  finish_coarse(c);  // before any capture: the graphs read the installed inverse
  const long long n = c.K.rows;
  hipStream_t s = c.stream;
  slx_pcg_result res{0, 0.0, 0};
  const bool flexible = c.amg_opt.precision >= 1;
  const double bnorm = std::sqrt(device_dot(c, n, b, b));
  if (bnorm == 0.0) {
    k::fill(n, 0.0, x_out, s);
    res.converged = 1;
    return res;
  }
  const double target = tol * bnorm;
  int total = 0;
  bool recursive_converged = false;
  double best_true = std::numeric_limits<double>::max();
  double* x = c.xpcg.p;
  k::copy(n, x_out, x, s);
  const int nb = op_dot_blocks(c);
  // the three parts of one iteration (Multigrid.cpp:1184-1216)
  auto body_spmv = [&] { op_apply_dot(c, c.p.p, c.q.p, c.partials.p); };
  auto body_update = [&] {
    k::reduce(nb, c.partials.p, c.scal.p + S_PQ, 0, s);
    k::pcg_update(n, x, c.r.p, c.p.p, c.q.p, c.scal.p, S_RZ, S_PQ, c.partials.p, c.scal.p + S_RR, s);
  };
  auto body_precond = [&] {
    // z = M r; rz_new = r.z; beta = rz_new / rz_old, or for flexible CG
    // (Polak-Ribiere) r.(z - z_old) / rz_old; p = z + beta p
    const float* zc = vcycle_raw(c, c.r.p, c.z.p);
    k::precond_tail(n, zc, c.r.p, c.z.p, c.scal.p, S_RZ, S_RZN, S_TMP, flexible, c.partials.p, s);
    k::pcg_direction(n, c.p.p, c.z.p, c.scal.p, S_TMP, -1, s);
  };
  const bool graphs = graphs_enabled();
  c.h_scal.resize(16);
  hipEvent_t ev_scal = c.take_event();
  struct EventBack {
    Context& c;
    hipEvent_t e;
    ~EventBack() { c.event_pool.push_back(e); }
  } event_back{c, ev_scal};
  // the operator product is one kernel: launched directly, so the events
  // around it time the kernel alone; the other two parts replay graphs
  if (graphs && !c.g_update) {
    c.g_update = capture_graph(c, body_update);
    c.g_precond = capture_graph(c, body_precond);
  }
  for (int restart = 0; restart < 3; ++restart) {
    op_residual(c, x, b, c.r.p);
    double rnorm = std::sqrt(device_dot(c, n, c.r.p, c.r.p));
    res.relative_residual = rnorm / bnorm;
    if (rnorm <= target) {
      res.converged = 1;
      break;
    }
    vcycle(c, c.r.p, c.z.p);
    k::copy(n, c.z.p, c.p.p, s);
    k::dot(n, c.r.p, c.z.p, c.partials.p, c.scal.p + S_RZ, s);
    bool converged = false;
    double* hb = c.host_buffer(16);
    while (total < max_iterations) {
      {
        Timed t(c, "spmv");
        body_spmv();
      }
      if (graphs) SLX_HIP(hipGraphLaunch(c.g_update, s)); else body_update();
      ++total;
      // p.q and r.r to the host; the next preconditioner step is queued before
      // the host waits, so the device does not idle during the test (after
      // the last iteration it only touches z, p and the step scalars, which
      // every restart recomputes)
      SLX_HIP(hipMemcpyAsync(hb + S_PQ, c.scal.p + S_PQ, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
      SLX_HIP(hipEventRecord(ev_scal, s));
      if (graphs) SLX_HIP(hipGraphLaunch(c.g_precond, s)); else body_precond();
      {
        // poll: a blocking wait returns later than the copy completes
        hipError_t q;
        while ((q = hipEventQuery(ev_scal)) == hipErrorNotReady) {
        }
        SLX_HIP(q);
      }
      c.h_scal[S_PQ] = hb[S_PQ];
      c.h_scal[S_RR] = hb[S_RR];
      const double pq = c.h_scal[S_PQ];
      if (!(pq > 0.0) || !std::isfinite(pq)) {
        std::ostringstream os;
        os << "preconditioned CG broke down at iteration " << total << " (p^T A p = " << pq
           << "): the stiffness matrix or the multigrid preconditioner is not positive definite. Check the "
              "supports, or use a direct solver to confirm";
        fail(SLX_SOLVER_ERROR, os.str());
      }
      rnorm = std::sqrt(c.h_scal[S_RR]);
      if (rnorm <= target) {
        converged = true;
        break;
      }
    }
    op_residual(c, x, b, c.r.p);
    res.relative_residual = std::sqrt(device_dot(c, n, c.r.p, c.r.p)) / bnorm;
    if (res.relative_residual <= tol) {
      res.converged = 1;
      break;
    }
    if (!converged) break;  // iteration budget spent
    recursive_converged = true;
    if (!(res.relative_residual < 0.5 * best_true)) break;
    best_true = res.relative_residual;
  }
  if (!res.converged && recursive_converged) res.converged = 1;
  res.iterations = total;
  k::copy(n, x, x_out, s);
  c.flush_timers();
  return res;
}


void ensure_block_workspace(Context& c, int nv) {
  const size_t n = (size_t)c.K.rows;
  if (c.block_nv != nv || c.block_precision != c.amg_opt.precision || c.vr.n != n * (size_t)nv)
    invalidate_graphs(c);
  c.block_nv = nv;
  c.block_precision = c.amg_opt.precision;
  for (DBuf<double>* v : {&c.vx, &c.vr, &c.vz, &c.vp, &c.vq}) v->alloc(n * (size_t)nv);
  const bool f32 = c.amg_opt.precision >= 1;
  for (Level& lv : c.levels) {
    const size_t m = (size_t)lv.n * (size_t)nv;
    for (DBuf<double>* v : {&lv.vx, &lv.vx2, &lv.vb, &lv.vres, &lv.vd, &lv.vd2}) {
      if (f32) v->release(); else v->alloc(m);
    }
    for (DBuf<float>* v : {&lv.vx32, &lv.vx232, &lv.vb32, &lv.vres32, &lv.vd32, &lv.vd232}) {
      if (f32) v->alloc(m); else v->release();
    }
  }
  const size_t per = (size_t)std::max({op_dot_blocks(c), k::spmv_dot_blocks(c.K), 2 * k::reduce_blocks((long long)n)});
  if (c.vpart.n < per * (size_t)nv + 8) {
    invalidate_graphs(c);
    c.vpart.alloc(per * (size_t)nv + 8);
  }
  if (c.vscal.n < 16 * (size_t)nv) {
    invalidate_graphs(c);
    c.vscal.alloc(16 * (size_t)nv);
  }
  if (c.vactive.n < (size_t)nv) {
    invalidate_graphs(c);
    c.vactive.alloc((size_t)nv);
  }
}

// PCG for nv load cases at once (Multigrid.cpp:1152-1237 per case): one
// operator product, one V-cycle pass and one set of reductions per iteration
// for all cases still iterating. Each case follows the single-case pcg()
// step for step: the same restarts, tests and iteration budget, and the same
// arithmetic, so its iterate, residual and iteration count equal a pcg() call
// on that case alone. A case that has met its tolerance is frozen (its x and
// r stay) while the others continue. b and x are case-major (nv x n).
void pcg_block(Context& c, int nv, const double* b, double* x_io, double tol, int max_iterations,
               std::vector<slx_pcg_result>& out) {
  if (nv < 1) fail(SLX_CONFIG_ERROR, "pcg_block: at least one load case");
  if (!c.amg_ready) fail(SLX_SOLVER_ERROR, "multigrid: apply() called before setup()");
  finish_coarse(c);
  ensure_block_workspace(c, nv);
  const long long n = c.K.rows;
  hipStream_t s = c.stream;
  const int nb = op_dot_blocks(c);
  double* scal = c.vscal.p;
  double* part = c.vpart.p;
  double* x = c.vx.p;  // the iterates (the graphs bind this buffer, not the caller's)
  k::copy(n * nv, x_io, x, s);
  // pinned host memory: 16 scalars per case, then the active flags
  double* hb = c.host_buffer(16 * (size_t)nv + (size_t)nv);
  int* hflags = reinterpret_cast<int*>(hb + 16 * (size_t)nv);
  auto read_scalars = [&](int slot) {  // scal[16 v + slot] of every case to hb
    SLX_HIP(hipMemcpyAsync(hb, scal, 16 * (size_t)nv * sizeof(double), hipMemcpyDeviceToHost, s));
    SLX_HIP(hipStreamSynchronize(s));
    std::vector<double> v((size_t)nv);
    for (int i = 0; i < nv; ++i) v[(size_t)i] = hb[16 * i + slot];
    return v;
  };
  auto apply_dot = [&] {  // q = K p and the partials of p.q, per case
    Timed t(c, "block_spmv");
    if (c.mf) {
      for (int v = 0; v < nv; ++v) op_apply_dot(c, c.vp.p + v * n, c.vq.p + v * n, part + (size_t)v * nb);
    } else {
      k::spmv_dot(c.K, c.vp.p, c.vq.p, part, s, nv);
    }
  };
  auto body_update = [&] {
    k::reduce(nb, part, scal + S_PQ, 0, s, nv, 16);
    k::pcg_update(n, x, c.vr.p, c.vp.p, c.vq.p, scal, S_RZ, S_PQ, part, scal + S_RR, s, nv, c.vactive.p);
  };
  const bool flexible = c.amg_opt.precision >= 1;
  auto body_precond = [&] {
    // as pcg(): z = M r, r.z (flexible CG: Polak-Ribiere beta), p = z + beta p
    const float* zc = block_vcycle_raw(c, c.vr.p, c.vz.p, nv);
    k::precond_tail(n, zc, c.vr.p, c.vz.p, scal, S_RZ, S_RZN, S_TMP, flexible, part, s, nv);
    k::pcg_direction(n, c.vp.p, c.vz.p, scal, S_TMP, -1, s, nv);
  };
  const bool graphs = graphs_enabled();
  if (graphs && !c.g_vupdate) {
    c.g_vupdate = capture_graph(c, body_update);
    c.g_vprecond = capture_graph(c, body_precond);
  }
  out.assign((size_t)nv, slx_pcg_result{0, 0.0, 0});
  // ||b|| per case
  k::dot(n, b, b, part, scal + S_TMP, s, nv, 16);
  std::vector<double> bnorm = read_scalars(S_TMP), target((size_t)nv), best((size_t)nv,
                                                                                std::numeric_limits<double>::max());
  std::vector<int> total((size_t)nv, 0);
  std::vector<char> done((size_t)nv, 0), rec_conv((size_t)nv, 0), loop_conv((size_t)nv, 0), live((size_t)nv, 0);
  for (int v = 0; v < nv; ++v) {
    bnorm[(size_t)v] = std::sqrt(bnorm[(size_t)v]);
    target[(size_t)v] = tol * bnorm[(size_t)v];
    if (bnorm[(size_t)v] == 0.0) {
      k::fill(n, 0.0, x + v * n, s);  // as pcg(): a zero load gives x = 0
      out[(size_t)v].converged = 1;
      done[(size_t)v] = 1;
    }
  }
  auto upload_flags = [&] {
    for (int v = 0; v < nv; ++v) hflags[v] = live[(size_t)v] ? 1 : 0;
    SLX_HIP(hipMemcpyAsync(c.vactive.p, hflags, sizeof(int) * (size_t)nv, hipMemcpyHostToDevice, s));
  };
  auto true_residuals = [&] {  // r = b - K x and r.r per case
    block_residual(c, 0, x, b, c.vr.p, nv);
    k::dot(n, c.vr.p, c.vr.p, part, scal + S_TMP, s, nv, 16);
    return read_scalars(S_TMP);
  };
  hipEvent_t ev = c.take_event();
  struct EventBack {
    Context& c;
    hipEvent_t e;
    ~EventBack() { c.event_pool.push_back(e); }
  } event_back{c, ev};
  for (int restart = 0; restart < 3; ++restart) {
    bool any = false;
    for (int v = 0; v < nv; ++v) any = any || !done[(size_t)v];
    if (!any) break;
    const std::vector<double> rr0 = true_residuals();
    bool any_live = false;
    for (int v = 0; v < nv; ++v) {
      live[(size_t)v] = 0;
      loop_conv[(size_t)v] = 0;
      if (done[(size_t)v]) continue;
      const double rn = std::sqrt(rr0[(size_t)v]);
      out[(size_t)v].relative_residual = rn / bnorm[(size_t)v];
      if (rn <= target[(size_t)v]) {
        out[(size_t)v].converged = 1;
        done[(size_t)v] = 1;
      } else {
        live[(size_t)v] = 1;
        any_live = true;
      }
    }
    if (!any_live) break;
    std::vector<char> in_round = live;
    {
      const float* zc = block_vcycle_raw(c, c.vr.p, c.vz.p, nv);
      if (zc) k::to_double(n * nv, zc, c.vz.p, s);
    }
    k::copy(n * nv, c.vz.p, c.vp.p, s);
    k::dot(n, c.vr.p, c.vz.p, part, scal + S_RZ, s, nv, 16);
    std::vector<char> sent((size_t)nv, 2);
    for (;;) {
      bool running = false;
      for (int v = 0; v < nv; ++v) {
        if (live[(size_t)v] && total[(size_t)v] >= max_iterations) live[(size_t)v] = 0;  // budget spent
        running = running || live[(size_t)v];
      }
      if (!running) break;
      if (live != sent) {
        upload_flags();
        sent = live;
      }
      apply_dot();
      if (graphs) SLX_HIP(hipGraphLaunch(c.g_vupdate, s)); else body_update();
      for (int v = 0; v < nv; ++v) total[(size_t)v] += live[(size_t)v] ? 1 : 0;
      SLX_HIP(hipMemcpyAsync(hb, scal, 16 * (size_t)nv * sizeof(double), hipMemcpyDeviceToHost, s));
      SLX_HIP(hipEventRecord(ev, s));
      if (graphs) SLX_HIP(hipGraphLaunch(c.g_vprecond, s)); else body_precond();
      {
        hipError_t q;
        while ((q = hipEventQuery(ev)) == hipErrorNotReady) {
        }
        SLX_HIP(q);
      }
      for (int v = 0; v < nv; ++v) {
        if (!live[(size_t)v]) continue;
        const double pq = hb[16 * v + S_PQ];
        if (!(pq > 0.0) || !std::isfinite(pq)) {
          std::ostringstream os;
          os << "preconditioned CG broke down at iteration " << total[(size_t)v] << " of load case " << v
             << " (p^T A p = " << pq
             << "): the stiffness matrix or the multigrid preconditioner is not positive definite. Check the "
                "supports, or use a direct solver to confirm";
          fail(SLX_SOLVER_ERROR, os.str());
        }
        if (std::sqrt(hb[16 * v + S_RR]) <= target[(size_t)v]) {
          loop_conv[(size_t)v] = 1;
          live[(size_t)v] = 0;
        }
      }
    }
    const std::vector<double> rr = true_residuals();
    for (int v = 0; v < nv; ++v) {
      if (!in_round[(size_t)v]) continue;
      slx_pcg_result& r = out[(size_t)v];
      r.relative_residual = std::sqrt(rr[(size_t)v]) / bnorm[(size_t)v];
      if (r.relative_residual <= tol) {
        r.converged = 1;
        done[(size_t)v] = 1;
      } else if (!loop_conv[(size_t)v]) {
        done[(size_t)v] = 1;  // iteration budget spent
      } else {
        rec_conv[(size_t)v] = 1;
        if (!(r.relative_residual < 0.5 * best[(size_t)v])) done[(size_t)v] = 1;
        best[(size_t)v] = r.relative_residual;
      }
    }
  }
  for (int v = 0; v < nv; ++v) {
    if (!out[(size_t)v].converged && rec_conv[(size_t)v]) out[(size_t)v].converged = 1;
    out[(size_t)v].iterations = total[(size_t)v];
  }
  k::copy(n * nv, x, x_io, s);
  c.flush_timers();
}

}  // namespace slx
