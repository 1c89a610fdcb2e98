// Free vibration of a structured model: the lowest eigenpairs of (K_ff, M_ff)
// by shift-invert subspace iteration (src/fem/ModalAnalysis.cpp:120-418), or,
// for a system of at most 400 free DOFs or one the subspace covers, by the dense
// generalised eigensolve of the reference's small-system branch. The
// q shifted systems of each iteration are solved with the context's AMG-PCG,
// each warm started from the Ritz vector divided by its Ritz value; the mass
// products, the projections x^T K x and x^T M x, the rotation to the Ritz basis
// and the eigenpair residuals run on the device. Only the q x q projected
// pencil goes to the host, where it is solved by Cholesky reduction and cyclic
// Jacobi rotations.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>
#include <sstream>

#include "slx_kernels.hpp"
#include "slx_modal.hpp"
#include "slx_solver.hpp"

namespace slx {
namespace {

// a = L L^T in place (row-major, lower triangle); false when a is not positive definite
bool cholesky(std::vector<double>& a, int n) {
  for (int j = 0; j < n; ++j) {
    double d = a[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    const double l = std::sqrt(d);
    a[(size_t)j * n + j] = l;
    for (int i = j + 1; i < n; ++i) {
      double v = a[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) v -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
      a[(size_t)i * n + j] = v / l;
    }
    for (int k = j + 1; k < n; ++k) a[(size_t)j * n + k] = 0.0;
  }
  return true;
}

// eigenvalues w (ascending) and orthonormal eigenvectors v (column j = v[i * n + j])
// of the symmetric matrix a, by cyclic Jacobi rotations (Golub and Van Loan, 8.5)
void jacobi_eigen(std::vector<double> a, int n, std::vector<double>& w, std::vector<double>& v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) (i == j ? diag : off) += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = cs * akp - sn * akq;
          a[(size_t)k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = cs * apk - sn * aqk;
          a[(size_t)q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = cs * vkp - sn * vkq;
          v[(size_t)k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  std::vector<int> order((size_t)n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  w.resize((size_t)n);
  std::vector<double> vs((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[(size_t)j] = a[(size_t)order[(size_t)j] * n + order[(size_t)j]];
    for (int i = 0; i < n; ++i) vs[(size_t)i * n + j] = v[(size_t)i * n + order[(size_t)j]];
  }
  v.swap(vs);
}

// eigenpairs of the pencil (A, B), B = L L^T: C = L^-1 A L^-T, vectors L^-T y
// (B-orthonormal); lam ascending
void reduced_pencil(const std::vector<double>& A, const std::vector<double>& L, int n, std::vector<double>& lam,
                    std::vector<double>& vec) {
  // W = L^-1 A (forward substitution on each column), C = L^-1 W^T
  std::vector<double> W((size_t)n * n), C((size_t)n * n);
  for (int col = 0; col < n; ++col)
    for (int i = 0; i < n; ++i) {
      double v = A[(size_t)i * n + col];
      for (int k = 0; k < i; ++k) v -= L[(size_t)i * n + k] * W[(size_t)k * n + col];
      W[(size_t)i * n + col] = v / L[(size_t)i * n + i];
    }
  for (int col = 0; col < n; ++col)
    for (int i = 0; i < n; ++i) {
      double v = W[(size_t)col * n + i];
      for (int k = 0; k < i; ++k) v -= L[(size_t)i * n + k] * C[(size_t)k * n + col];
      C[(size_t)i * n + col] = v / L[(size_t)i * n + i];
    }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double s = 0.5 * (C[(size_t)i * n + j] + C[(size_t)j * n + i]);
      C[(size_t)i * n + j] = C[(size_t)j * n + i] = s;
    }
  std::vector<double> Y;
  jacobi_eigen(C, n, lam, Y);
  // vec = L^-T Y (back substitution on each column)
  vec.assign((size_t)n * n, 0.0);
  for (int col = 0; col < n; ++col)
    for (int i = n - 1; i >= 0; --i) {
      double v = Y[(size_t)i * n + col];
      for (int k = i + 1; k < n; ++k) v -= L[(size_t)k * n + i] * vec[(size_t)k * n + col];
      vec[(size_t)i * n + col] = v / L[(size_t)i * n + i];
    }
}

// lowest_pencil_pairs (ModalAnalysis.cpp:27-50): the `count` lowest pairs of
// (k, m), through the reversed pencil (m, k) when m is not positive definite.
// vectors: row-major n x count
bool lowest_pencil_pairs(const std::vector<double>& k, const std::vector<double>& m, int n, int count,
                         std::vector<double>& lambda, std::vector<double>& vectors) {
  std::vector<double> lam, vec;
  std::vector<double> L = m;
  if (cholesky(L, n)) {
    reduced_pencil(k, L, n, lam, vec);
    lambda.assign(lam.begin(), lam.begin() + count);
    vectors.resize((size_t)n * count);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < count; ++j) vectors[(size_t)i * count + j] = vec[(size_t)i * n + j];
    return true;
  }
  L = k;
  if (!cholesky(L, n)) return false;
  reduced_pencil(m, L, n, lam, vec);
  lambda.resize((size_t)count);
  vectors.resize((size_t)n * count);
  for (int j = 0; j < count; ++j) {
    const double mu = lam[(size_t)(n - 1 - j)];
    if (!(mu > 0.0)) return false;
    lambda[(size_t)j] = 1.0 / mu;
    for (int i = 0; i < n; ++i) vectors[(size_t)i * count + j] = vec[(size_t)i * n + (n - 1 - j)];
  }
  return true;
}

}  // namespace

void modal_solve(Context& c, const slx_modal_options& o, const double* d_mscale, double total_mass, double* eig_out,
                 double* res_out, double* shapes, slx_modal_result& R) {
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = c.stream;
  const int n = c.K.rows;
  if (n == 0) fail(SLX_MODEL_ERROR, "modal analysis has no free degrees of freedom; every DOF is constrained");
  const int m = std::min(o.num_modes, n);
  const int q = std::min(n, std::max(2 * m, m + 8));
  // a subspace that covers the system, or a small system, takes the dense
  // generalised eigensolve (ModalAnalysis.cpp:166-179)
  const bool dense = q >= n || n <= 400;
  if (!dense && q > 64)
    fail(SLX_CONFIG_ERROR, "modal analysis on the device supports at most 32 modes (a subspace of 64)");
  R = slx_modal_result{};
  R.num_modes = m;
  R.subspace_size = q;
  R.total_mass = total_mass;
  const double rho_v = o.density * c.volumes[0];
  const bool lumped = o.lumped != 0;
  const size_t nq = (size_t)n * q;
  DBuf<double> X, Xb, Y, KXb, MXb, KX, MX, kd, md, G, parts, dmu, dritz;
  X.alloc(nq);
  Xb.alloc(nq);
  Y.alloc(nq);
  KXb.alloc(nq);
  MXb.alloc(nq);
  KX.alloc((size_t)n * m);
  MX.alloc((size_t)n * m);
  kd.alloc((size_t)n);
  md.alloc((size_t)n);
  G.alloc((size_t)2 * q * q);
  parts.alloc((size_t)k::modal_partials(n, q));
  dmu.alloc((size_t)m);
  dritz.alloc((size_t)q * q);
  auto mass = [&](const double* x, double* y) {
    k::mass_apply(c.dim, c.nx, c.ny, c.nz, rho_v, lumped, false, d_mscale, c.d_reduced.p, x, y, s);
  };

  // starting_subspace (ModalAnalysis.cpp:52-86): the mass diagonal, then unit
  // vectors at the DOFs with the largest m_ii / k_ii (stable order). With
  // q <= n the seeded random columns of the reference are never reached.
  std::vector<double> hk((size_t)n), hm((size_t)n);
  k::csr_diagonal(c.K, kd.p, s);
  k::mass_apply(c.dim, c.nx, c.ny, c.nz, rho_v, lumped, true, d_mscale, c.d_reduced.p, nullptr, md.p, s);
  kd.download(hk.data(), (size_t)n, s);
  md.download(hm.data(), (size_t)n, s);
  SLX_HIP(hipStreamSynchronize(s));
  {
    std::vector<double> x0(nq, 0.0);
    std::copy(hm.begin(), hm.end(), x0.begin());
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const double ra = hk[(size_t)a] > 0.0 ? hm[(size_t)a] / hk[(size_t)a] : 0.0;
      const double rb = hk[(size_t)b] > 0.0 ? hm[(size_t)b] / hk[(size_t)b] : 0.0;
      return ra > rb;
    });
    for (int col = 1; col < q; ++col) x0[(size_t)col * n + order[(size_t)(col - 1)]] = 1.0;
    X.upload(x0.data(), nq, s);
  }
  // the largest coefficient of a symmetric positive definite matrix is on its
  // diagonal, so max |K_ij| / max |M_ij| (ModalAnalysis.cpp:326-331) is the ratio of the diagonal maxima
  const double kmax = *std::max_element(hk.begin(), hk.end());
  const double mmax = *std::max_element(hm.begin(), hm.end());
  const double lambda_scale = mmax > 0.0 ? kmax / mmax : 1.0;

  std::vector<double> prev((size_t)m, std::numeric_limits<double>::max()), mu_prev, mu, ritz, g((size_t)2 * q * q);
  std::vector<double> ritz_cm((size_t)q * q), resid((size_t)2 * m), last_res((size_t)m, 0.0);
  const double* phi_src = nullptr;
  if (dense) {
    // K_ff and M_ff as dense matrices (K from its CSR, M column by column from
    // the device's mass product), the pencil solved as the projected one is
    std::vector<double> kn((size_t)n * n, 0.0), mn((size_t)n * n);
    {
      std::vector<int> hp((size_t)n + 1), hc((size_t)c.K.nnz);
      std::vector<double> hv((size_t)c.K.nnz);
      c.K.ptr.download(hp.data(), hp.size(), s);
      c.K.col.download(hc.data(), hc.size(), s);
      c.K.val.download(hv.data(), hv.size(), s);
      SLX_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < n; ++i)
        for (int k = hp[(size_t)i]; k < hp[(size_t)i + 1]; ++k) kn[(size_t)i * n + hc[(size_t)k]] = hv[(size_t)k];
      DBuf<double> E, ME;
      std::vector<double> eye((size_t)n * n, 0.0);
      for (int i = 0; i < n; ++i) eye[(size_t)i * n + i] = 1.0;
      E.upload(eye, s);
      ME.alloc((size_t)n * n);
      for (int col = 0; col < n; ++col) mass(E.p + (size_t)col * n, ME.p + (size_t)col * n);
      ME.download(mn.data(), mn.size(), s);  // column j = M e_j; M is symmetric
      SLX_HIP(hipStreamSynchronize(s));
    }
    std::vector<double> lam, vec;
    if (!lowest_pencil_pairs(kn, mn, n, m, lam, vec))
      fail(SLX_SOLVER_ERROR,
           "dense generalised eigensolve failed: the mass matrix is not positive definite and the stiffness matrix "
           "is not either, or fewer DOFs than the modes requested carry mass (check material.density and the mass "
           "interpolation floor)");
    // phi^T M phi = 1 and the eigenpair residuals |K phi - lambda M phi| / |lambda M phi|
    // (ModalAnalysis.cpp:311-322,366-368)
    std::vector<double> phi((size_t)n * m), kp((size_t)n), mp((size_t)n);
    for (int j = 0; j < m; ++j) {
      for (int i = 0; i < n; ++i) {
        double a = 0.0;
        for (int k = 0; k < n; ++k) a += mn[(size_t)i * n + k] * vec[(size_t)k * m + j];
        mp[(size_t)i] = a;
      }
      double mm = 0.0;
      for (int i = 0; i < n; ++i) mm += vec[(size_t)i * m + j] * mp[(size_t)i];
      const double mnorm = std::sqrt(std::max(mm, 0.0));
      if (!(mnorm > 0.0))
        fail(SLX_SOLVER_ERROR,
             "a computed mode shape has zero modal mass; the mass matrix is singular on the free DOFs");
      for (int i = 0; i < n; ++i) phi[(size_t)j * n + i] = vec[(size_t)i * m + j] / mnorm;
      double rr = 0.0, ref = 0.0;
      for (int i = 0; i < n; ++i) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < n; ++k) {
          a += kn[(size_t)i * n + k] * phi[(size_t)j * n + k];
          b += mn[(size_t)i * n + k] * phi[(size_t)j * n + k];
        }
        const double lm = lam[(size_t)j] * b;
        rr += (a - lm) * (a - lm);
        ref += lm * lm;
      }
      prev[(size_t)j] = lam[(size_t)j];
      last_res[(size_t)j] = std::sqrt(rr) / std::max(std::sqrt(ref), 1.0e-300);
    }
    KX.upload(phi.data(), phi.size(), s);
    R.converged = 1;
    R.iterations = 0;
    R.final_change = 0.0;
  }
  for (int iter = 1; !dense && iter <= o.max_iterations; ++iter) {
    for (int col = 0; col < q; ++col) mass(X.p + (size_t)col * n, Y.p + (size_t)col * n);
    for (int col = 0; col < q; ++col) {
      double* xb = Xb.p + (size_t)col * n;
      // near convergence K^-1 M x = x / lambda: the last Ritz pair is the guess
      // (with warm_start set, as linear.warm_start in the reference)
      if (o.warm_start && mu_prev.size() == (size_t)q && mu_prev[(size_t)col] > 0.0)
        k::scale_copy(n, 1.0 / mu_prev[(size_t)col], X.p + (size_t)col * n, xb, s);
      else
        k::fill(n, 0.0, xb, s);
      const slx_pcg_result r = pcg(c, Y.p + (size_t)col * n, xb, o.linear_tolerance,
                                   o.linear_max_iterations > 0 ? o.linear_max_iterations : 1000);
      R.linear_iterations += r.iterations;
      if (!r.converged) {
        std::ostringstream os;
        os << "AMG-CG did not converge in a modal subspace solve: " << r.iterations
           << " iterations left a relative residual of " << r.relative_residual << " against a tolerance of "
           << o.linear_tolerance;
        fail(SLX_CONVERGENCE_ERROR, os.str());
      }
    }
    for (int col = 0; col < q; ++col) {
      op_apply(c, Xb.p + (size_t)col * n, KXb.p + (size_t)col * n);
      mass(Xb.p + (size_t)col * n, MXb.p + (size_t)col * n);
    }
    k::gram(n, q, Xb.p, KXb.p, parts.p, G.p, s);
    k::gram(n, q, Xb.p, MXb.p, parts.p, G.p + (size_t)q * q, s);
    G.download(g.data(), (size_t)2 * q * q, s);
    SLX_HIP(hipStreamSynchronize(s));
    std::vector<double> kr((size_t)q * q), mr((size_t)q * q);
    bool finite = true;
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j) {
        kr[(size_t)i * q + j] = 0.5 * (g[(size_t)i * q + j] + g[(size_t)j * q + i]);
        mr[(size_t)i * q + j] = 0.5 * (g[(size_t)q * q + (size_t)i * q + j] + g[(size_t)q * q + (size_t)j * q + i]);
        finite = finite && std::isfinite(kr[(size_t)i * q + j]) && std::isfinite(mr[(size_t)i * q + j]);
      }
    if (!finite)
      fail(SLX_SOLVER_ERROR,
           "subspace iteration produced a non-finite basis; the shifted matrix K - sigma M is singular. Reduce "
           "modal.shift_factor");
    if (!lowest_pencil_pairs(kr, mr, q, q, mu, ritz)) {
      std::ostringstream os;
      os << "the projected " << q << " x " << q << " eigenproblem became numerically singular at subspace iteration "
         << iter << "; the trial vectors lost independence. Reduce modal.num_modes or tighten modal.tolerance";
      fail(SLX_SOLVER_ERROR, os.str());
    }
    // x = xbar ritz (Ritz vectors column-major on the device)
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j) ritz_cm[(size_t)j * q + i] = ritz[(size_t)i * q + j];
    dritz.upload(ritz_cm.data(), (size_t)q * q, s);
    k::block_combine(n, q, q, Xb.p, dritz.p, X.p, s);
    mu_prev = mu;
    double change = 0.0;
    for (int i = 0; i < m; ++i) {
      const double denom = std::max(std::fabs(mu[(size_t)i]), 1.0e-300);
      change = std::max(change, std::fabs(mu[(size_t)i] - prev[(size_t)i]) / denom);
    }
    for (int i = 0; i < m; ++i) prev[(size_t)i] = mu[(size_t)i];
    R.iterations = iter;
    R.final_change = change;
    // eigenpair residuals of the m requested modes: K x = (K xbar) ritz, M x = (M xbar) ritz
    k::block_combine(n, q, m, KXb.p, dritz.p, KX.p, s);
    k::block_combine(n, q, m, MXb.p, dritz.p, MX.p, s);
    dmu.upload(mu.data(), (size_t)m, s);
    k::eig_residuals(n, m, KX.p, MX.p, dmu.p, parts.p, G.p, s);
    G.download(resid.data(), (size_t)2 * m, s);
    SLX_HIP(hipStreamSynchronize(s));
    double max_res = 0.0;
    for (int i = 0; i < m; ++i) {
      const double ref = std::max(std::sqrt(resid[(size_t)2 * i + 1]), 1.0e-300);
      last_res[(size_t)i] = std::sqrt(resid[(size_t)2 * i]) / ref;
      max_res = std::max(max_res, last_res[(size_t)i]);
    }
    if (change < o.tolerance && max_res < o.residual_tolerance) {
      R.converged = 1;
      phi_src = X.p;
      break;
    }
  }
  if (!R.converged) {
    std::ostringstream os;
    os << "subspace iteration did not converge in " << o.max_iterations
       << " iterations; the relative eigenvalue change stalled at " << R.final_change << " against a tolerance of "
       << o.tolerance << " (the stopping rule also requires every eigenpair residual below " << o.residual_tolerance
       << "). Increase modal.max_iterations, loosen modal.tolerance or modal.residual_tolerance, or request fewer "
          "modes";
    fail(SLX_CONVERGENCE_ERROR, os.str());
  }

  // phi^T M phi = 1 (ModalAnalysis.cpp:311-322), through M x = (M xbar) ritz
  // (the dense branch normalised its vectors on the host)
  if (!dense) {
    k::gram(n, m, phi_src, MX.p, parts.p, G.p, s);
    std::vector<double> gm((size_t)m * m);
    G.download(gm.data(), (size_t)m * m, s);
    SLX_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < m; ++i) {
      const double mnorm = std::sqrt(std::max(gm[(size_t)i * m + i], 0.0));
      if (!(mnorm > 0.0))
        fail(SLX_SOLVER_ERROR,
             "a computed mode shape has zero modal mass; the mass matrix is singular on the free DOFs");
      k::scale_copy(n, 1.0 / mnorm, phi_src + (size_t)i * n, KX.p + (size_t)i * n, s);
    }
  }
  // validity screening (ModalAnalysis.cpp:333-375)
  for (int i = 0; i < m; ++i) {
    const double lam = prev[(size_t)i];
    if (!std::isfinite(lam)) {
      std::ostringstream os;
      os << "eigenvalue " << i << " is not finite; the eigenproblem is ill-posed";
      fail(SLX_SOLVER_ERROR, os.str());
    }
    if (lam < -o.rigid_body_ratio * lambda_scale) {
      std::ostringstream os;
      os << "eigenvalue " << i << " is negative (" << lam
         << " 1/s^2). K_ff is not positive definite, which is physically impossible for a properly constrained "
            "linear-elastic model: check the boundary conditions and the SIMP stiffness floor";
      fail(SLX_SOLVER_ERROR, os.str());
    }
    if (lam < o.rigid_body_ratio * lambda_scale) ++R.rigid_body_modes;
    if (last_res[(size_t)i] > o.residual_tolerance) ++R.residual_warnings;
    if (eig_out) eig_out[i] = lam;
    if (res_out) res_out[i] = last_res[(size_t)i];
  }
  if (shapes) {
    std::vector<double> phi((size_t)n * m);
    SLX_HIP(hipMemcpyAsync(phi.data(), KX.p, sizeof(double) * phi.size(), hipMemcpyDeviceToHost, s));
    SLX_HIP(hipStreamSynchronize(s));
    const size_t nd = (size_t)c.ndofs;
    std::fill(shapes, shapes + nd * m, 0.0);
    for (int i = 0; i < m; ++i)
      for (int r = 0; r < n; ++r) shapes[(size_t)i * nd + c.free_dofs[(size_t)r]] = phi[(size_t)i * n + r];
  }
  SLX_HIP(hipStreamSynchronize(s));
  R.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace slx
