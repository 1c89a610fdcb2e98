// SA-AMG hierarchy of the sparlab_hip library (src/fem/Multigrid.cpp:383-1111).
//
// Split between host and device:
//  * once per sparsity structure, on the host: node blocks and the rigid-body
//    near-null space, strength of connection and the three greedy aggregation
//    passes (serial by definition: the aggregates must equal the reference's bit
//    for bit), the tentative prolongators (one small column-pivoted QR per
//    aggregate) and the symbolic patterns of P, R = P^T, A P, R (A P) and the
//    symmetrised coarse operator;
//  * at every setup, on the device: diagonal, Lanczos estimate of
//    lambda_max(D^-1 A), prolongator smoothing, the two Galerkin products and
//    the symmetrisation, level by level; the coarsest operator (at most a few
//    thousand unknowns) comes back to the host once for its factorisation and
//    returns as a dense inverse, which the V-cycle applies as one GEMV.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <sstream>

#include "slx_kernels.hpp"
#include "slx_mf.hpp"
#include "slx_solver.hpp"

namespace slx {

void DCsr::upload_pattern(const HCsr& h, hipStream_t s) {
  rows = h.rows;
  cols = h.cols;
  nnz = h.nnz();
  max_row = 0;
  for (int i = 0; i < h.rows; ++i) max_row = std::max(max_row, h.ptr[(size_t)i + 1] - h.ptr[(size_t)i]);
  ptr.upload(h.ptr, s);
  col.upload(h.col, s);
  val.alloc((size_t)nnz + 8);  // padding: the row products read up to 8 values from a run's start
  cbase.release();
  ncol.release();
}

namespace {

using Clock = std::chrono::steady_clock;

// SPARLAB_SETUP_TRACE=1: the host steps of a full setup, timed to stderr
bool setup_trace_on() {
  static const bool on = [] {
    const char* e = std::getenv("SPARLAB_SETUP_TRACE");
    return e && e[0] == '1';
  }();
  return on;
}
Clock::time_point g_trace_clock;
void setup_trace_start() { g_trace_clock = Clock::now(); }
void setup_trace(const char* what, int level, hipStream_t s) {
  if (!setup_trace_on()) return;
  SLX_HIP(hipStreamSynchronize(s));
  const auto now = Clock::now();
  std::fprintf(stderr, "setup trace: level %d %-28s %9.3f ms\n", level, what,
               1e3 * std::chrono::duration<double>(now - g_trace_clock).count());
  g_trace_clock = now;
}

// ---------------------------------------------------------------------------
// Symbolic sparse algebra (patterns only; rows sorted by column)
// ---------------------------------------------------------------------------
HCsr pattern_product(const HCsr& a, const HCsr& b) {
  // two passes over the rows (count, then fill and sort in place), each row on
  // one thread with its own marker array
  HCsr c;
  c.rows = a.rows;
  c.cols = b.cols;
  c.ptr.assign((size_t)a.rows + 1, 0);
  auto walk = [&](bool fill) {
#pragma omp parallel
    {
      std::vector<int> marker((size_t)b.cols, -1);
#pragma omp for schedule(dynamic, 512)
      for (int i = 0; i < a.rows; ++i) {
        int cnt = 0;
        int* out = fill ? c.col.data() + c.ptr[(size_t)i] : nullptr;
        for (int k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
          const int j = a.col[(size_t)k];
          for (int kk = b.ptr[(size_t)j]; kk < b.ptr[(size_t)j + 1]; ++kk) {
            const int cc = b.col[(size_t)kk];
            if (marker[(size_t)cc] != i) {
              marker[(size_t)cc] = i;
              if (fill) out[cnt] = cc;
              ++cnt;
            }
          }
        }
        // (the marker holds row numbers: each row of one pass stamps its own)
        if (fill) std::sort(out, out + cnt);
        else c.ptr[(size_t)i + 1] = cnt;
      }
    }
  };
  walk(false);
  for (int i = 0; i < a.rows; ++i) {
    const long long next = (long long)c.ptr[(size_t)i + 1] + c.ptr[(size_t)i];
    if (next > std::numeric_limits<int>::max()) fail(SLX_SOLVER_ERROR, "multigrid: a Galerkin product exceeds 2^31 entries");
    c.ptr[(size_t)i + 1] = (int)next;
  }
  c.col.resize((size_t)c.nnz());
  walk(true);
  return c;
}

// R = P^T with src[k] = index in P of R's entry k; columns of R ascend.
void transpose_pattern(const HCsr& p, HCsr& r, std::vector<int>& src) {
  // row blocks per thread: per-block column counts give every block its start
  // in each column, so the entries of a column keep ascending rows
  r.rows = p.cols;
  r.cols = p.rows;
  r.ptr.assign((size_t)p.cols + 1, 0);
  r.col.resize((size_t)p.nnz());
  src.resize((size_t)p.nnz());
  const int nt = std::max(1, std::min(omp_get_max_threads(),
                                      (int)std::min<long long>(64, (1LL << 26) / std::max(1, p.cols))));
  std::vector<int> cnt((size_t)nt * (size_t)p.cols, 0);
  auto block = [&](int t) { return (int)((long long)p.rows * t / nt); };
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; ++t) {
    int* ct = cnt.data() + (size_t)t * p.cols;
    for (int k = p.ptr[(size_t)block(t)]; k < p.ptr[(size_t)block(t + 1)]; ++k) ++ct[p.col[(size_t)k]];
  }
  // r.ptr, then each block's start in each column
  for (int j = 0; j < p.cols; ++j) {
    int run = r.ptr[(size_t)j];
    for (int t = 0; t < nt; ++t) {
      const int v = cnt[(size_t)t * p.cols + j];
      cnt[(size_t)t * p.cols + j] = run;
      run += v;
    }
    r.ptr[(size_t)j + 1] = run;
  }
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; ++t) {
    int* next = cnt.data() + (size_t)t * p.cols;
    for (int i = block(t); i < block(t + 1); ++i)
      for (int k = p.ptr[(size_t)i]; k < p.ptr[(size_t)i + 1]; ++k) {
        const int q = next[p.col[(size_t)k]]++;
        r.col[(size_t)q] = i;
        src[(size_t)q] = k;
      }
  }
}

// S = (C + C^T)/2 pattern; ia / ib: index in C of (i, j) / (j, i) or -1.
void symmetrize_pattern(const HCsr& c, HCsr& s, std::vector<int>& ia, std::vector<int>& ib) {
  HCsr t;
  std::vector<int> tsrc;
  transpose_pattern(c, t, tsrc);
  s.rows = c.rows;
  s.cols = c.cols;
  s.ptr.assign((size_t)c.rows + 1, 0);
  const int big = std::numeric_limits<int>::max();
  // the merge of row i of C and of C^T: counted, then written at the row's offset
  auto merge = [&](int i, bool fill) {
    int p = c.ptr[(size_t)i], q = t.ptr[(size_t)i], o = fill ? s.ptr[(size_t)i] : 0, n = 0;
    const int pe = c.ptr[(size_t)i + 1], qe = t.ptr[(size_t)i + 1];
    while (p < pe || q < qe) {
      const int cp = p < pe ? c.col[(size_t)p] : big;
      const int cq = q < qe ? t.col[(size_t)q] : big;
      int col, a = -1, b = -1;
      if (cp == cq) {
        col = cp;
        a = p++;
        b = tsrc[(size_t)q++];
      } else if (cp < cq) {
        col = cp;
        a = p++;
      } else {
        col = cq;
        b = tsrc[(size_t)q++];
      }
      if (fill) {
        s.col[(size_t)o + n] = col;
        ia[(size_t)o + n] = a;
        ib[(size_t)o + n] = b;
      }
      ++n;
    }
    return n;
  };
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < c.rows; ++i) s.ptr[(size_t)i + 1] = merge(i, false);
  for (int i = 0; i < c.rows; ++i) s.ptr[(size_t)i + 1] += s.ptr[(size_t)i];
  s.col.resize((size_t)s.ptr[(size_t)c.rows]);
  ia.resize(s.col.size());
  ib.resize(s.col.size());
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < c.rows; ++i) merge(i, true);
}

// ---------------------------------------------------------------------------
// Aggregation (Multigrid.cpp:457-565): block Frobenius strength, theta test,
// greedy passes in node order.
// ---------------------------------------------------------------------------
// The node blocks of a matrix for the strength test: per node, its neighbour
// nodes in ascending order with the sum of squares of the block's entries
// (rows in order, columns in order within a row), and the diagonal block's sum.
struct NodeBlocks {
  std::vector<int> ptr, node;
  std::vector<double> sumsq, diag;
};

NodeBlocks host_node_blocks(const HCsr& a, const std::vector<int>& node_ptr) {
  const int nn = (int)node_ptr.size() - 1;
  std::vector<int> node_of((size_t)a.rows);
  for (int nd = 0; nd < nn; ++nd)
    for (int i = node_ptr[(size_t)nd]; i < node_ptr[(size_t)nd + 1]; ++i) node_of[(size_t)i] = nd;
  NodeBlocks b;
  b.diag.assign((size_t)nn, 0.0);
  std::vector<std::vector<std::pair<int, double>>> per((size_t)nn);
#pragma omp parallel
  {
    std::vector<int> marker((size_t)nn, -1), touched;
    std::vector<double> acc((size_t)nn, 0.0);
#pragma omp for schedule(dynamic, 1024)
    for (int nd = 0; nd < nn; ++nd) {
      touched.clear();
      double sd = 0.0;
      for (int i = node_ptr[(size_t)nd]; i < node_ptr[(size_t)nd + 1]; ++i)
        for (int k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
          const int o = node_of[(size_t)a.col[(size_t)k]];
          const double v = a.val[(size_t)k];
          if (o == nd) {
            sd += v * v;
            continue;
          }
          if (marker[(size_t)o] != nd) {
            marker[(size_t)o] = nd;
            acc[(size_t)o] = 0.0;
            touched.push_back(o);
          }
          acc[(size_t)o] += v * v;
        }
      std::sort(touched.begin(), touched.end());
      b.diag[(size_t)nd] = sd;
      auto& out = per[(size_t)nd];
      out.reserve(touched.size());
      for (int o : touched) out.emplace_back(o, acc[(size_t)o]);
    }
  }
  b.ptr.assign((size_t)nn + 1, 0);
  for (int nd = 0; nd < nn; ++nd) b.ptr[(size_t)nd + 1] = b.ptr[(size_t)nd] + (int)per[(size_t)nd].size();
  b.node.resize((size_t)b.ptr[(size_t)nn]);
  b.sumsq.resize(b.node.size());
#pragma omp parallel for schedule(static)
  for (int nd = 0; nd < nn; ++nd) {
    int q = b.ptr[(size_t)nd];
    for (auto& pr : per[(size_t)nd]) {
      b.node[(size_t)q] = pr.first;
      b.sumsq[(size_t)q++] = pr.second;
    }
  }
  return b;
}

// Aggregation (Multigrid.cpp:457-565) from the node blocks: block Frobenius
// strength against the diagonal blocks' norms, theta test, greedy passes in
// node order.
void aggregate_blocks(const NodeBlocks& b, double theta, std::vector<int>& agg) {
  const int nn = (int)b.diag.size();
  std::vector<double> bdiag((size_t)nn);
  for (int nd = 0; nd < nn; ++nd) bdiag[(size_t)nd] = std::sqrt(b.diag[(size_t)nd]);
  // strong neighbours per node, ascending, with their relative strength
  std::vector<std::vector<std::pair<int, double>>> strong((size_t)nn);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int nd = 0; nd < nn; ++nd) {
    const double di = bdiag[(size_t)nd];
    auto& out = strong[(size_t)nd];
    for (int q = b.ptr[(size_t)nd]; q < b.ptr[(size_t)nd + 1]; ++q) {
      const int o = b.node[(size_t)q];
      const double s = std::sqrt(b.sumsq[(size_t)q]);
      const double ref = std::sqrt(di * bdiag[(size_t)o]);
      if (s > 0.0 && s >= theta * ref) out.emplace_back(o, ref > 0.0 ? s / ref : 0.0);
    }
  }
  agg.assign((size_t)nn, -1);
  int count = 0;
  for (int nd = 0; nd < nn; ++nd) {  // pass 1: whole free neighbourhoods
    if (agg[(size_t)nd] >= 0 || strong[(size_t)nd].empty()) continue;
    bool free_nb = true;
    for (auto& pr : strong[(size_t)nd])
      if (agg[(size_t)pr.first] >= 0) { free_nb = false; break; }
    if (!free_nb) continue;
    agg[(size_t)nd] = count;
    for (auto& pr : strong[(size_t)nd]) agg[(size_t)pr.first] = count;
    ++count;
  }
  const std::vector<int> first = agg;
  for (int nd = 0; nd < nn; ++nd) {  // pass 2: join the strongest pass-1 aggregate
    if (first[(size_t)nd] >= 0) continue;
    int best = -1;
    double best_v = -1.0;
    for (auto& pr : strong[(size_t)nd]) {
      const int target = first[(size_t)pr.first];
      if (target >= 0 && pr.second > best_v) { best = target; best_v = pr.second; }
    }
    if (best >= 0) agg[(size_t)nd] = best;
  }
  for (int nd = 0; nd < nn; ++nd) {  // pass 3: new aggregates of what is left
    if (agg[(size_t)nd] >= 0) continue;
    agg[(size_t)nd] = count;
    for (auto& pr : strong[(size_t)nd])
      if (agg[(size_t)pr.first] < 0) agg[(size_t)pr.first] = count;
    ++count;
  }
}

void aggregate_nodes(const HCsr& a, const std::vector<int>& node_ptr, double theta, std::vector<int>& agg) {
  aggregate_blocks(host_node_blocks(a, node_ptr), theta, agg);
}

// Sums in the order of the SSE2 kernels (two-lane packets) of Eigen 3.4.0,
// the library the reference's QR runs in: a vector reduction starts packet
// accumulators on lanes {0,1} and {2,3}, adds {i,i+1} and {i+2,i+3} per step of
// four, combines them, adds one more packet if two even entries remain, then
// the two lanes, then the odd last entry.
template <class F>
double lane_sum(int n, F&& x) {
  if (n <= 0) return 0.0;
  const int n2 = n & ~1;
  if (n2 == 0) return x(0);
  double a0 = x(0), a1 = x(1);
  if (n2 > 2) {
    double b0 = x(2), b1 = x(3);
    const int n4 = n & ~3;
    for (int i = 4; i < n4; i += 4) {
      a0 = a0 + x(i);
      a1 = a1 + x(i + 1);
      b0 = b0 + x(i + 2);
      b1 = b1 + x(i + 3);
    }
    a0 = a0 + b0;
    a1 = a1 + b1;
    if (n2 > n4) {
      a0 = a0 + x(n4);
      a1 = a1 + x(n4 + 1);
    }
  }
  double s = a0 + a1;
  for (int i = n2; i < n; ++i) s = s + x(i);
  return s;
}

// One output of Eigen's row-major matrix-vector kernel: even and odd products
// accumulate in the two lanes, then the lanes are added, then the odd last
// product.
double lane_gemv(const double* a, const double* b, int n) {
  double e = 0.0, o = 0.0;
  int j = 0;
  for (; j + 2 <= n; j += 2) {
    e = a[j] * b[j] + e;
    o = a[j + 1] * b[j + 1] + o;
  }
  double s = e + o;
  for (; j < n; ++j) s = s + a[j] * b[j];
  return 0.0 + s;
}

// H = I - tau [1; v][1; v]^T on a rows x cols column-major block (leading
// dimension ld), as MatrixBase::applyHouseholderOnTheLeft forms it.
void reflect_left(double* blk, int ld, int rows, int cols, const double* v, double tau, std::vector<double>& w,
                  std::vector<double>& tv) {
  if (rows == 1) {
    const double f = 1.0 - tau;
    for (int j = 0; j < cols; ++j) blk[(size_t)j * ld] *= f;
    return;
  }
  if (tau == 0.0 || cols == 0) return;
  const int nb = rows - 1;
  w.assign((size_t)cols, 0.0);
  if (cols == 1) {
    const double* b = blk + 1;
    w[0] = 0.0 + lane_sum(nb, [&](int i) { return v[i] * b[i]; });
  } else {
    for (int j = 0; j < cols; ++j) w[(size_t)j] = lane_gemv(blk + (size_t)j * ld + 1, v, nb);
  }
  for (int j = 0; j < cols; ++j) w[(size_t)j] = w[(size_t)j] + blk[(size_t)j * ld];
  for (int j = 0; j < cols; ++j) blk[(size_t)j * ld] = blk[(size_t)j * ld] - tau * w[(size_t)j];
  tv.resize((size_t)nb);
  for (int i = 0; i < nb; ++i) tv[(size_t)i] = tau * v[i];
  for (int j = 0; j < cols; ++j) {
    double* c = blk + (size_t)j * ld + 1;
    for (int i = 0; i < nb; ++i) c[i] = c[i] - w[(size_t)j] * tv[(size_t)i];
  }
}

// Column-pivoted Householder QR of one aggregate's near-null-space rows (a:
// rows x m, row-major), as Eigen::ColPivHouseholderQR (Multigrid.cpp:606-617):
// Businger-Golub pivoting on the updated column norms (first largest wins),
// LAPACK xGEQP3 norm downdating with direct recomputation below sqrt(eps),
// reflectors beta = -sign(x0) sqrt(x0^2 + |tail|^2), v = tail / (x0 - beta),
// tau = (beta - x0) / beta (none when |tail|^2 <= DBL_MIN), Q = the reflectors
// applied last to first to I(rows, rank); rank at 1e-10 |R00|. Every sum runs
// in the packet order above, so P-hat carries the reference's bits.
int qr_colpiv(std::vector<double>& a, int rows, int m, std::vector<double>& q, std::vector<double>& r) {
  const int size = std::min(rows, m);
  std::vector<double> c((size_t)rows * m);  // column-major working copy
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < m; ++j) c[(size_t)j * rows + i] = a[(size_t)i * m + j];
  auto at = [&](int i, int j) -> double& { return c[(size_t)j * rows + i]; };
  auto norm_from = [&](int j, int from) {
    const double* col = c.data() + (size_t)j * rows + from;
    return std::sqrt(lane_sum(rows - from, [&](int i) { return col[i] * col[i]; }));
  };
  std::vector<int> perm((size_t)m);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<double> tau((size_t)size, 0.0), nu((size_t)m), nd((size_t)m), w, tv;
  for (int j = 0; j < m; ++j) nu[(size_t)j] = nd[(size_t)j] = norm_from(j, 0);
  const double thr = std::sqrt(std::numeric_limits<double>::epsilon());
  for (int k = 0; k < size; ++k) {
    int piv = k;
    for (int j = k + 1; j < m; ++j)
      if (nu[(size_t)j] > nu[(size_t)piv]) piv = j;
    if (piv != k) {
      for (int i = 0; i < rows; ++i) std::swap(at(i, k), at(i, piv));
      std::swap(nu[(size_t)k], nu[(size_t)piv]);
      std::swap(nd[(size_t)k], nd[(size_t)piv]);
      std::swap(perm[(size_t)k], perm[(size_t)piv]);
    }
    double* col = c.data() + (size_t)k * rows;
    const double x0 = col[k];
    const int nt = rows - k - 1;
    const double tail = nt > 0 ? lane_sum(nt, [&](int i) { return col[k + 1 + i] * col[k + 1 + i]; }) : 0.0;
    double beta, t;
    if (tail <= std::numeric_limits<double>::min()) {
      t = 0.0;
      beta = x0;
      for (int i = k + 1; i < rows; ++i) col[i] = 0.0;
    } else {
      beta = std::sqrt(x0 * x0 + tail);
      if (x0 >= 0.0) beta = -beta;
      const double den = x0 - beta;
      for (int i = k + 1; i < rows; ++i) col[i] = col[i] / den;
      t = (beta - x0) / beta;
    }
    tau[(size_t)k] = t;
    col[k] = beta;
    reflect_left(c.data() + (size_t)(k + 1) * rows + k, rows, rows - k, m - k - 1, col + k + 1, t, w, tv);
    for (int j = k + 1; j < m; ++j) {
      if (nu[(size_t)j] == 0.0) continue;
      double g = std::abs(at(k, j)) / nu[(size_t)j];
      g = (1.0 + g) * (1.0 - g);
      if (g < 0.0) g = 0.0;
      const double ratio = nu[(size_t)j] / nd[(size_t)j];
      if (g * (ratio * ratio) <= thr) {
        nd[(size_t)j] = norm_from(j, k + 1);
        nu[(size_t)j] = nd[(size_t)j];
      } else {
        nu[(size_t)j] *= std::sqrt(g);
      }
    }
  }
  const double lead = size > 0 ? std::abs(at(0, 0)) : 0.0;
  int rank = 0;
  for (int i = 0; i < size; ++i)
    if (std::abs(at(i, i)) > 1.0e-10 * lead) ++rank;
  rank = std::max(rank, 1);
  std::vector<double> qc((size_t)rows * rank, 0.0);
  for (int j = 0; j < rank; ++j) qc[(size_t)j * rows + j] = 1.0;
  for (int k = size - 1; k >= 0; --k)
    reflect_left(qc.data() + k, rows, rows - k, rank, c.data() + (size_t)k * rows + k + 1, tau[(size_t)k], w, tv);
  q.assign((size_t)rows * rank, 0.0);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < rank; ++j) q[(size_t)i * rank + j] = qc[(size_t)j * rows + i];
  r.assign((size_t)rank * m, 0.0);
  for (int i = 0; i < rank; ++i)
    for (int j = i; j < m; ++j) r[(size_t)i * m + perm[(size_t)j]] = at(i, j);
  return rank;
}

// Tentative prolongator and the next level's node blocks / near-null space
// (Multigrid.cpp:567-669).
void tentative(Level& lvl, Level& next, int m, HCsr& pt) {
  const int n = lvl.n, nn = lvl.num_nodes;
  const std::vector<int>& agg = lvl.aggregate;
  const int na = agg.empty() ? 0 : *std::max_element(agg.begin(), agg.end()) + 1;
  std::vector<int> aptr((size_t)na + 1, 0), anodes((size_t)nn);
  for (int nd = 0; nd < nn; ++nd) ++aptr[(size_t)agg[(size_t)nd] + 1];
  for (int a = 0; a < na; ++a) aptr[(size_t)a + 1] += aptr[(size_t)a];
  {
    std::vector<int> fill(aptr.begin(), aptr.end() - 1);
    for (int nd = 0; nd < nn; ++nd) anodes[(size_t)fill[(size_t)agg[(size_t)nd]]++] = nd;
  }
  std::vector<std::vector<double>> qs((size_t)na), rs((size_t)na);
  std::vector<int> rank((size_t)na);
#pragma omp parallel for schedule(dynamic, 64)
  for (int a = 0; a < na; ++a) {
    int rows = 0;
    for (int k = aptr[(size_t)a]; k < aptr[(size_t)a + 1]; ++k) {
      const int nd = anodes[(size_t)k];
      rows += lvl.node_ptr[(size_t)nd + 1] - lvl.node_ptr[(size_t)nd];
    }
    std::vector<double> block((size_t)rows * m);
    int row = 0;
    for (int k = aptr[(size_t)a]; k < aptr[(size_t)a + 1]; ++k) {
      const int nd = anodes[(size_t)k];
      for (int i = lvl.node_ptr[(size_t)nd]; i < lvl.node_ptr[(size_t)nd + 1]; ++i, ++row)
        std::memcpy(&block[(size_t)row * m], &lvl.nullspace[(size_t)i * m], sizeof(double) * m);
    }
    rank[(size_t)a] = qr_colpiv(block, rows, m, qs[(size_t)a], rs[(size_t)a]);
  }
  next.node_ptr.assign((size_t)na + 1, 0);
  for (int a = 0; a < na; ++a) next.node_ptr[(size_t)a + 1] = next.node_ptr[(size_t)a] + rank[(size_t)a];
  next.num_nodes = na;
  next.n = next.node_ptr.back();
  next.nullspace.assign((size_t)next.n * m, 0.0);
  for (int a = 0; a < na; ++a)
    std::copy(rs[(size_t)a].begin(), rs[(size_t)a].end(), next.nullspace.begin() + (size_t)next.node_ptr[(size_t)a] * m);
  std::vector<int> agg_of_row((size_t)n), local((size_t)n);
  for (int a = 0; a < na; ++a) {
    int loc = 0;
    for (int k = aptr[(size_t)a]; k < aptr[(size_t)a + 1]; ++k) {
      const int nd = anodes[(size_t)k];
      for (int i = lvl.node_ptr[(size_t)nd]; i < lvl.node_ptr[(size_t)nd + 1]; ++i) {
        agg_of_row[(size_t)i] = a;
        local[(size_t)i] = loc++;
      }
    }
  }
  pt.rows = n;
  pt.cols = next.n;
  pt.ptr.assign((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) pt.ptr[(size_t)i + 1] = pt.ptr[(size_t)i] + rank[(size_t)agg_of_row[(size_t)i]];
  pt.col.resize((size_t)pt.nnz());
  pt.val.resize((size_t)pt.nnz());
  for (int i = 0; i < n; ++i) {
    const int a = agg_of_row[(size_t)i], ra = rank[(size_t)a];
    for (int j = 0; j < ra; ++j) {
      pt.col[(size_t)pt.ptr[(size_t)i] + j] = next.node_ptr[(size_t)a] + j;
      pt.val[(size_t)pt.ptr[(size_t)i] + j] = qs[(size_t)a][(size_t)local[(size_t)i] * ra + j];
    }
  }
}

// Largest eigenvalue of the symmetric tridiagonal Lanczos matrix, as the
// reference takes it from Eigen 3.4.0's SelfAdjointEigenSolver::
// computeFromTridiagonal (Multigrid.cpp:324-326): implicit symmetric QR steps
// with the Wilkinson shift (Golub & Van Loan, Algorithm 8.3.2) on the trailing
// unreduced block, an off-diagonal entry deflated when it is below DBL_MIN or
// when (e_i / eps)^2 <= |d_i| + |d_(i+1)|, at most 30 k sweeps.
double tridiag_max(const std::vector<double>& diag, const std::vector<double>& sub) {
  const int k = (int)diag.size();
  if (k == 1) return diag[0];
  std::vector<double> d(diag), e(sub.begin(), sub.begin() + (k - 1));
  auto rotation = [](double p, double q, double& c, double& s) {  // [c s; -s c]^T [p; q] = [r; 0]
    if (q == 0.0) {
      c = p < 0.0 ? -1.0 : 1.0;
      s = 0.0;
    } else if (p == 0.0) {
      c = 0.0;
      s = q < 0.0 ? 1.0 : -1.0;
    } else if (std::abs(p) > std::abs(q)) {
      const double t = q / p;
      double u = std::sqrt(1.0 + t * t);
      if (p < 0.0) u = -u;
      c = 1.0 / u;
      s = -t * c;
    } else {
      const double t = p / q;
      double u = std::sqrt(1.0 + t * t);
      if (q < 0.0) u = -u;
      s = -1.0 / u;
      c = -t * s;
    }
  };
  auto qr_step = [&](int start, int end) {
    const double td = (d[(size_t)end - 1] - d[(size_t)end]) * 0.5;
    const double f = e[(size_t)end - 1];
    double mu = d[(size_t)end];
    if (td == 0.0) {
      mu -= std::abs(f);
    } else if (f != 0.0) {
      const double hi = std::max(std::abs(td), std::abs(f)), lo = std::min(std::abs(f), std::abs(td));
      const double h = hi == 0.0 ? 0.0 : hi * std::sqrt(1.0 + (lo / hi) * (lo / hi));
      if (f * f == 0.0)
        mu -= f / ((td + (td > 0.0 ? h : -h)) / f);
      else
        mu -= (f * f) / (td + (td > 0.0 ? h : -h));
    }
    double x = d[(size_t)start] - mu, z = e[(size_t)start];
    for (int i = start; i < end && z != 0.0; ++i) {
      double c, s;
      rotation(x, z, c, s);
      const double sdk = s * d[(size_t)i] + c * e[(size_t)i];
      const double dkp1 = s * e[(size_t)i] + c * d[(size_t)i + 1];
      d[(size_t)i] = c * (c * d[(size_t)i] - s * e[(size_t)i]) - s * (c * e[(size_t)i] - s * d[(size_t)i + 1]);
      d[(size_t)i + 1] = s * sdk + c * dkp1;
      e[(size_t)i] = c * sdk - s * dkp1;
      if (i > start) e[(size_t)i - 1] = c * e[(size_t)i - 1] - s * z;
      x = e[(size_t)i];
      if (i < end - 1) {
        z = -s * e[(size_t)i + 1];
        e[(size_t)i + 1] = c * e[(size_t)i + 1];
      }
    }
  };
  const double tiny = std::numeric_limits<double>::min();
  const double inv_eps = 1.0 / std::numeric_limits<double>::epsilon();
  int end = k - 1, start = 0;
  long long sweeps = 0;
  while (end > 0) {
    for (int i = start; i < end; ++i) {
      if (std::abs(e[(size_t)i]) < tiny) {
        e[(size_t)i] = 0.0;
      } else {
        const double se = inv_eps * e[(size_t)i];
        if (se * se <= std::abs(d[(size_t)i]) + std::abs(d[(size_t)i + 1])) e[(size_t)i] = 0.0;
      }
    }
    while (end > 0 && e[(size_t)end - 1] == 0.0) --end;
    if (end <= 0) break;
    if (++sweeps > 30LL * k) break;
    start = end - 1;
    while (start > 0 && e[(size_t)start - 1] != 0.0) --start;
    qr_step(start, end);
  }
  return *std::max_element(d.begin(), d.end());
}

// P-hat's aggregate blocks for the transfer kernels (k::TentBlocks): every row
// holds one run of at most 6 consecutive columns, the runs of an aggregate's
// rows are the same, and the aggregates' runs tile the coarse columns. Leaves
// tent_ok false (the transfers then use the CSR products) when they do not.
void build_tent_blocks(Level& lv, const HCsr& pt, hipStream_t s) {
  lv.tent_ok = false;
  const int n = pt.rows;
  std::vector<int> c0len((size_t)n, 0);
  std::vector<double> ell((size_t)6 * n, 0.0);
  std::vector<int> start_len((size_t)pt.cols, -1);  // run length at an aggregate's first column
  for (int r = 0; r < n; ++r) {
    const int k0 = pt.ptr[(size_t)r], len = pt.ptr[(size_t)r + 1] - k0;
    if (len == 0) continue;
    if (len > 6) return;
    const int c0 = pt.col[(size_t)k0];
    for (int k = 0; k < len; ++k) {
      if (pt.col[(size_t)k0 + k] != c0 + k) return;
      ell[(size_t)k * n + r] = pt.val[(size_t)k0 + k];
    }
    if (start_len[(size_t)c0] >= 0 && start_len[(size_t)c0] != len) return;
    start_len[(size_t)c0] = len;
    c0len[(size_t)r] = c0 * 8 + len;
  }
  std::vector<int> agg_of((size_t)pt.cols, -1), agg_c0len;
  for (int col = 0; col < pt.cols;) {  // the runs must tile [0, cols)
    const int len = start_len[(size_t)col];
    if (len <= 0) return;
    agg_of[(size_t)col] = (int)agg_c0len.size();
    agg_c0len.push_back(col * 8 + len);
    col += len;
  }
  const int na = (int)agg_c0len.size();
  std::vector<int> aptr((size_t)na + 1, 0), rows;
  for (int r = 0; r < n; ++r)
    if (c0len[(size_t)r]) ++aptr[(size_t)agg_of[(size_t)(c0len[(size_t)r] >> 3)] + 1];
  for (int a = 0; a < na; ++a) aptr[(size_t)a + 1] += aptr[(size_t)a];
  rows.resize((size_t)aptr[(size_t)na]);
  {
    std::vector<int> fill(aptr.begin(), aptr.end() - 1);
    for (int r = 0; r < n; ++r)
      if (c0len[(size_t)r]) rows[(size_t)fill[(size_t)agg_of[(size_t)(c0len[(size_t)r] >> 3)]]++] = r;
  }
  std::vector<double> ella((size_t)6 * n, 0.0);
  for (size_t i = 0; i < rows.size(); ++i)
    for (int k = 0; k < 6; ++k) ella[(size_t)k * n + i] = ell[(size_t)k * n + (size_t)rows[i]];
  lv.tent_c0len.upload(c0len, s);
  lv.tent_ell.upload(ell, s);
  lv.tent_ell32.upload(std::vector<float>(ell.begin(), ell.end()), s);
  lv.tent_ella.upload(ella, s);
  lv.tent_ella32.upload(std::vector<float>(ella.begin(), ella.end()), s);
  lv.tent_agg_ptr.upload(aptr, s);
  lv.tent_agg_rows.upload(rows, s);
  lv.tent_agg_c0len.upload(agg_c0len, s);
  lv.tent_naggs = na;
  lv.tent_ok = true;
}

void build_finest_nodes(Context& c, Level& lvl) {
  const int n = lvl.n, dim = c.layout_dim, ndpn = c.layout_ndpn;
  const std::vector<int>& u = c.layout_unknowns;
  lvl.node_ptr.assign(1, 0);
  int prev = -1;
  for (int i = 0; i < n; ++i) {
    const int g = u[(size_t)i], node = g / ndpn;
    if (g < 0 || node >= c.layout_nnodes) fail(SLX_SOLVER_ERROR, "multigrid: an unknown maps outside the model's nodes");
    if (i > 0 && g <= u[(size_t)i - 1]) fail(SLX_SOLVER_ERROR, "multigrid: the DOF layout must list unknowns in ascending order");
    if (node != prev) {
      if (i > 0) lvl.node_ptr.push_back(i);
      prev = node;
    }
  }
  lvl.node_ptr.push_back(n);
  lvl.num_nodes = (int)lvl.node_ptr.size() - 1;
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = std::numeric_limits<double>::max();
    hi[d] = -std::numeric_limits<double>::max();
  }
  for (int nd = 0; nd < c.layout_nnodes; ++nd)
    for (int d = 0; d < 3; ++d) {
      const double x = d < dim ? c.layout_coords[(size_t)nd * dim + d] : 0.0;
      lo[d] = std::min(lo[d], x);
      hi[d] = std::max(hi[d], x);
    }
  double ctr[3], scale = 0.0;
  for (int d = 0; d < 3; ++d) {
    ctr[d] = 0.5 * (lo[d] + hi[d]);
    scale = std::max(scale, hi[d] - lo[d]);
  }
  scale = std::max(scale, 1.0e-300);
  const int m = dim == 2 ? 3 : 6;
  c.nullspace_dim = m;
  lvl.nullspace.assign((size_t)n * m, 0.0);
  for (int i = 0; i < n; ++i) {
    const int g = u[(size_t)i], node = g / ndpn, k = g % ndpn;
    double r[3];
    for (int d = 0; d < 3; ++d) r[d] = ((d < dim ? c.layout_coords[(size_t)node * dim + d] : 0.0) - ctr[d]) / scale;
    double* row = &lvl.nullspace[(size_t)i * m];
    if (k >= 3) {
      row[k] = 1.0 / scale;
      continue;
    }
    row[k] = 1.0;
    if (dim == 2) {
      row[2] = k == 0 ? -r[1] : r[0];
    } else {
      const double rx[3] = {0.0, -r[2], r[1]}, ry[3] = {r[2], 0.0, -r[0]}, rz[3] = {-r[1], r[0], 0.0};
      row[3] = rx[k];
      row[4] = ry[k];
      row[5] = rz[k];
    }
  }
}

// ---------------------------------------------------------------------------
// Node-row product metadata
// ---------------------------------------------------------------------------
std::vector<int> node_of_rows(const std::vector<int>& node_ptr) {
  std::vector<int> of((size_t)node_ptr.back());
  for (size_t nd = 0; nd + 1 < node_ptr.size(); ++nd)
    for (int i = node_ptr[nd]; i < node_ptr[nd + 1]; ++i) of[(size_t)i] = (int)nd;
  return of;
}

// all rows of each node share one column pattern
bool node_uniform(const HCsr& m, const std::vector<int>& node_ptr) {
  bool ok = true;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int nd = 0; nd < (int)node_ptr.size() - 1; ++nd) {
    const int r0 = node_ptr[(size_t)nd];
    const int l0 = m.ptr[(size_t)r0 + 1] - m.ptr[(size_t)r0];
    for (int r = r0 + 1; r < node_ptr[(size_t)nd + 1] && ok; ++r) {
      if (m.ptr[(size_t)r + 1] - m.ptr[(size_t)r] != l0) { ok = false; break; }
      if (!std::equal(m.col.begin() + m.ptr[(size_t)r], m.col.begin() + m.ptr[(size_t)r + 1], m.col.begin() + m.ptr[(size_t)r0]))
        ok = false;
    }
  }
  return ok;
}

// One column list per node for the V-cycle's products when the rows of every
// node share their pattern (node_ptr covers all rows); otherwise none.
void share_columns(DCsr& a, const HCsr& m, const std::vector<int>& node_ptr, hipStream_t s) {
  a.cbase.release();
  a.ncol.release();
  if (node_ptr.size() < 2 || node_ptr.back() != m.rows || (int)node_ptr.size() - 1 == m.rows ||
      !node_uniform(m, node_ptr))
    return;
  std::vector<int> cbase((size_t)m.rows), ncol;
  ncol.reserve((size_t)m.nnz() / 2 + 1);
  for (size_t nd = 0; nd + 1 < node_ptr.size(); ++nd) {
    const int r0 = node_ptr[nd];
    const int base = (int)ncol.size();
    ncol.insert(ncol.end(), m.col.begin() + m.ptr[(size_t)r0], m.col.begin() + m.ptr[(size_t)r0 + 1]);
    for (int r = r0; r < node_ptr[nd + 1]; ++r) cbase[(size_t)r] = base - m.ptr[(size_t)r];
  }
  a.cbase.upload(cbase, s);
  a.ncol.upload(ncol, s);
}

// the columns of every row come in complete runs of the column nodes
bool runs_complete(const HCsr& m, const std::vector<int>& col_node_ptr, const std::vector<int>& col_node_of) {
  bool ok = true;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int i = 0; i < m.rows; ++i) {
    int k = m.ptr[(size_t)i];
    while (k < m.ptr[(size_t)i + 1] && ok) {
      const int j = m.col[(size_t)k];
      const int J = col_node_of[(size_t)j];
      const int j0 = col_node_ptr[(size_t)J], dj = col_node_ptr[(size_t)J + 1] - j0;
      if (j != j0 || k + dj > m.ptr[(size_t)i + 1]) { ok = false; break; }
      for (int q = 0; q < dj; ++q)
        if (m.col[(size_t)k + q] != j0 + q) { ok = false; break; }
      k += dj;
    }
  }
  return ok;
}

// per row node: the aggregate blocks of its (shared) row pattern, with the
// position of each block in the row; false if a block is not a complete
// aggregate column range
bool block_lists(const HCsr& m, const std::vector<int>& node_ptr, const std::vector<int>& agg_of_col,
                 const std::vector<int>& agg_off, std::vector<int>& bptr, std::vector<int>& bagg,
                 std::vector<int>& bpos) {
  const int nn = (int)node_ptr.size() - 1;
  bptr.assign((size_t)nn + 1, 0);
  std::atomic<bool> ok{true};
  // node nd's blocks: counted (and checked), then written at the node's offset
  auto walk = [&](int nd, bool fill) {
    const int r0 = node_ptr[(size_t)nd];
    int k = m.ptr[(size_t)r0], n = 0, o = fill ? bptr[(size_t)nd] : 0;
    const int k1 = m.ptr[(size_t)r0 + 1];
    while (k < k1) {
      const int c = m.col[(size_t)k];
      const int a = agg_of_col[(size_t)c];
      const int w = agg_off[(size_t)a + 1] - agg_off[(size_t)a];
      if (c != agg_off[(size_t)a] || k + w > k1) return -1;
      for (int q = 0; q < w; ++q)
        if (m.col[(size_t)k + q] != c + q) return -1;
      if (fill) {
        bagg[(size_t)o + n] = a;
        bpos[(size_t)o + n] = k - m.ptr[(size_t)r0];
      }
      ++n;
      k += w;
    }
    return n;
  };
#pragma omp parallel for schedule(dynamic, 1024)
  for (int nd = 0; nd < nn; ++nd) {
    const int n = walk(nd, false);
    if (n < 0) ok.store(false, std::memory_order_relaxed);
    bptr[(size_t)nd + 1] = std::max(n, 0);
  }
  if (!ok.load()) return false;
  for (int nd = 0; nd < nn; ++nd) bptr[(size_t)nd + 1] += bptr[(size_t)nd];
  bagg.resize((size_t)bptr[(size_t)nn]);
  bpos.resize(bagg.size());
#pragma omp parallel for schedule(dynamic, 1024)
  for (int nd = 0; nd < nn; ++nd) walk(nd, true);
  return true;
}

// lanes x slots needed by the widest node-row of c
// runs of each node-row of m (its first row stands for all): per run the
// offset in the row, the first B row, the number of B rows and the B node
bool run_table(const HCsr& m, const std::vector<int>& node_ptr, const std::vector<int>& col_node_ptr,
               const std::vector<int>& col_node_of, std::vector<int>& run_ptr, std::vector<int>& runs) {
  const int nn = (int)node_ptr.size() - 1;
  run_ptr.assign((size_t)nn + 1, 0);
  std::atomic<bool> ok{true};
  // node nd's runs: counted (and checked), then written at the node's offset
  auto walk = [&](int nd, bool fill) {
    const int r0 = node_ptr[(size_t)nd];
    const int k0 = m.ptr[(size_t)r0], k1 = m.ptr[(size_t)r0 + 1];
    int k = k0, n = 0;
    int* out = fill ? runs.data() + (size_t)4 * run_ptr[(size_t)nd] : nullptr;
    while (k < k1) {
      const int j = m.col[(size_t)k];
      const int J = col_node_of[(size_t)j];
      const int dj = col_node_ptr[(size_t)J + 1] - col_node_ptr[(size_t)J];
      if (dj > 6) return -1;
      if (fill) {
        out[4 * n] = k - k0;
        out[4 * n + 1] = j;
        out[4 * n + 2] = dj;
        out[4 * n + 3] = J;
      }
      ++n;
      k += dj;
    }
    return n;
  };
#pragma omp parallel for schedule(dynamic, 1024)
  for (int nd = 0; nd < nn; ++nd) {
    const int n = walk(nd, false);
    if (n < 0) ok.store(false, std::memory_order_relaxed);
    run_ptr[(size_t)nd + 1] = std::max(n, 0);
  }
  if (!ok.load()) return false;
  for (int nd = 0; nd < nn; ++nd) run_ptr[(size_t)nd + 1] += run_ptr[(size_t)nd];
  runs.resize((size_t)4 * run_ptr[(size_t)nn]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int nd = 0; nd < nn; ++nd) walk(nd, true);
  return true;
}

// Contribution lists of C = A B: the items are C's blocks (block lists cbp /
// cba / cbq over the C nodes); item q of node I lists, in the order of A's row,
// every run t of I's A row whose B node J holds q's aggregate, with that
// block's position in J's rows. False if a list entry does not fit 16 + 16 bits
// or a product block is missing from C's pattern.
bool build_plan(const std::vector<int>& cnode_ptr, const std::vector<int>& cbp, const std::vector<int>& cba,
                const std::vector<int>& cbq, const std::vector<int>& run_ptr, const std::vector<int>& runs,
                const std::vector<int>& bbp, const std::vector<int>& bba, const std::vector<int>& bbq,
                const std::vector<int>& bptr, const std::vector<int>& agg_off, const std::vector<int>* order,
                ProductPlan& plan, hipStream_t s) {
  const int nn = (int)cnode_ptr.size() - 1;
  const int naggs = (int)agg_off.size() - 1;
  const long long nitems = cbp[(size_t)nn];
  std::vector<long long> begin((size_t)nitems + 1, 0);
  std::atomic<bool> ok{true};
  auto walk = [&](bool fill, std::vector<uint2>* list, std::vector<long long>* pos) {
#pragma omp parallel
    {
      std::vector<int> item_of((size_t)naggs, -1);
#pragma omp for schedule(dynamic, 256)
      for (int I = 0; I < nn; ++I) {
        for (int q = cbp[(size_t)I]; q < cbp[(size_t)I + 1]; ++q) item_of[(size_t)cba[(size_t)q]] = q;
        for (int t = run_ptr[(size_t)I]; t < run_ptr[(size_t)I + 1]; ++t) {
          const int koff = runs[(size_t)4 * t], brow = runs[(size_t)4 * t + 1], dj = runs[(size_t)4 * t + 2];
          const int J = runs[(size_t)4 * t + 3];
          const int lb = bptr[(size_t)brow + 1] - bptr[(size_t)brow];
          for (int u = bbp[(size_t)J]; u < bbp[(size_t)J + 1]; ++u) {
            const int q = item_of[(size_t)bba[(size_t)u]];
            if (q < 0 || koff > 0xffff || lb > 0x1fff || dj > 7) {
              ok.store(false, std::memory_order_relaxed);
              continue;
            }
            if (fill)
              (*list)[(size_t)(*pos)[(size_t)q]++] =
                  make_uint2((unsigned)(bptr[(size_t)brow] + bbq[(size_t)u]),
                             ((unsigned)koff << 16) | ((unsigned)lb << 3) | (unsigned)dj);
            else
              ++begin[(size_t)q + 1];
          }
        }
        for (int q = cbp[(size_t)I]; q < cbp[(size_t)I + 1]; ++q) item_of[(size_t)cba[(size_t)q]] = -1;
      }
    }
  };
  walk(false, nullptr, nullptr);
  if (!ok.load()) return false;
  for (long long q = 0; q < nitems; ++q) begin[(size_t)q + 1] += begin[(size_t)q];
  std::vector<uint2> list((size_t)begin[(size_t)nitems]);
  std::vector<long long> pos(begin.begin(), begin.end() - 1);
  walk(true, &list, &pos);
  std::vector<int> node((size_t)nitems);
  int dmax = 1, wmax = 1;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : dmax, wmax)
  for (int I = 0; I < nn; ++I) {
    dmax = std::max(dmax, cnode_ptr[(size_t)I + 1] - cnode_ptr[(size_t)I]);
    for (int q = cbp[(size_t)I]; q < cbp[(size_t)I + 1]; ++q) {
      node[(size_t)q] = I;
      const int g = cba[(size_t)q];
      wmax = std::max(wmax, agg_off[(size_t)g + 1] - agg_off[(size_t)g]);
    }
  }
  if (dmax > 6 || wmax > 64) return false;
  if (order) {
    // emit the items node by node in the given order (the lists move with them):
    // each node's first item and first list entry in the new order, then the moves
    if ((int)order->size() != nn) return false;
    std::vector<long long> item0((size_t)nn + 1, 0), entry0((size_t)nn + 1, 0);
    for (int r = 0; r < nn; ++r) {
      const int I = (*order)[(size_t)r];
      item0[(size_t)r + 1] = item0[(size_t)r] + (cbp[(size_t)I + 1] - cbp[(size_t)I]);
      entry0[(size_t)r + 1] = entry0[(size_t)r] + (begin[(size_t)cbp[(size_t)I + 1]] - begin[(size_t)cbp[(size_t)I]]);
    }
    if (item0[(size_t)nn] != nitems) return false;
    std::vector<int> n2((size_t)nitems), a2((size_t)nitems), p2((size_t)nitems);
    std::vector<long long> b2((size_t)nitems + 1, 0);
    std::vector<uint2> l2(list.size());
#pragma omp parallel for schedule(dynamic, 1024)
    for (int r = 0; r < nn; ++r) {
      const int I = (*order)[(size_t)r];
      long long q2 = item0[(size_t)r], e2 = entry0[(size_t)r];
      for (int q = cbp[(size_t)I]; q < cbp[(size_t)I + 1]; ++q, ++q2) {
        n2[(size_t)q2] = I;
        a2[(size_t)q2] = cba[(size_t)q];
        p2[(size_t)q2] = cbq[(size_t)q];
        for (long long e = begin[(size_t)q]; e < begin[(size_t)q + 1]; ++e) l2[(size_t)e2++] = list[(size_t)e];
        b2[(size_t)q2 + 1] = e2;
      }
    }
    plan.nitems = nitems;
    plan.dmax = dmax;
    plan.wmax = wmax;
    plan.item_node.upload(n2, s);
    plan.item_agg.upload(a2, s);
    plan.item_pos.upload(p2, s);
    plan.item_begin.upload(b2, s);
    plan.list.upload(l2, s);
    return true;
  }
  plan.nitems = nitems;
  plan.dmax = dmax;
  plan.wmax = wmax;
  plan.item_node.upload(node, s);
  plan.item_agg.upload(cba, s);
  plan.item_pos.upload(cbq, s);
  plan.item_begin.upload(begin, s);
  plan.list.upload(list, s);
  return true;
}

// Row plan of C = A B (one wavefront per C node): the runs of each C node's A
// row (run table: offset in A's row, first B row, B rows, B node) and, per run,
// one entry for each aggregate block of the B node's rows: its position in B's
// row, its position in C's row (the C block of the same aggregate) and its
// width. False if a field does not fit its bits, a block is missing from C's
// pattern or a row is too long for the kernel's LDS.
bool build_row_plan(const std::vector<int>& cnode_ptr, const std::vector<int>& cptr, const std::vector<int>& cbp,
                    const std::vector<int>& cba, const std::vector<int>& cbq, const std::vector<int>& run_ptr,
                    const std::vector<int>& runs, const std::vector<int>& bbp, const std::vector<int>& bba,
                    const std::vector<int>& bbq, const std::vector<int>& bptr, const std::vector<int>& agg_off,
                    const std::vector<int>* order, RowPlan& plan, hipStream_t s) {
  const int nn = (int)cnode_ptr.size() - 1;
  const int naggs = (int)agg_off.size() - 1;
  const long long nruns = run_ptr[(size_t)nn];
  std::vector<int> run_entry((size_t)nruns + 1, 0);
  long long total = 0;
  int dmax = 1, jmax = 1, wmax = 1, lenc_max = 1, emax = 0, lena_max = 1;
  for (int I = 0; I < nn; ++I) {
    const int r0 = cnode_ptr[(size_t)I];
    dmax = std::max(dmax, cnode_ptr[(size_t)I + 1] - r0);
    lenc_max = std::max(lenc_max, cptr[(size_t)r0 + 1] - cptr[(size_t)r0]);
    const long long e0 = total;
    for (int t = run_ptr[(size_t)I]; t < run_ptr[(size_t)I + 1]; ++t) {
      const int J = runs[(size_t)4 * t + 3];
      const int nb = bbp[(size_t)J + 1] - bbp[(size_t)J];
      if (nb < 1) return false;
      lena_max = std::max(lena_max, runs[(size_t)4 * t] + runs[(size_t)4 * t + 2]);
      jmax = std::max(jmax, runs[(size_t)4 * t + 2]);
      run_entry[(size_t)t] = (int)total;
      total += nb;
      if (total > 0x7fffffffLL) return false;
    }
    emax = std::max(emax, (int)(total - e0));
  }
  run_entry[(size_t)nruns] = (int)total;
  for (int g = 0; g < naggs; ++g) wmax = std::max(wmax, agg_off[(size_t)g + 1] - agg_off[(size_t)g]);
  if (dmax > 6 || jmax > 6 || wmax > 7 || lenc_max > 0xffff) return false;
  std::vector<uint4> run((size_t)nruns);
  std::vector<unsigned> entry((size_t)total);
  std::atomic<bool> ok{true};
#pragma omp parallel
  {
    std::vector<int> pos_of((size_t)naggs, -1);
#pragma omp for schedule(dynamic, 256)
    for (int I = 0; I < nn; ++I) {
      for (int q = cbp[(size_t)I]; q < cbp[(size_t)I + 1]; ++q) pos_of[(size_t)cba[(size_t)q]] = cbq[(size_t)q];
      for (int t = run_ptr[(size_t)I]; t < run_ptr[(size_t)I + 1]; ++t) {
        const int koff = runs[(size_t)4 * t], brow = runs[(size_t)4 * t + 1], dj = runs[(size_t)4 * t + 2];
        const int J = runs[(size_t)4 * t + 3];
        const int lb = bptr[(size_t)brow + 1] - bptr[(size_t)brow];
        if (koff > 0xffff || lb > 0x1fff || dj > 6 || dj < 1) {
          ok.store(false, std::memory_order_relaxed);
          continue;
        }
        run[(size_t)t] = make_uint4((unsigned)bptr[(size_t)brow], ((unsigned)koff << 16) | ((unsigned)lb << 3) | (unsigned)dj,
                                    (unsigned)run_entry[(size_t)t], (unsigned)run_entry[(size_t)t + 1]);
        int e = run_entry[(size_t)t];
        for (int u = bbp[(size_t)J]; u < bbp[(size_t)J + 1]; ++u, ++e) {
          const int g = bba[(size_t)u];
          const int cpos = pos_of[(size_t)g], bpos = bbq[(size_t)u];
          const int w = agg_off[(size_t)g + 1] - agg_off[(size_t)g];
          if (cpos < 0 || bpos > 0x1fff || cpos > 0xffff) {
            ok.store(false, std::memory_order_relaxed);
            continue;
          }
          entry[(size_t)e] = ((unsigned)cpos << 16) | ((unsigned)bpos << 3) | (unsigned)w;
        }
      }
      for (int q = cbp[(size_t)I]; q < cbp[(size_t)I + 1]; ++q) pos_of[(size_t)cba[(size_t)q]] = -1;
    }
  }
  if (!ok.load()) return false;
  // LDS per wavefront: accumulators (the kernel's row bound: 3 or 6) and, when
  // small enough, the node's entries
  const int dk = dmax <= 3 && jmax <= 3 ? 3 : 6;
  const size_t budget = 16 * 1024;
  const bool sa = dk == 3 && k::row_product_lds(dk, lenc_max, 0, lena_max) <= budget;
  const bool pre = k::row_product_lds(dk, lenc_max, emax, sa ? lena_max : 0) <= budget;
  const size_t wave_lds = k::row_product_lds(dk, lenc_max, pre ? emax : 0, sa ? lena_max : 0);
  int wpb = 4;
  while (wpb > 1 && wave_lds * (size_t)wpb > 64 * 1024) wpb /= 2;
  if (wave_lds * (size_t)wpb > 64 * 1024) return false;
  std::vector<int> ord((size_t)nn);
  if (order) {
    if ((int)order->size() != nn) return false;
    ord = *order;
  } else {
    for (int I = 0; I < nn; ++I) ord[(size_t)I] = I;
  }
  plan.nnodes = nn;
  plan.dmax = dmax;
  plan.jmax = jmax;
  plan.wmax = wmax;
  plan.lenc_max = lenc_max;
  plan.emax = emax;
  plan.wpb = wpb;
  plan.pre = pre;
  plan.sa = sa;
  plan.lena_max = lena_max;
  plan.runs = nruns;
  plan.entries = total;
  plan.order.upload(ord, s);
  plan.run_ptr.upload(run_ptr, s);
  plan.run.upload(run, s);
  plan.entry.upload(entry, s);
  return true;
}

// Nodes sorted along a Morton (Z-order) curve of their coordinates (n x 3):
// products visit nodes that share neighbours close together in time.
std::vector<int> morton_order(const std::vector<double>& xyz) {
  const size_t n = xyz.size() / 3;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (size_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], xyz[3 * i + (size_t)d]);
      hi[d] = std::max(hi[d], xyz[3 * i + (size_t)d]);
    }
  auto spread = [](unsigned long long v) {  // 21 bits -> every third of 63
    v &= 0x1fffffULL;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  std::vector<std::pair<unsigned long long, int>> key(n);
  for (size_t i = 0; i < n; ++i) {
    unsigned long long code = 0;
    for (int d = 0; d < 3; ++d) {
      const double span = hi[d] - lo[d];
      const double t = span > 0.0 ? (xyz[3 * i + (size_t)d] - lo[d]) / span : 0.0;
      code |= spread((unsigned long long)(t * 2097151.0)) << d;
    }
    key[i] = {code, (int)i};
  }
  std::stable_sort(key.begin(), key.end());
  std::vector<int> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = key[i].second;
  return order;
}

void prepare_node_products(Level& lv, const Level& next, const HCsr& a, const HCsr& pt, const HCsr& ppat,
                           const HCsr& rpat, const HCsr& appat, const HCsr& rawpat, const std::vector<double>* xyz,
                           hipStream_t s) {
  lv.node_ok = false;
  const std::vector<int> f_of = node_of_rows(lv.node_ptr);
  const std::vector<int> c_of = node_of_rows(next.node_ptr);
  const std::vector<int>& off = next.node_ptr;
  if (!node_uniform(a, lv.node_ptr) || !node_uniform(ppat, lv.node_ptr) || !node_uniform(appat, lv.node_ptr) ||
      !node_uniform(rpat, next.node_ptr) || !node_uniform(rawpat, next.node_ptr))
    return;
  if (!runs_complete(a, lv.node_ptr, f_of) || !runs_complete(rpat, lv.node_ptr, f_of)) return;
  setup_trace("  plans: pattern checks", -1, s);
  std::vector<int> tp, ta, tq, pp, pa, pq, qp, qa, qq, wp, wa, wq;
  if (!block_lists(pt, lv.node_ptr, c_of, off, tp, ta, tq) || !block_lists(ppat, lv.node_ptr, c_of, off, pp, pa, pq) ||
      !block_lists(appat, lv.node_ptr, c_of, off, qp, qa, qq) ||
      !block_lists(rawpat, next.node_ptr, c_of, off, wp, wa, wq))
    return;
  setup_trace("  plans: block lists", -1, s);
  std::vector<int> arp, ar, rrp, rr;
  if (!run_table(a, lv.node_ptr, lv.node_ptr, f_of, arp, ar) || !run_table(rpat, next.node_ptr, lv.node_ptr, f_of, rrp, rr))
    return;
  setup_trace("  plans: run tables", -1, s);
  for (size_t nd = 0; nd + 1 < tp.size(); ++nd)
    if (tp[nd + 1] - tp[nd] != 1) return;  // P-hat: one block per node
  // with node coordinates, fine nodes and aggregates (at their first node) in Morton order
  std::vector<int> f_order, c_order;
  if (xyz) {
    f_order = morton_order(*xyz);
    const int nagg = (int)next.node_ptr.size() - 1;
    std::vector<double> axyz((size_t)nagg * 3, 0.0);
    std::vector<char> seen((size_t)nagg, 0);
    for (int nd = 0; nd + 1 < (int)lv.node_ptr.size(); ++nd) {
      const int g = lv.aggregate[(size_t)nd];
      if (g < 0 || seen[(size_t)g]) continue;
      seen[(size_t)g] = 1;
      for (int d = 0; d < 3; ++d) axyz[(size_t)g * 3 + d] = (*xyz)[(size_t)nd * 3 + d];
    }
    c_order = morton_order(axyz);
  }
  const std::vector<int>* fo = xyz ? &f_order : nullptr;
  const std::vector<int>* co = xyz ? &c_order : nullptr;
  setup_trace("  plans: Morton orders", -1, s);
  if (!build_plan(lv.node_ptr, pp, pa, pq, arp, ar, tp, ta, tq, pt.ptr, off, fo, lv.plan_p, s)) return;
  setup_trace("  plans: P plan", -1, s);
  if (!build_row_plan(lv.node_ptr, appat.ptr, qp, qa, qq, arp, ar, pp, pa, pq, ppat.ptr, off, fo, lv.plan_ap, s))
    return;
  setup_trace("  plans: A P row plan", -1, s);
  if (!build_row_plan(next.node_ptr, rawpat.ptr, wp, wa, wq, rrp, rr, qp, qa, qq, appat.ptr, off, co, lv.plan_raw, s))
    return;
  setup_trace("  plans: R A P row plan", -1, s);
  std::vector<int> row_agg((size_t)lv.n);
  for (int i = 0; i < lv.n; ++i) row_agg[(size_t)i] = c_of[(size_t)pt.col[(size_t)pt.ptr[(size_t)i]]];
  lv.f_node_ptr.upload(lv.node_ptr, s);
  lv.c_node_ptr.upload(next.node_ptr, s);
  lv.row_agg.upload(row_agg, s);
  lv.node_ok = true;
}

k::BlockProductArgs plan_args(const Level& lv, const ProductPlan& plan, bool a_is_r, const DCsr& A, const DCsr& B,
                              DCsr& C) {
  k::BlockProductArgs g{};
  g.nitems = plan.nitems;
  g.dmax = plan.dmax;
  g.wmax = plan.wmax;
  g.item_node = plan.item_node.p;
  g.item_agg = plan.item_agg.p;
  g.item_pos = plan.item_pos.p;
  g.item_begin = plan.item_begin.p;
  g.list = plan.list.p;
  g.out_node_ptr = a_is_r ? lv.c_node_ptr.p : lv.f_node_ptr.p;
  g.aptr = A.ptr.p;
  g.aval = A.val.p;
  g.bval = B.val.p;
  g.cptr = C.ptr.p;
  g.cval = C.val.p;
  g.agg_off = lv.c_node_ptr.p;
  return g;
}

k::RowProductArgs row_args(const Level& lv, const RowPlan& plan, bool a_is_r, const DCsr& A, const DCsr& B, DCsr& C) {
  k::RowProductArgs g{};
  g.nnodes = plan.nnodes;
  g.dmax = plan.dmax;
  g.jmax = plan.jmax;
  g.wmax = plan.wmax;
  g.lenc_max = plan.lenc_max;
  g.emax = plan.emax;
  g.wpb = plan.wpb;
  g.pre = plan.pre;
  g.sa = plan.sa;
  g.lena_max = plan.lena_max;
  g.order = plan.order.p;
  g.node_ptr = a_is_r ? lv.c_node_ptr.p : lv.f_node_ptr.p;
  g.aptr = A.ptr.p;
  g.aval = A.val.p;
  g.run_ptr = plan.run_ptr.p;
  g.run = plan.run.p;
  g.entry = plan.entry.p;
  g.bval = B.val.p;
  g.cptr = C.ptr.p;
  g.cval = C.val.p;
  return g;
}

DCsr& level_op(Context& c, size_t l) { return l == 0 ? c.K : c.levels[l].A; }

HCsr download_csr(const DCsr& a, hipStream_t s) {
  HCsr h;
  h.rows = a.rows;
  h.cols = a.cols;
  h.ptr.resize((size_t)a.rows + 1);
  h.col.resize((size_t)a.nnz);
  h.val.resize((size_t)a.nnz);
  a.ptr.download(h.ptr.data(), h.ptr.size(), s);
  a.col.download(h.col.data(), h.col.size(), s);
  a.val.download(h.val.data(), h.val.size(), s);
  SLX_HIP(hipStreamSynchronize(s));
  return h;
}

// The level-0 node blocks from K's host pattern and its device values: the
// runs of each node's first row (one neighbour node each) are counted on the
// host, their sums of squares over the node's rows computed on the device in
// the host's order. False when the rows of a node do not share one pattern of
// complete node runs (the host path then reads K).
bool device_node_blocks(Context& c, const std::vector<int>& node_ptr, NodeBlocks& b) {
  const HCsr& a = c.k_host;
  const std::vector<int> node_of = node_of_rows(node_ptr);
  if (!node_uniform(a, node_ptr) || !runs_complete(a, node_ptr, node_of)) return false;
  const int nn = (int)node_ptr.size() - 1;
  std::vector<int> run_ptr((size_t)nn + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int nd = 0; nd < nn; ++nd) {
    const int r0 = node_ptr[(size_t)nd];
    int runs = 0;
    for (int k = a.ptr[(size_t)r0]; k < a.ptr[(size_t)r0 + 1];) {
      const int J = node_of[(size_t)a.col[(size_t)k]];
      k += node_ptr[(size_t)J + 1] - node_ptr[(size_t)J];
      ++runs;
    }
    run_ptr[(size_t)nd + 1] = runs;
  }
  for (int nd = 0; nd < nn; ++nd) run_ptr[(size_t)nd + 1] += run_ptr[(size_t)nd];
  hipStream_t s = c.stream;
  DBuf<int> d_node_ptr, d_node_of, d_run_ptr;
  DBuf<double> d_sums;
  d_node_ptr.upload(node_ptr, s);
  d_node_of.upload(node_of, s);
  d_run_ptr.upload(run_ptr, s);
  d_sums.alloc((size_t)std::max(1, run_ptr[(size_t)nn]));
  k::node_block_sums(nn, d_node_ptr.p, d_node_of.p, c.K.ptr.p, c.K.col.p, c.K.val.p, d_run_ptr.p, d_sums.p, s);
  std::vector<double> sums((size_t)run_ptr[(size_t)nn]);
  d_sums.download(sums.data(), sums.size(), s);
  SLX_HIP(hipStreamSynchronize(s));
  // the node's own run is its diagonal block; the others ascend with the columns
  b.diag.assign((size_t)nn, 0.0);
  b.ptr.assign((size_t)nn + 1, 0);
  for (int nd = 0; nd < nn; ++nd) b.ptr[(size_t)nd + 1] = b.ptr[(size_t)nd] + (run_ptr[(size_t)nd + 1] - run_ptr[(size_t)nd]) - 1;
  b.node.resize((size_t)b.ptr[(size_t)nn]);
  b.sumsq.resize(b.node.size());
  std::atomic<bool> ok{true};
#pragma omp parallel for schedule(dynamic, 1024)
  for (int nd = 0; nd < nn; ++nd) {
    const int r0 = node_ptr[(size_t)nd];
    int t = run_ptr[(size_t)nd], q = b.ptr[(size_t)nd];
    bool own = false;
    for (int k = a.ptr[(size_t)r0]; k < a.ptr[(size_t)r0 + 1]; ++t) {
      const int J = node_of[(size_t)a.col[(size_t)k]];
      k += node_ptr[(size_t)J + 1] - node_ptr[(size_t)J];
      if (J == nd) {
        b.diag[(size_t)nd] = sums[(size_t)t];
        own = true;
      } else if (q < b.ptr[(size_t)nd + 1]) {
        b.node[(size_t)q] = J;
        b.sumsq[(size_t)q++] = sums[(size_t)t];
      }
    }
    if (!own) ok.store(false, std::memory_order_relaxed);
  }
  return ok.load();
}

// Lanczos estimate of lambda_max(D^-1 A) on the device (Multigrid.cpp:288-327).
// All steps are queued with their scalars in device memory (lanczos_queue
// returns the step count); one read brings alpha and beta back, and
// lanczos_lambda cuts the recurrence where the reference breaks. Steps queued
// after such a breakdown compute values nobody reads.
int lanczos_queue(Context& c, size_t l) {
  Level& lv = c.levels[l];
  DCsr& A = level_op(c, l);
  const int n = lv.n;
  hipStream_t s = c.stream;
  if (lv.lz_start.n != (size_t)n) {
    std::vector<double> v0((size_t)n);
    for (int i = 0; i < n; ++i) v0[(size_t)i] = 1.0 + 0.5 * std::sin(1.7 * double(i) + 0.3);
    lv.lz_start.upload(v0, s);
  }
  const int m = std::max(1, std::min(c.amg_opt.lanczos_steps, n));
  lv.sq.alloc((size_t)n);
  lv.lz_v.alloc((size_t)n);
  lv.lz_vprev.alloc((size_t)n);
  lv.lz_w.alloc((size_t)n);
  lv.lz_part.alloc((size_t)k::axmy_blocks(n));
  DBuf<double>& sc = lv.lz_scal;  // [0] norm, [1..m] alpha, [m+1..2m] beta
  sc.alloc((size_t)(2 * m + 1));
  k::vsqrt(n, lv.dinv.p, lv.sq.p, s);
  k::dot(n, lv.lz_start.p, lv.lz_start.p, c.partials.p, sc.p, s);
  k::sqrt_scalar(sc.p, s);
  k::copy(n, lv.lz_start.p, lv.lz_w.p, s);
  k::lanczos_shift(n, sc.p, lv.lz_w.p, lv.lz_v.p, lv.lz_vprev.p, s);  // v = v0 / |v0|
  k::fill(n, 0.0, lv.lz_vprev.p, s);
  double* alpha = sc.p + 1;
  double* beta = sc.p + 1 + m;
  for (int j = 0; j < m; ++j) {
    // w = s A (s v) - beta_{j-1} vprev with the partials of w.v; alpha_j = w.v;
    // w -= alpha_j v with the partials of w.w; beta_j = |w|; v, vprev advance
    const double* bprev = j > 0 ? beta + j - 1 : nullptr;
    const int nb = &A == &c.K ? op_lanczos(c, lv.sq.p, lv.lz_v.p, lv.lz_vprev.p, bprev, lv.lz_w.p, c.partials.p)
                              : (k::lanczos_apply(A, lv.sq.p, lv.lz_v.p, lv.lz_vprev.p, bprev, lv.lz_w.p,
                                                  c.partials.p, s),
                                 k::lanczos_blocks(n));
    k::axmy_dot_sum(n, nb, c.partials.p, alpha + j, lv.lz_v.p, lv.lz_w.p, lv.lz_part.p, s);
    k::lanczos_shift_sum(n, k::axmy_blocks(n), lv.lz_part.p, beta + j, lv.lz_w.p, lv.lz_v.p, lv.lz_vprev.p, s);
  }
  return m;
}

// lambda_max from the Lanczos scalars read back (h = lz_scal): the recurrence
// is cut where the reference breaks (beta <= 1e-12 |alpha|)
double lanczos_lambda(const std::vector<double>& h, int m) {
  std::vector<double> a, b;
  for (int j = 0; j < m; ++j) {
    const double aj = h[(size_t)(1 + j)], bj = h[(size_t)(1 + m + j)];
    a.push_back(aj);
    if (j + 1 == m || !(bj > 1.0e-12 * std::abs(aj))) break;
    b.push_back(bj);
  }
  std::vector<double> sub(b.begin(), b.begin() + (long)(a.size() - 1));
  return tridiag_max(a, sub);
}

// whether R = P^T can be gathered by coarse node: every row of a coarse node
// has the length of the node's first row, and its entry k reads the value after
// the one the row before reads
bool node_gather_ok(const HCsr& r, const std::vector<int>& src, const std::vector<int>& cnode) {
  if (cnode.size() < 2 || cnode.back() != r.rows) return false;
  const int nn = (int)cnode.size() - 1;
  bool ok = true;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int a = 0; a < nn; ++a) {
    const int j0 = cnode[(size_t)a], d = cnode[(size_t)a + 1] - j0;
    const int b0 = r.ptr[(size_t)j0], len = r.ptr[(size_t)j0 + 1] - b0;
    for (int q = 1; q < d && ok; ++q) {
      const int bq = r.ptr[(size_t)j0 + q];
      if (r.ptr[(size_t)j0 + q + 1] - bq != len) {
        ok = false;
        break;
      }
      for (int k = 0; k < len; ++k)
        if (src[(size_t)bq + k] != src[(size_t)b0 + k] + q) {
          ok = false;
          break;
        }
    }
  }
  return ok;
}

void r_transpose(Level& lv, int ncoarse_nodes, hipStream_t s) {
  if (lv.r_nodes)
    k::gather_nodes(ncoarse_nodes, lv.r_cnode.p, lv.R.ptr.p, lv.r_src.p, lv.P.val.p, lv.R.val.p, s);
  else
    k::gather(lv.R.nnz, lv.r_src.p, lv.P.val.p, lv.R.val.p, s);
}

// Dependency levels of the symmetric Gauss-Seidel sweeps of one level
// (Multigrid.cpp:788-806): in the forward sweep row i reads the new values of
// the rows j < i in its pattern, so its level is one above the highest of
// theirs; in the backward sweep likewise for the rows j > i. Rows of one level
// never read each other's unknowns. Each level's rows are kept ascending.
void gs_levels(const int* ptr, const int* col, int n, bool forward, std::vector<int>& lptr, std::vector<int>& rows) {
  std::vector<int> lev((size_t)n, 0);
  int nl = 0;
  for (int q = 0; q < n; ++q) {
    const int i = forward ? q : n - 1 - q;
    int m = 0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int j = col[k];
      if (forward ? j < i : j > i) m = std::max(m, lev[(size_t)j] + 1);
    }
    lev[(size_t)i] = m;
    nl = std::max(nl, m + 1);
  }
  lptr.assign((size_t)nl + 1, 0);
  for (int i = 0; i < n; ++i) ++lptr[(size_t)lev[(size_t)i] + 1];
  for (int L = 0; L < nl; ++L) lptr[(size_t)L + 1] += lptr[(size_t)L];
  rows.assign((size_t)n, 0);
  std::vector<int> fill(lptr.begin(), lptr.end() - 1);
  for (int i = 0; i < n; ++i) rows[(size_t)fill[(size_t)lev[(size_t)i]]++] = i;
}

// the sweeps' schedules of every smoothed level, built once per structure from
// the level's pattern (level 0: the host copy of K's pattern when it is current)
void gs_schedules(Context& c) {
  hipStream_t s = c.stream;
  for (size_t l = 0; l + 1 < c.levels.size(); ++l) {
    Level& lv = c.levels[l];
    if (!lv.gs_fptr.empty() && lv.gs_frows.n == (size_t)lv.n) continue;
    const DCsr& A = level_op(c, l);
    std::vector<int> hptr, hcol;
    const int *ptr, *col;
    if (l == 0 && c.k_host_version == c.k_pattern_version && c.k_host.rows == A.rows) {
      ptr = c.k_host.ptr.data();
      col = c.k_host.col.data();
    } else {
      hptr.resize((size_t)A.rows + 1);
      hcol.resize((size_t)A.nnz);
      A.ptr.download(hptr.data(), hptr.size(), s);
      A.col.download(hcol.data(), hcol.size(), s);
      SLX_HIP(hipStreamSynchronize(s));
      ptr = hptr.data();
      col = hcol.data();
    }
    std::vector<int> rows;
    gs_levels(ptr, col, A.rows, true, lv.gs_fptr, rows);
    lv.gs_frows.upload(rows, s);
    gs_levels(ptr, col, A.rows, false, lv.gs_bptr, rows);
    lv.gs_brows.upload(rows, s);
    SLX_HIP(hipStreamSynchronize(s));
  }
}

// D^-1 and lambda_max(D^-1 A) of level l with one read from the device: the
// Lanczos scalars, the level's diagonal flag and, at level 0, the finiteness
// sum of the matrix queued by amg_setup.
void numeric_level(Context& c, size_t l) {
  Level& lv = c.levels[l];
  DCsr& A = level_op(c, l);
  hipStream_t s = c.stream;
  lv.dinv.alloc((size_t)lv.n);
  int* bad = c.setup_flags.p + 1 + l;
  constexpr int kNoBadRow = 0x7F7F7F7F;  // above any row index
  SLX_HIP(hipMemsetAsync(bad, 0x7F, sizeof(int), s));
  if (l == 0 && c.mf)  // from the element factors: the assembled diagonal bit for bit, K not read
    kmf::inv_diag(c.nx + 1, c.ny + 1, c.nz + 1, c.d_reduced.p, c.mf_scale.p, c.d_ke0.p, lv.dinv.p, bad, s);
  else
    k::inv_diag(A, lv.dinv.p, bad, s);
  const int m = lanczos_queue(c, l);
  std::vector<double> h((size_t)(2 * m + 1));
  lv.lz_scal.download(h.data(), h.size(), s);
  int b = kNoBadRow;
  SLX_HIP(hipMemcpyAsync(&b, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  double sumsq = 0.0;
  if (l == 0) SLX_HIP(hipMemcpyAsync(&sumsq, c.scal.p + 9, sizeof(double), hipMemcpyDeviceToHost, s));
  SLX_HIP(hipStreamSynchronize(s));
  // finiteness of every entry, each call (Multigrid.cpp:992-996): the sum of
  // squares is finite exactly when every entry is
  if (l == 0 && !std::isfinite(sumsq)) fail(SLX_SOLVER_ERROR, "multigrid: the system matrix has a non-finite entry");
  if (b != kNoBadRow) {
    std::ostringstream os;
    os << "multigrid level " << l << ": diagonal entry " << b
       << " is not positive; a stiffness matrix must have a positive diagonal. The model has an unconnected degree "
          "of freedom or a non-positive stiffness";
    fail(SLX_SOLVER_ERROR, os.str());
  }
  lv.lambda_max = lanczos_lambda(h, m);
  if (!(lv.lambda_max > 0.0) || !std::isfinite(lv.lambda_max)) {
    std::ostringstream os;
    os << "multigrid level " << l << ": the spectral estimate of D^-1 A is " << lv.lambda_max
       << "; the matrix is not positive definite";
    fail(SLX_SOLVER_ERROR, os.str());
  }
}

// The CSR kernels' lanes per row, once per structure. A row's sum order depends
// on the lane count, so the FP64 path keeps the kernels' fixed rule from the mean
// row length (lanes 0), and its results are the same on every run; the FP32
// V-cycle of the mixed path times each operator and keeps the fastest.
void autotune(Context& c) {
  const char* e = std::getenv("SPARLAB_NO_AUTOTUNE");
  if (e && e[0] == '1') return;
  hipStream_t s = c.stream;
  const bool f32 = c.amg_opt.precision >= 1;
  std::vector<DCsr*> mats;
  for (size_t l = 0; l + 1 < c.levels.size(); ++l) {
    if (l > 0 || !c.mf) {
      mats.push_back(&level_op(c, l));
      mats.push_back(&c.levels[l].P);
      mats.push_back(&c.levels[l].R);
    } else {
      mats.push_back(&c.levels[l].Ptent);
      mats.push_back(&c.levels[l].Rt);
    }
  }
  if (!f32) {
    for (DCsr* m : mats) m->lanes = 0;
    return;
  }
  size_t rows = 1, cols = 1;
  for (DCsr* m : mats) {
    rows = std::max(rows, (size_t)m->rows);
    cols = std::max(cols, (size_t)m->cols);
  }
  DBuf<double> x, y;
  DBuf<float> x32, y32;
  x.alloc(cols);
  k::fill((long long)cols, 1.0, x.p, s);
  if (f32) {
    x32.alloc(cols);
    y32.alloc(rows);
    k::to_float((long long)cols, x.p, x32.p, s);
  } else {
    y.alloc(rows);
  }
  hipEvent_t a = c.take_event(), b = c.take_event();
  for (DCsr* m : mats) {
    if (m->rows == 0 || (f32 && !m->val32.p)) continue;
    float best = std::numeric_limits<float>::max();
    int pick = 0;
    for (int lanes : {1, 4, 8, 16, 64}) {
      m->lanes = lanes;
      auto run = [&] {
        if (f32) k::spmv32(*m, x32.p, y32.p, s); else k::spmv(*m, x.p, y.p, s);
      };
      run();  // warm
      // the best of three batches: one noisy batch must not pick a width
      float ms = std::numeric_limits<float>::max();
      for (int batch = 0; batch < 3; ++batch) {
        SLX_HIP(hipEventRecord(a, s));
        for (int r = 0; r < 4; ++r) run();
        SLX_HIP(hipEventRecord(b, s));
        SLX_HIP(hipEventSynchronize(b));
        float t = 0.0f;
        SLX_HIP(hipEventElapsedTime(&t, a, b));
        ms = std::min(ms, t);
      }
      if (ms < best) {
        best = ms;
        pick = lanes;
      }
    }
    m->lanes = pick;
  }
  c.event_pool.push_back(a);
  c.event_pool.push_back(b);
}

}  // namespace

double device_dot(Context& c, long long n, const double* a, const double* b) {
  k::dot(n, a, b, c.partials.p, c.scal.p + 15, c.stream);
  double* hb = c.host_buffer(16);
  SLX_HIP(hipMemcpyAsync(hb + 15, c.scal.p + 15, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SLX_HIP(hipStreamSynchronize(c.stream));
  return hb[15];
}

// The pivot check of the coarsest level's inverse (Multigrid.cpp:766-781): the
// device inverse records the pivots of the unpivoted LDL^T; an unresisted
// rigid-body mode shows as a round-off sized (or non-positive) pivot.
void finish_coarse(Context& c) {
  if (!c.coarse_pending) return;
  if (c.capturing) fail(SLX_SOLVER_ERROR, "multigrid: internal error (coarse inverse pending during graph capture)");
  c.coarse_pending = false;
  std::vector<double> d((size_t)c.coarse_n);
  c.coarse_piv.download(d.data(), d.size(), c.stream);
  SLX_HIP(hipStreamSynchronize(c.stream));
  double mx = 0.0, mn = std::numeric_limits<double>::max();
  bool finite = true;
  for (double v : d) {
    finite = finite && std::isfinite(v);
    mx = std::max(mx, std::abs(v));
    mn = std::min(mn, v);
  }
  c.coarse_pivot_ratio = finite && mx > 0.0 ? mn / mx : 0.0;
  if (!(c.coarse_pivot_ratio > c.amg_opt.coarse_pivot_tolerance)) {
    c.amg_ready = false;
    std::ostringstream os;
    os << "multigrid: the coarsest operator (" << c.coarse_n << " unknowns) is singular or indefinite (smallest / "
       << "largest pivot " << c.coarse_pivot_ratio << "). A rigid-body mode survived to the coarse level, so the model "
       << "is under-constrained: too few supports, or a part connected to nothing";
    fail(SLX_SOLVER_ERROR, os.str());
  }
}

// Level 0's diagonal and spectral bound from the current element factors, the
// coarse levels of the last full numeric setup kept (coarse_reuse of the mixed
// path: A_1 and below are those of an earlier design). The level-0 smoother and
// transfers use the current operator: P = (I - omega D^-1 K) P-hat is applied
// through it, with D and omega refreshed here. Matrix-free Hex8 models only.
void amg_refresh_level0(Context& c) {
  const auto t0 = Clock::now();
  if (!c.amg_ready || !c.mf || c.levels.empty()) fail(SLX_SOLVER_ERROR, "multigrid: no hierarchy to refresh");
  invalidate_graphs(c);  // the Chebyshev and transfer coefficients are kernel arguments
  hipStream_t s = c.stream;
  Level& lv = c.levels[0];
  int* bad = c.setup_flags.p + 1;
  constexpr int kNoBadRow = 0x7F7F7F7F;
  SLX_HIP(hipMemsetAsync(bad, 0x7F, sizeof(int), s));
  kmf::inv_diag(c.nx + 1, c.ny + 1, c.nz + 1, c.d_reduced.p, c.mf_scale.p, c.d_ke0.p, lv.dinv.p, bad, s);
  k::dot(c.nelem, c.mf_scale.p, c.mf_scale.p, c.partials.p, c.scal.p + 9, s);
  const int m = lanczos_queue(c, 0);
  std::vector<double> h((size_t)(2 * m + 1));
  lv.lz_scal.download(h.data(), h.size(), s);
  int b = kNoBadRow;
  double sumsq = 0.0;
  SLX_HIP(hipMemcpyAsync(&b, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  SLX_HIP(hipMemcpyAsync(&sumsq, c.scal.p + 9, sizeof(double), hipMemcpyDeviceToHost, s));
  SLX_HIP(hipStreamSynchronize(s));
  if (!std::isfinite(sumsq)) fail(SLX_SOLVER_ERROR, "multigrid: the system matrix has a non-finite entry");
  if (b != kNoBadRow) {
    std::ostringstream os;
    os << "multigrid level 0: diagonal entry " << b
       << " is not positive; a stiffness matrix must have a positive diagonal. The model has an unconnected degree "
          "of freedom or a non-positive stiffness";
    fail(SLX_SOLVER_ERROR, os.str());
  }
  lv.lambda_max = lanczos_lambda(h, m);
  if (!(lv.lambda_max > 0.0) || !std::isfinite(lv.lambda_max))
    fail(SLX_SOLVER_ERROR, "multigrid level 0: the spectral estimate of D^-1 A is not positive and finite");
  lv.omega = c.amg_opt.prolongator_damping / lv.lambda_max;
  if (c.amg_opt.precision >= 1) k::to_float(lv.n, lv.dinv.p, lv.dinv32.p, s);
  for (double& ph : c.phase_seconds) ph = 0.0;
  c.setup_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
  c.phase_seconds[1] = c.setup_seconds;
  c.reused = true;
}

void amg_setup(Context& c, bool defer_coarse) {
  const auto t0 = Clock::now();
  hipStream_t s = c.stream;
  const int n = c.K.rows;
  if (c.k_stale) {  // the last evaluation refreshed level 0 only: assemble its factors now
    k::assemble_structured(c.dim, c.K, c.d_free.p, c.mf_scale.p, c.d_ke0.p, c.nx, c.ny, c.nz, s, c.d_dofxyz.p,
                           c.d_nodemap.p);
    c.k_stale = false;
  }
  // coarse_reuse restarts from this setup: the next evaluation sets its base count
  c.coarse_lag = 0;
  c.lag_base_its = 0;
  c.coarse_pending = false;  // a previous setup's inverse nobody used
  invalidate_graphs(c);
  for (double& ph : c.phase_seconds) ph = 0.0;
  // phase times from device events (no waits between the phases): event k
  // closes phase lap_phase[k]
  std::vector<std::pair<int, hipEvent_t>> laps;
  hipEvent_t lap_start = c.take_event();
  SLX_HIP(hipEventRecord(lap_start, s));
  auto lap = [&](int phase) {
    hipEvent_t e = c.take_event();
    SLX_HIP(hipEventRecord(e, s));
    laps.push_back({phase, e});
  };
  if (n == 0) fail(SLX_SOLVER_ERROR, "multigrid: the system matrix must be square and non-empty");
  if (c.layout_dim != 2 && c.layout_dim != 3) fail(SLX_SOLVER_ERROR, "multigrid: the DOF layout must be two- or three-dimensional");
  if (c.layout_ndpn != c.layout_dim && !(c.layout_dim == 3 && c.layout_ndpn == 6))
    fail(SLX_SOLVER_ERROR, "multigrid: unsupported DOFs per node (translations only, or 3 + 3 in 3-D)");
  if ((int)c.layout_unknowns.size() != n) fail(SLX_SOLVER_ERROR, "multigrid: the DOF layout does not match the matrix");
  const slx_amg_options& o = c.amg_opt;
  if (o.smoother != 0 && o.smoother != 1)
    fail(SLX_CONFIG_ERROR, "multigrid: unknown smoother (0 Chebyshev, 1 symmetric Gauss-Seidel)");
  ensure_workspace(c);
  const bool same = c.amg_ready && c.fp_rows == n && c.fp_nnz == c.K.nnz && c.fp_layout_version == c.layout_version &&
                    c.fp_pattern_version == c.k_pattern_version;
  const bool reuse = o.reuse_aggregates && same;
  c.amg_ready = false;  // until this setup completes
  // device flags: [0] a Galerkin pattern mismatch, [1 + l] level l's first
  // non-positive diagonal row; read with the level's Lanczos scalars and at the end
  c.setup_flags.alloc((size_t)o.max_levels + 2);
  SLX_HIP(hipMemsetAsync(c.setup_flags.p, 0, sizeof(int), s));
  // finiteness of every entry (Multigrid.cpp:992-996): the sum of squares,
  // checked with level 0's read; a matrix-free model's K = sum_e s_e Ke0 is
  // finite with its element factors, whose squares are summed instead
  if (c.mf)
    k::dot(c.nelem, c.mf_scale.p, c.mf_scale.p, c.partials.p, c.scal.p + 9, s);
  else
    k::dot(c.K.nnz, c.K.val.p, c.K.val.p, c.partials.p, c.scal.p + 9, s);
  setup_trace_start();
  auto trace = [&](const char* what, size_t lvl) { setup_trace(what, (int)lvl, s); };
  HCsr host_a;  // the level being aggregated (first setup only): pattern and values
  // Level 0: K's pattern is on the host where it was uploaded; the symmetry
  // check and the strength sums read the values on the device, so K itself is
  // not brought back.
  const bool dev0 = !reuse && c.k_host_version == c.k_pattern_version && c.k_host.rows == n &&
                    c.k_host.nnz() == c.K.nnz;
  if (!reuse) {
    // symmetry check once per structure (Multigrid.cpp:968-991)
    int bad_i = -1, bad_j = -1;
    if (dev0) {
      DBuf<unsigned long long> bad;
      bad.alloc(1);
      SLX_HIP(hipMemsetAsync(bad.p, 0xff, sizeof(unsigned long long), s));
      k::symmetry_check(c.K, 1.0e-10, c.partials.p, bad.p, s);
      unsigned long long hb = ~0ULL;
      bad.download(&hb, 1, s);
      SLX_HIP(hipStreamSynchronize(s));
      if (hb != ~0ULL) {
        bad_i = (int)(hb >> 32);
        bad_j = (int)(hb & 0xffffffffULL);
      }
    } else {
      host_a = download_csr(c.K, s);
      trace("download K", 0);
      double max_abs = 0.0;
      for (double v : host_a.val) max_abs = std::max(max_abs, std::abs(v));
#pragma omp parallel for schedule(dynamic, 4096)
      for (int i = 0; i < n; ++i)
        for (int k = host_a.ptr[(size_t)i]; k < host_a.ptr[(size_t)i + 1]; ++k) {
          const int j = host_a.col[(size_t)k];
          const int* b0 = host_a.col.data() + host_a.ptr[(size_t)j];
          const int* b1 = host_a.col.data() + host_a.ptr[(size_t)j + 1];
          const int* hit = std::lower_bound(b0, b1, i);
          const double other = (hit != b1 && *hit == i) ? host_a.val[(size_t)(hit - host_a.col.data())] : 0.0;
          if (std::abs(other - host_a.val[(size_t)k]) > 1.0e-10 * max_abs) {
#pragma omp critical
            {
              bad_i = i;
              bad_j = j;
            }
          }
        }
    }
    if (bad_i >= 0) {
      std::ostringstream os;
      os << "multigrid: the system matrix is not symmetric (entry (" << bad_i << ", " << bad_j
         << ")); conjugate gradients needs a symmetric positive-definite matrix";
      fail(SLX_SOLVER_ERROR, os.str());
    }
    c.levels.clear();
    c.levels.reserve((size_t)o.max_levels + 1);
    c.levels.emplace_back();
    c.levels[0].n = n;
    trace("symmetry check", 0);
    build_finest_nodes(c, c.levels[0]);
    c.amg_ready = false;
    trace("finest nodes", 0);
  }
  lap(0);
  int* flag = c.setup_flags.p;
  size_t l = 0;
  const size_t reused_levels = c.levels.size();
  while (true) {
    numeric_level(c, l);
    lap(1);
    trace("numeric level (Lanczos)", l);
    Level& lv = c.levels[l];
    const bool stop = reuse ? (l + 1 == reused_levels)
                            : (lv.n <= o.coarse_size || (int)l + 1 >= o.max_levels);
    if (stop) break;
    if (!reuse) {
      if (l > 0) host_a = download_csr(lv.A, s);
      const HCsr* hap = &host_a;  // the level's pattern (values only where downloaded)
      if (l == 0 && dev0) {
        NodeBlocks nb;
        if (device_node_blocks(c, lv.node_ptr, nb)) {
          aggregate_blocks(nb, o.strength_threshold, lv.aggregate);
          hap = &c.k_host;
        } else {  // a pattern the device sums do not cover: the host reads K
          host_a = download_csr(c.K, s);
          aggregate_nodes(host_a, lv.node_ptr, o.strength_threshold, lv.aggregate);
        }
      } else {
        aggregate_nodes(host_a, lv.node_ptr, o.strength_threshold, lv.aggregate);
      }
      const HCsr& ha = *hap;
      lap(2);
      trace("aggregation", l);
      Level next;
      HCsr pt;
      tentative(lv, next, c.nullspace_dim, pt);
      lap(6);
      trace("tentative prolongator", l);
      if (next.n >= (int)(0.85 * double(lv.n)) || next.n == 0) {
        lv.aggregate.clear();
        break;
      }
      // symbolic patterns of P, R, A P, R (A P) and the coarse operator
      HCsr ppat = pattern_product(ha, pt);
      trace("pattern P", l);
      HCsr rpat;
      std::vector<int> rsrc;
      transpose_pattern(ppat, rpat, rsrc);
      trace("pattern R", l);
      HCsr appat = pattern_product(ha, ppat);
      trace("pattern A P", l);
      HCsr rawpat = pattern_product(rpat, appat);
      trace("pattern R A P", l);
      HCsr spat;
      std::vector<int> ia, ib;
      symmetrize_pattern(rawpat, spat, ia, ib);
      trace("pattern symmetrised", l);
      lv.Ptent.upload(pt, s);
      if (l == 0 && c.mf) build_tent_blocks(lv, pt, s);
      {
        HCsr rt;
        std::vector<int> rtsrc;
        transpose_pattern(pt, rt, rtsrc);
        rt.val.resize(rtsrc.size());
        for (size_t q = 0; q < rtsrc.size(); ++q) rt.val[q] = pt.val[(size_t)rtsrc[q]];
        lv.Rt.upload(rt, s);
        share_columns(lv.Rt, rt, next.node_ptr, s);
      }
      lv.P.upload_pattern(ppat, s);
      lv.R.upload_pattern(rpat, s);
      lv.r_src.upload(rsrc, s);
      lv.r_nodes = node_gather_ok(rpat, rsrc, next.node_ptr);
      if (lv.r_nodes) lv.r_cnode.upload(next.node_ptr, s);
      lv.AP.upload_pattern(appat, s);
      lv.RAW.upload_pattern(rawpat, s);
      lv.sym_a.upload(ia, s);
      lv.sym_b.upload(ib, s);
      next.A.upload_pattern(spat, s);
      trace("uploads", l);
      share_columns(next.A, spat, next.node_ptr, s);
      share_columns(lv.P, ppat, lv.node_ptr, s);
      share_columns(lv.R, rpat, next.node_ptr, s);
      share_columns(lv.Ptent, pt, lv.node_ptr, s);
      if (l == 0 && !c.mf) share_columns(c.K, ha, lv.node_ptr, s);
      trace("shared columns", l);
      std::vector<double> xyz;  // level-0 node coordinates (mesh node of each node block's first unknown)
      if (l == 0 && !c.layout_coords.empty() && c.layout_ndpn > 0) {
        const int nn = (int)lv.node_ptr.size() - 1;
        xyz.assign((size_t)nn * 3, 0.0);
        for (int nd = 0; nd < nn; ++nd) {
          const int mesh = c.layout_unknowns[(size_t)lv.node_ptr[(size_t)nd]] / c.layout_ndpn;
          for (int d = 0; d < c.layout_dim; ++d)
            xyz[(size_t)nd * 3 + d] = c.layout_coords[(size_t)mesh * c.layout_dim + d];
        }
      }
      prepare_node_products(lv, next, ha, pt, ppat, rpat, appat, rawpat, xyz.empty() ? nullptr : &xyz, s);
      lap(4);  // the symbolic products and their plans count as Galerkin time
      trace("product plans", l);
      lv.has_next = true;
      SLX_HIP(hipStreamSynchronize(s));  // host pattern arrays go out of scope
      c.levels.push_back(std::move(next));
    }
    Level& cur = c.levels[l];
    Level& coarse = c.levels[l + 1];
    DCsr& A = level_op(c, l);
    const double omega = o.prolongator_damping / cur.lambda_max;
    cur.omega = omega;
    if (cur.node_ok) {
      k::BlockProductArgs g = plan_args(cur, cur.plan_p, false, A, cur.Ptent, cur.P);
      g.dinv = cur.dinv.p;
      g.omega = omega;
      g.tptr = cur.Ptent.ptr.p;
      g.tval = cur.Ptent.val.p;
      g.row_agg = cur.row_agg.p;
      k::block_product(g, true, s);
      r_transpose(cur, (int)coarse.node_ptr.size() - 1, s);
      lap(3);
      {
        Timed t(c, l == 0 ? "galerkin_ap" : nullptr);
        k::row_product(row_args(cur, cur.plan_ap, false, A, cur.P, cur.AP), s);
      }
      {
        Timed t(c, l == 0 ? "galerkin_raw" : nullptr);
        k::row_product(row_args(cur, cur.plan_raw, true, cur.R, cur.AP, cur.RAW), s);
      }
    } else {
      k::smooth_prolongator(A, cur.Ptent, cur.P, cur.dinv.p, omega, flag, s);
      r_transpose(cur, (int)coarse.node_ptr.size() - 1, s);
      k::spgemm_numeric(A, cur.P, cur.AP, flag, s);
      k::spgemm_numeric(cur.R, cur.AP, cur.RAW, flag, s);
    }
    k::symmetrize_values(coarse.A.nnz, cur.sym_a.p, cur.sym_b.p, cur.RAW.val.p, coarse.A.val.p, s);
    lap(4);
    trace("numeric products", l);
    ++l;
  }
  c.levels.resize(l + 1);
  c.levels.back().has_next = false;
  // coarsest level: dense inverse on the device (its pivots are checked by
  // finish_coarse before the first V-cycle)
  {
    const int nc = c.levels.back().n;
    if (nc > kDenseCoarseMax) {
      std::ostringstream os;
      os << "multigrid: the coarsest level has " << nc << " unknowns, above the " << kDenseCoarseMax
         << " of the device's dense coarse solver; lower amg.coarse_size or raise amg.max_levels";
      fail(SLX_SOLVER_ERROR, os.str());
    }
    const DCsr& ac = level_op(c, l);
    c.coarse_inv.alloc((size_t)nc * nc);
    c.coarse_piv.alloc((size_t)nc);
    c.coarse_work.alloc((size_t)k::dense_inverse_work(nc));
    k::dense_inverse(nc, ac.ptr.p, ac.col.p, ac.val.p, c.coarse_inv.p, c.coarse_piv.p, c.coarse_work.p, s);
    if (o.precision >= 1) {
      c.coarse_inv32.alloc(c.coarse_inv.n);
      k::to_float((long long)c.coarse_inv.n, c.coarse_inv.p, c.coarse_inv32.p, s);
    }
    c.coarse_n = nc;
    c.coarse_pending = true;
  }
  lap(5);
  for (auto& lv : c.levels) {
    lv.x.alloc((size_t)lv.n);
    lv.x2.alloc((size_t)lv.n);
    lv.b.alloc((size_t)lv.n);
    lv.res.alloc((size_t)lv.n);
    lv.d.alloc((size_t)lv.n);
    lv.d2.alloc((size_t)lv.n);
  }
  if (o.smoother == 1) gs_schedules(c);
  if (o.precision >= 1) prepare_mixed(c);
  if (!reuse || c.tuned_precision != o.precision) {
    autotune(c);
    tune_tile(c);
    c.tuned_precision = o.precision;
  }
  int galerkin_flag = 0;
  SLX_HIP(hipMemcpyAsync(&galerkin_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  SLX_HIP(hipStreamSynchronize(s));
  {
    float ms = 0.0f;
    hipEvent_t prev = lap_start;
    for (auto& pe : laps) {
      SLX_HIP(hipEventElapsedTime(&ms, prev, pe.second));
      c.phase_seconds[pe.first] += 1.0e-3 * ms;
      c.event_pool.push_back(prev);
      prev = pe.second;
    }
    c.event_pool.push_back(prev);
  }
  if (galerkin_flag != 0) fail(SLX_SOLVER_ERROR, "multigrid: internal error in the Galerkin product (pattern mismatch)");
  ++c.setup_id;
  c.fp_rows = n;
  c.fp_nnz = c.K.nnz;
  c.fp_layout_version = c.layout_version;
  c.fp_pattern_version = c.k_pattern_version;
  c.amg_ready = true;
  c.reused = reuse;
  if (!defer_coarse) finish_coarse(c);
  c.setup_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace slx
