// Segmented (batched) CG: a batch of independent systems -- designs of
// different sizes, or the load cases of one design -- assembled as one
// block-diagonal matrix and solved together. The multigrid hierarchy of a
// block-diagonal matrix is block-diagonal (aggregates never cross blocks that
// share no entry), so every segment runs the reference's PCG
// (src/fem/Multigrid.cpp:1152-1237) with its own step lengths, residual,
// convergence test and restarts, while every kernel launch covers the whole
// batch and the iteration is replayed from captured graphs.
#include <algorithm>
#include <cmath>
#include <limits>
#include <sstream>

#include "slx_kernels.hpp"
#include "slx_seg.hpp"
#include "slx_solver.hpp"

namespace slx {

struct SegState {
  std::vector<int> seg_ptr;
  int nseg = 0;
  DBuf<int> ch_begin, ch_end, ch_seg, seg_chunk, d_seg_ptr, row_seg;
  kseg::Chunks ch;
  DBuf<double> pq, rr, rz, rzn, rzx, beta, part;
  DBuf<int> active;
  DBuf<int> flag;
  hipGraphExec_t g1 = nullptr, g2 = nullptr, g3 = nullptr;
  long long setup_id = -1;
  const void* key = nullptr;  // workspace the graphs were captured on
  bool flexible = false;
  void drop_graphs() {
    for (hipGraphExec_t* g : {&g1, &g2, &g3}) {
      if (*g) (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  }
  ~SegState() { drop_graphs(); }
};

namespace {

constexpr int kChunkRows = 4096;

SegState& prepare(Context& c, const std::vector<int>& seg_ptr) {
  if (!c.seg) c.seg = std::make_shared<SegState>();
  SegState& st = *c.seg;
  const int n = c.K.rows;
  const bool same = st.seg_ptr == seg_ptr && st.setup_id == c.setup_id;
  if (!same) {
    st.drop_graphs();
    st.seg_ptr = seg_ptr;
    st.nseg = (int)seg_ptr.size() - 1;
    std::vector<int> cb, ce, cs, sc(1, 0), rs((size_t)n);
    for (int s = 0; s < st.nseg; ++s) {
      for (int i = seg_ptr[(size_t)s]; i < seg_ptr[(size_t)s + 1]; i += kChunkRows) {
        cb.push_back(i);
        ce.push_back(std::min(i + kChunkRows, seg_ptr[(size_t)s + 1]));
        cs.push_back(s);
      }
      sc.push_back((int)cb.size());
      for (int i = seg_ptr[(size_t)s]; i < seg_ptr[(size_t)s + 1]; ++i) rs[(size_t)i] = s;
    }
    hipStream_t s = c.stream;
    st.ch_begin.upload(cb, s);
    st.ch_end.upload(ce, s);
    st.ch_seg.upload(cs, s);
    st.seg_chunk.upload(sc, s);
    st.d_seg_ptr.upload(seg_ptr, s);
    st.row_seg.upload(rs, s);
    st.ch.count = (int)cb.size();
    st.ch.begin = st.ch_begin.p;
    st.ch.end = st.ch_end.p;
    st.ch.seg = st.ch_seg.p;
    st.ch.seg_chunk = st.seg_chunk.p;
    for (DBuf<double>* b : {&st.pq, &st.rr, &st.rz, &st.rzn, &st.rzx, &st.beta}) b->alloc((size_t)st.nseg);
    st.part.alloc(std::max<size_t>(1, cb.size()));
    st.active.alloc((size_t)st.nseg);
    st.flag.alloc(1);
    // the matrix must not couple segments
    const int zero = 0;
    st.flag.upload(&zero, 1, s);
    kseg::check_blocks(n, c.K.ptr.p, c.K.col.p, st.row_seg.p, st.d_seg_ptr.p, st.flag.p, s);
    int bad = 0;
    st.flag.download(&bad, 1, s);
    SLX_HIP(hipStreamSynchronize(s));
    if (bad) {
      st.seg_ptr.clear();
      fail(SLX_CONFIG_ERROR, "segmented solve: the matrix couples rows of different segments (it must be block "
                             "diagonal along the segments)");
    }
    st.setup_id = c.setup_id;
  }
  if (st.key != (const void*)c.r.p) st.drop_graphs();
  return st;
}

std::vector<double> download(Context& c, const DBuf<double>& d, int n) {
  double* hb = c.host_buffer((size_t)n + 16);
  SLX_HIP(hipMemcpyAsync(hb + 16, d.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c.stream));
  SLX_HIP(hipStreamSynchronize(c.stream));
  return std::vector<double>(hb + 16, hb + 16 + n);
}

}  // namespace

void seg_pcg(Context& c, const std::vector<int>& seg_ptr, const double* b, double* x_out, double tol,
             int max_iterations, std::vector<slx_pcg_result>& res) {
  finish_coarse(c);
  SegState& st = prepare(c, seg_ptr);
  const int ns = st.nseg;
  const long long n = c.K.rows;
  hipStream_t s = c.stream;
  const kseg::Chunks& ch = st.ch;
  const bool flexible = c.amg_opt.precision >= 1;
  if (flexible != st.flexible) st.drop_graphs();
  st.flexible = flexible;
  res.assign((size_t)ns, slx_pcg_result{0, 0.0, 0});
  kseg::dot(ch, ns, b, b, st.part.p, st.rr.p, s);
  std::vector<double> bnorm = download(c, st.rr, ns);
  for (double& v : bnorm) v = std::sqrt(v);
  double* x = c.xpcg.p;
  k::copy(n, x_out, x, s);
  std::vector<char> done((size_t)ns, 0), recursive((size_t)ns, 0);
  std::vector<int> act((size_t)ns, 0);
  std::vector<double> best((size_t)ns, std::numeric_limits<double>::max()), rel((size_t)ns, 0.0);
  for (int g = 0; g < ns; ++g)
    if (bnorm[(size_t)g] == 0.0) {  // zero right-hand side: x = 0
      k::fill(seg_ptr[(size_t)g + 1] - seg_ptr[(size_t)g], 0.0, x + seg_ptr[(size_t)g], s);
      done[(size_t)g] = 1;
      res[(size_t)g].converged = 1;
    }
  auto body1 = [&] {
    op_apply(c, c.p.p, c.q.p);
    kseg::dot(ch, ns, c.p.p, c.q.p, st.part.p, st.pq.p, s);
  };
  auto body2 = [&] { kseg::update(ch, ns, x, c.r.p, c.p.p, c.q.p, st.rz.p, st.pq.p, st.active.p, st.part.p, st.rr.p, s); };
  auto body3 = [&] {
    if (flexible) k::copy(n, c.z.p, c.zold.p, s);
    vcycle(c, c.r.p, c.z.p);
    kseg::dot(ch, ns, c.r.p, c.z.p, st.part.p, st.rzn.p, s);
    if (flexible) kseg::dot(ch, ns, c.r.p, c.zold.p, st.part.p, st.rzx.p, s);
    kseg::beta(ns, st.rzn.p, flexible ? st.rzx.p : nullptr, st.rz.p, st.beta.p, s);
    kseg::direction(ch, c.p.p, c.z.p, st.beta.p, s);
  };
  const bool graphs = graphs_enabled();
  if (graphs && !st.g1) {
    st.g1 = capture_graph(c, body1);
    st.g2 = capture_graph(c, body2);
    st.g3 = capture_graph(c, body3);
    st.key = c.r.p;
  }
  auto true_residuals = [&] {
    op_residual(c, x, b, c.r.p);
    kseg::dot(ch, ns, c.r.p, c.r.p, st.part.p, st.rr.p, s);
    std::vector<double> rr = download(c, st.rr, ns);
    for (int g = 0; g < ns; ++g) rel[(size_t)g] = bnorm[(size_t)g] > 0.0 ? std::sqrt(rr[(size_t)g]) / bnorm[(size_t)g] : 0.0;
  };
  for (int restart = 0; restart < 3; ++restart) {
    true_residuals();
    int live = 0;
    for (int g = 0; g < ns; ++g) {
      act[(size_t)g] = 0;
      if (done[(size_t)g]) continue;
      res[(size_t)g].relative_residual = rel[(size_t)g];
      if (rel[(size_t)g] <= tol) {
        done[(size_t)g] = 1;
        res[(size_t)g].converged = 1;
        continue;
      }
      act[(size_t)g] = 1;
      ++live;
    }
    if (!live) break;
    st.active.upload(act.data(), (size_t)ns, s);
    vcycle(c, c.r.p, c.z.p);
    k::copy(n, c.z.p, c.p.p, s);
    kseg::dot(ch, ns, c.r.p, c.z.p, st.part.p, st.rz.p, s);
    std::vector<char> conv((size_t)ns, 0);
    while (live) {
      int most = 0;
      for (int g = 0; g < ns; ++g)
        if (act[(size_t)g]) most = std::max(most, res[(size_t)g].iterations);
      if (most >= max_iterations) break;
      {
        Timed t(c, "spmv");
        if (graphs) SLX_HIP(hipGraphLaunch(st.g1, s)); else body1();
      }
      if (graphs) SLX_HIP(hipGraphLaunch(st.g2, s)); else body2();
      const std::vector<double> pq = download(c, st.pq, ns);
      const std::vector<double> rr = download(c, st.rr, ns);
      bool changed = false;
      for (int g = 0; g < ns; ++g) {
        if (!act[(size_t)g]) continue;
        ++res[(size_t)g].iterations;
        if (!(pq[(size_t)g] > 0.0) || !std::isfinite(pq[(size_t)g])) {
          std::ostringstream os;
          os << "segmented CG: segment " << g << " broke down at iteration " << res[(size_t)g].iterations
             << " (p^T A p = " << pq[(size_t)g] << "): its block of the matrix or of the preconditioner is not "
             << "positive definite";
          fail(SLX_SOLVER_ERROR, os.str());
        }
        if (std::sqrt(rr[(size_t)g]) <= tol * bnorm[(size_t)g] || res[(size_t)g].iterations >= max_iterations) {
          conv[(size_t)g] = std::sqrt(rr[(size_t)g]) <= tol * bnorm[(size_t)g];
          act[(size_t)g] = 0;
          changed = true;
          --live;
        }
      }
      if (!live) break;
      if (changed) st.active.upload(act.data(), (size_t)ns, s);
      if (graphs) SLX_HIP(hipGraphLaunch(st.g3, s)); else body3();
    }
    true_residuals();
    int again = 0;
    for (int g = 0; g < ns; ++g) {
      if (done[(size_t)g]) continue;
      res[(size_t)g].relative_residual = rel[(size_t)g];
      if (rel[(size_t)g] <= tol) {
        done[(size_t)g] = 1;
        res[(size_t)g].converged = 1;
      } else if (!conv[(size_t)g] || !(rel[(size_t)g] < 0.5 * best[(size_t)g])) {
        done[(size_t)g] = 1;  // iteration budget spent, or restarts stopped helping
        res[(size_t)g].converged = conv[(size_t)g] || recursive[(size_t)g];
      } else {
        recursive[(size_t)g] = 1;
        best[(size_t)g] = rel[(size_t)g];
        ++again;
      }
    }
    if (!again) break;
  }
  for (int g = 0; g < ns; ++g)
    if (!done[(size_t)g] && recursive[(size_t)g]) res[(size_t)g].converged = 1;
  k::copy(n, x, x_out, s);
  c.flush_timers();
}

}  // namespace slx
