// HIP kernels of the SA-AMG numeric setup (src/fem/Multigrid.cpp:199-225,
// 671-747): numeric sparse products into a pattern fixed by the host's
// symbolic pass, the prolongator smoothing epilogue, value gathers for the
// transpose and the symmetrisation, the diagonal and the Lanczos steps.
//
// The numeric product C = A B runs one wavefront per row of C with the row's
// columns and accumulators in LDS. The A entries of the row are taken in order
// and, for each, the lanes take the entries of the matching B row in parallel:
// one B row holds distinct columns, so no two lanes touch the same
// accumulator within a step and every entry of C sums its products in the
// order of A's row, the order of the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "slx_kernels.hpp"

namespace slx {
namespace k {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int lower_bound_lds(const int* cols, int len, int c) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// EPI 0: C = A B.  EPI 1: P = (I - omega D^-1 A) P-hat with B = P-hat, i.e.
// C_row = -(omega dinv_i) (A P-hat)_row + P-hat_row (Multigrid.cpp:686-701).
template <int EPI>
__global__ void __launch_bounds__(kWave) k_spgemm_numeric(int rows, const int* __restrict__ aptr,
                                                          const int* __restrict__ acol, const double* __restrict__ aval,
                                                          const int* __restrict__ bptr, const int* __restrict__ bcol,
                                                          const double* __restrict__ bval, const int* __restrict__ cptr,
                                                          const int* __restrict__ ccol, double* __restrict__ cval,
                                                          int lds_cap, const double* __restrict__ dinv, double omega,
                                                          int* __restrict__ error_flag) {
  extern __shared__ double smem[];
  double* acc = smem;
  int* cols = reinterpret_cast<int*>(smem + lds_cap);
  const int lane = threadIdx.x;
  for (int row = blockIdx.x; row < rows; row += gridDim.x) {
    const int cb = cptr[row], len = cptr[row + 1] - cb;
    if (len > lds_cap) {  // the host sizes lds_cap from the longest row
      if (lane == 0) atomicOr(error_flag, 1);
      continue;
    }
    for (int t = lane; t < len; t += kWave) {
      cols[t] = ccol[cb + t];
      acc[t] = 0.0;
    }
    __syncthreads();
    const int ab = aptr[row], ae = aptr[row + 1];
    for (int k0 = ab; k0 < ae; k0 += kWave) {
      // stage up to 64 A entries in registers, then walk them in order
      const int kk = k0 + lane;
      int my_j = 0, my_b0 = 0, my_b1 = 0;
      double my_av = 0.0;
      if (kk < ae) {
        my_j = acol[kk];
        my_av = aval[kk];
        my_b0 = bptr[my_j];
        my_b1 = bptr[my_j + 1];
      }
      const int cnt = min(kWave, ae - k0);
      for (int t = 0; t < cnt; ++t) {
        const double av = __shfl(my_av, t, kWave);
        const int b0 = __shfl(my_b0, t, kWave), b1 = __shfl(my_b1, t, kWave);
        for (int q = b0 + lane; q < b1; q += kWave) {
          const int c = bcol[q];
          const int s = lower_bound_lds(cols, len, c);
          if (s < len && cols[s] == c) acc[s] += av * bval[q];
          else atomicOr(error_flag, 2);
        }
        __syncthreads();
      }
    }
    if (EPI == 1) {
      const double scale = omega * dinv[row];
      for (int t = lane; t < len; t += kWave) acc[t] *= -scale;
      __syncthreads();
      for (int q = bptr[row] + lane; q < bptr[row + 1]; q += kWave) {
        const int c = bcol[q];
        const int s = lower_bound_lds(cols, len, c);
        if (s < len && cols[s] == c) acc[s] += bval[q];
        else atomicOr(error_flag, 4);
      }
      __syncthreads();
    }
    for (int t = lane; t < len; t += kWave) cval[cb + t] = acc[t];
    __syncthreads();
  }
}

__global__ void k_gather(long long n, const int* __restrict__ src, const double* __restrict__ in,
                         double* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[src[i]];
}

// R = P^T by coarse node: the d unknowns of an aggregate are consecutive
// columns of every P row, so their R rows share one pattern and row J0 + q,
// entry k, is P's value src[J0's entry k] + q. One block per coarse node reads
// each run of d values of P once and writes the d rows as coalesced streams.
__global__ void __launch_bounds__(128) k_gather_nodes(const int* __restrict__ cnode, const int* __restrict__ rptr,
                                                      const int* __restrict__ src, const double* __restrict__ in,
                                                      double* __restrict__ out) {
  const int A = blockIdx.x;
  const int j0 = cnode[A], d = cnode[A + 1] - j0;
  const int b0 = rptr[j0], len = rptr[j0 + 1] - b0;
  for (int k = threadIdx.x; k < len; k += blockDim.x) {
    const int s0 = src[b0 + k];
    for (int q = 0; q < d; ++q) out[rptr[j0 + q] + k] = in[s0 + q];
  }
}

// (C + C^T)/2 from the raw product (Multigrid.cpp:250-284): ia / ib index the
// raw values of (i, j) and (j, i), or -1 where absent.
__global__ void k_symmetrize(long long n, const int* __restrict__ ia, const int* __restrict__ ib,
                             const double* __restrict__ raw, double* __restrict__ out) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int a = ia[t], b = ib[t];
  double v;
  if (a >= 0 && b >= 0) v = 0.5 * (raw[a] + raw[b]);
  else if (a >= 0) v = 0.5 * raw[a];
  else v = 0.5 * raw[b];
  out[t] = v;
}

// inv_diag_i = 1 / the diagonal entry of row i (Multigrid.cpp:721-735); rows are
// sorted and hold each column once, so the entry is found by bisection.
__global__ void k_inv_diag(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                           const double* __restrict__ val, double* __restrict__ dinv, int* __restrict__ bad_row) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = ptr[i], hi = ptr[i + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] < i) lo = mid + 1; else hi = mid;
  }
  const double d = (lo < ptr[i + 1] && col[lo] == i) ? 0.0 + val[lo] : 0.0;
  if (!(d > 0.0)) atomicMin(bad_row, i);
  dinv[i] = 1.0 / d;
}

// ---------------------------------------------------------------------------
// Block products C = A B. Every operator of the hierarchy has rows grouped in
// nodes (the free DOFs of a mesh node, the coarse DOFs of an aggregate) whose
// rows share one column pattern, and the columns of the products fall in
// aggregate blocks (contiguous coarse columns). An item is one block of one
// C node-row; its contribution list, built once per sparsity structure, holds
// one entry per (A run, B block) pair that feeds it, in the order of A's row. A
// run is one B-node's stretch of A's row; the entry carries where it starts in
// A's row, how many B rows it spans, the offset of the item's aggregate block
// in the first of those B rows and the B row length (node-uniform rows), so a
// lane needs no further index lookups. One lane per block column t computes an
// item's column for all of the node's rows r: it sums A(row r, run) . B(run
// rows, t) over the list, in the order of A's row, the order of the CPU
// restatement (A's values are the same for the item's lanes: broadcast loads).
// EPI 1 forms P = (I - omega D^-1 A) P-hat with B = P-hat (Multigrid.cpp:686-701).
// ---------------------------------------------------------------------------
template <int EPI, int DMAX>
__global__ void __launch_bounds__(256) k_block_product(BlockProductArgs g) {
  const int lane = threadIdx.x & 63;
  const int gpw = kWave / g.wmax;  // items per wavefront: one lane per block column
  const int grp = lane / g.wmax, t = lane - grp * g.wmax;
  if (grp >= gpw) return;
  // consecutive items (one node-row's blocks, then its neighbours', which share A
  // and B rows) stay on one XCD and its L2: block b runs logical block
  // (b % 8) * (grid / 8) + b / 8 (the grid is a multiple of 8)
  const int blk = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  const long long wave = ((long long)blk * blockDim.x + threadIdx.x) >> 6;
  const long long item = wave * gpw + grp;
  if (item >= g.nitems) return;
  const int node = g.item_node[item];
  const int agg = g.item_agg[item];
  const int w = g.agg_off[agg + 1] - g.agg_off[agg];
  if (t >= w) return;
  const int r0 = g.out_node_ptr[node];
  const int d = min(g.out_node_ptr[node + 1] - r0, DMAX);
  const double* arow[DMAX];
  double acc[DMAX];
#pragma unroll
  for (int r = 0; r < DMAX; ++r) {
    arow[r] = g.aval + g.aptr[r0 + (r < d ? r : 0)];
    acc[r] = 0.0;
  }
  const long long e1 = g.item_begin[item + 1];
  for (long long e = g.item_begin[item]; e < e1; ++e) {
    const uint2 en = g.list[e];
    const int koff = (int)(en.y >> 16), lb = (int)((en.y >> 3) & 0x1fffu), dj = (int)(en.y & 7u);
    const double* b = g.bval + en.x + t;
    if (dj == 3) {
      const double b0 = b[0], b1 = b[lb], b2 = b[2 * lb];
#pragma unroll
      for (int r = 0; r < DMAX; ++r) {
        if (r >= d) break;
        const double* a = arow[r] + koff;
        acc[r] += a[0] * b0;
        acc[r] += a[1] * b1;
        acc[r] += a[2] * b2;
      }
    } else {
      for (int m = 0; m < dj; ++m) {
        const double bm = b[m * lb];
#pragma unroll
        for (int r = 0; r < DMAX; ++r)
          if (r < d) acc[r] += arow[r][koff + m] * bm;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < DMAX; ++r) {
    if (r >= d) break;
    const int row = r0 + r;
    double v = acc[r];
    if (EPI == 1) {
      v *= -(g.omega * g.dinv[row]);
      if (agg == g.row_agg[row]) v += g.tval[g.tptr[row] + t];
    }
    g.cval[g.cptr[row] + g.item_pos[item] + t] = v;
  }
}

// ---------------------------------------------------------------------------
// Row products C = A B with one wavefront per C node (the C and A rows of the
// node share one pattern). The node's A row is walked run by run; a run is the
// stretch of A's row that meets one B node, and its values are the same for
// every lane (scalar loads). For each run the lanes take the B node's blocks,
// one lane per block column, read the run's B rows with coalesced loads and add
// A(run) . B(run rows, column) into the node's accumulators in LDS at the
// block's position in C's row. Within one run no two lanes share an
// accumulator, so every entry of C sums its products in the order of A's row,
// the order of the CPU restatement (Multigrid.cpp:199-225), with fused
// multiply-adds (rounding differs from the restatement's separate products and
// sums). The next step's B values are loaded while the current step is added.
// With PRE the node's entries, with SA the node's A rows, are staged in LDS
// first (SA: A's values come from LDS instead of scalar loads).
// ---------------------------------------------------------------------------
struct Step {
  int t, eb;  // run, first entry of the pass
  uint4 rd;   // the run's descriptor (entries [rd.z, rd.w))
};

template <int DMAX, int JMAX, bool PRE, bool SA>
__global__ void __launch_bounds__(256) k_row_product(
    int nnodes, int wpb, int lenc_max, int emax, int lena_max, int wmax, const int* __restrict__ order,
    const int* __restrict__ node_ptr, const int* __restrict__ aptr, const double* __restrict__ aval,
    const int* __restrict__ run_ptr, const uint4* __restrict__ run,
    const unsigned* __restrict__ entry, const double* __restrict__ bval, const int* __restrict__ cptr,
    double* __restrict__ cval) {
  extern __shared__ double lds[];
  // the wavefront's index is uniform: its LDS bases stay in scalar registers
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // consecutive nodes (Morton order) stay on one XCD and its L2: block b runs
  // logical block (b % 8) * (grid / 8) + b / 8 (the grid is a multiple of 8)
  const int blk = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  const int slot = blk * wpb + wid;
  if (slot >= nnodes) return;
  double* acc = lds + (size_t)wid * ((size_t)DMAX * lenc_max + (SA ? (size_t)DMAX * lena_max : 0) +
                                     (PRE ? (size_t)(emax + 1) / 2 : 0));
  double* sa = acc + (size_t)DMAX * lenc_max;
  unsigned* ent = reinterpret_cast<unsigned*>(sa + (SA ? (size_t)DMAX * lena_max : 0));
  const int I = __builtin_amdgcn_readfirstlane(order[slot]);
  const int r0 = node_ptr[I];
  const int d = min(node_ptr[I + 1] - r0, DMAX);
  const int c0 = cptr[r0], lenc = cptr[r0 + 1] - c0;
  for (int p = lane; p < d * lenc; p += 64) acc[p] = 0.0;
  int abase[DMAX];
#pragma unroll
  for (int r = 0; r < DMAX; ++r) abase[r] = aptr[r0 + (r < d ? r : d - 1)];
  const int t0 = run_ptr[I], t1 = run_ptr[I + 1];
  if (t0 < t1) {
    const int E0 = (int)run[t0].z;
    const int lena = aptr[r0 + 1] - aptr[r0];
    if (SA)  // the node's rows are consecutive in A's values
      for (int p = lane; p < d * lena; p += 64) sa[p] = aval[abase[0] + p];
    if (PRE) {
      const int E1 = (int)run[t1 - 1].w;
      for (int e = E0 + lane; e < E1; e += 64) ent[e - E0] = entry[e];
    }
    const int per = 64 / wmax;
    const int ul = lane / wmax, tt = lane - ul * wmax;
    // one step = one pass of the lanes over (up to per) blocks of run t from
    // entry eb; fetch() gives the lane's C position (or -1) and its B values
    auto fetch = [&](const uint4& rd, int eb, double(&bv)[JMAX]) -> int {
      const int lb = (int)((rd.y >> 3) & 0x1fffu), dj = (int)(rd.y & 7u);
      const int e1 = (int)rd.w;
      const int e = min(eb + ul, e1 - 1);
      const unsigned en = PRE ? ent[e - E0] : entry[e];
      const int w = (int)(en & 7u), bpos = (int)((en >> 3) & 0x1fffu), cpos = (int)(en >> 16);
      const bool live = ul < per && eb + ul < e1 && tt < w;
      const unsigned idx = rd.x + (unsigned)(bpos + min(tt, w - 1));
#pragma unroll
      for (int m = 0; m < JMAX; ++m) bv[m] = bval[idx + (unsigned)(min(m, dj - 1) * lb)];
      return live ? cpos + tt : -1;
    };
    // one wide scalar load per row: JMAX values from the run's start (the
    // value arrays carry 8 doubles of padding past the last row)
    auto load_a = [&](const uint4& rd, double(&av)[DMAX][JMAX]) {
      const int koff = (int)(rd.y >> 16);
      if (SA) {
        const int dj = (int)(rd.y & 7u);
#pragma unroll
        for (int r = 0; r < DMAX; ++r)
#pragma unroll
          for (int m = 0; m < JMAX; ++m) av[r][m] = sa[min(r, d - 1) * lena + koff + min(m, dj - 1)];
        return;
      }
      struct Run {
        double v[JMAX == 3 ? 4 : 8];
      };
#pragma unroll
      for (int r = 0; r < DMAX; ++r) {
        Run t;
        __builtin_memcpy(&t, aval + abase[r] + koff, sizeof(Run));
#pragma unroll
        for (int m = 0; m < JMAX; ++m) av[r][m] = t.v[m];
      }
    };
    // the step after st (st.t == t1: past the last step, valid addresses, no lane live)
    auto next = [&](const Step& st) -> Step {
      Step n = st;
      n.eb = st.eb + per;
      if (n.eb >= (int)st.rd.w) {
        n.t = st.t + 1;
        n.eb = (int)st.rd.w;
        if (n.t < t1) n.rd = run[n.t]; else n.eb = (int)st.rd.w - 1;
      }
      return n;
    };
    // v += a b in A's order with fused multiply-adds; a run spanning JMAX B rows
    // (the common case) takes a branch-free path
    auto add = [&](const Step& st, int pos, const double(&bv)[JMAX], const double(&av)[DMAX][JMAX]) {
#pragma clang fp contract(fast)
      const int dj = (int)(st.rd.y & 7u);
      if (pos < 0) return;
      if (dj == JMAX) {
#pragma unroll
        for (int r = 0; r < DMAX; ++r) {
          if (r >= d) break;
          double v = acc[r * lenc + pos];
#pragma unroll
          for (int m = 0; m < JMAX; ++m) v += av[r][m] * bv[m];
          acc[r * lenc + pos] = v;
        }
        return;
      }
#pragma unroll
      for (int r = 0; r < DMAX; ++r) {
        if (r >= d) break;
        double v = acc[r * lenc + pos];
#pragma unroll
        for (int m = 0; m < JMAX; ++m)
          if (m < dj) v += av[r][m] * bv[m];
        acc[r * lenc + pos] = v;
      }
    };
    // two register sets in turn: the B values of the next step load while
    // the current step is added
    Step c{t0, E0, run[t0]};
    double b0[JMAX], b1[JMAX], a[DMAX][JMAX];
    int p0 = fetch(c.rd, c.eb, b0);
    load_a(c.rd, a);
    while (true) {
      const Step n = next(c);
      int p1 = fetch(n.rd, n.eb, b1);
      if (n.t >= t1) p1 = -1;
      add(c, p0, b0, a);
      if (n.t >= t1) break;
      if (n.t != c.t) load_a(n.rd, a);
      const Step m = next(n);
      p0 = fetch(m.rd, m.eb, b0);
      if (m.t >= t1) p0 = -1;
      add(n, p1, b1, a);
      if (m.t >= t1) break;
      if (m.t != n.t) load_a(m.rd, a);
      c = m;
    }
  }
  for (int p = lane; p < d * lenc; p += 64) cval[c0 + p] = acc[p];
}

// Lanczos on D^-1/2 A D^-1/2 (Multigrid.cpp:304-317). The step's scalars stay
// in device memory, so a level's steps run without host round trips:
// w_i = s_i (A (s v))_i - beta vprev_i, with s = sqrt(dinv), beta = *beta_ptr (0 if null)
// ... with the block partials of w.v (rows in a grid-stride loop; partials[block])
// cbase (may be null): rows of one node share one column list, entry k of row i
// reads its column at col[cbase[i] + k]. Each lane keeps four entries' loads in
// flight and adds them in its own order (k, k + 16, ...), as a one-at-a-time loop
// would.
__global__ void __launch_bounds__(256) k_lanczos_apply(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                                                       const int* __restrict__ cbase,
                                                       const double* __restrict__ val, const double* __restrict__ s,
                                                       const double* __restrict__ v, const double* __restrict__ vprev,
                                                       const double* __restrict__ beta_ptr, double* __restrict__ w,
                                                       double* __restrict__ partials) {
  constexpr int LANES = 16;
  __shared__ double sh[4];
  const int lane = threadIdx.x % LANES;
  const double beta = beta_ptr ? *beta_ptr : 0.0;
  double dot = 0.0;
  for (int row = (blockIdx.x * blockDim.x + threadIdx.x) / LANES; row < n; row += gridDim.x * (blockDim.x / LANES)) {
    const int* colr = cbase ? col + cbase[row] : col;
    const int e = ptr[row + 1];
    double acc = 0.0;
    int k = ptr[row] + lane;
    for (; k + 3 * LANES < e; k += 4 * LANES) {
      const int j0 = colr[k], j1 = colr[k + LANES], j2 = colr[k + 2 * LANES], j3 = colr[k + 3 * LANES];
      const double a0 = val[k], a1 = val[k + LANES], a2 = val[k + 2 * LANES], a3 = val[k + 3 * LANES];
      const double u0 = s[j0] * v[j0], u1 = s[j1] * v[j1], u2 = s[j2] * v[j2], u3 = s[j3] * v[j3];
      acc += a0 * u0;
      acc += a1 * u1;
      acc += a2 * u2;
      acc += a3 * u3;
    }
    for (; k < e; k += LANES) {
      const int j = colr[k];
      acc += val[k] * (s[j] * v[j]);
    }
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
    if (lane == 0) {
      const double wi = acc * s[row] - beta * vprev[row];
      w[row] = wi;
      dot += wi * v[row];
    }
  }
  for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = dot;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// w -= a v with a = *a_ptr, and the block partials of w.w (grid-stride)
// the sum of m partials in a fixed order, the same in every block of 256
__device__ __forceinline__ double sum_partials_256(int m, const double* __restrict__ p, double* sh) {
  double t = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) t += p[i];
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  const double r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

// alpha = the sum of the product's partials (block 0 stores it); w -= alpha v
// with the partials of w.w
__global__ void __launch_bounds__(256) k_axmy_dot_sum(int n, int m, const double* __restrict__ part_in,
                                                      double* __restrict__ alpha_out, const double* __restrict__ v,
                                                      double* __restrict__ w, double* __restrict__ part_out) {
  __shared__ double sh[4];
  const double a = sum_partials_256(m, part_in, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) *alpha_out = a;
  double dot = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double wi = w[i] - a * v[i];
    w[i] = wi;
    dot += wi * wi;
  }
  for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = dot;
  __syncthreads();
  if (threadIdx.x == 0) part_out[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// beta = sqrt(the sum of the partials) (block 0 stores it); vprev = v, v = w / beta
__global__ void __launch_bounds__(256) k_lanczos_shift_sum(int n, int m, const double* __restrict__ part_in,
                                                           double* __restrict__ beta_out,
                                                           const double* __restrict__ w, double* __restrict__ v,
                                                           double* __restrict__ vprev) {
  __shared__ double sh[4];
  const double b = sqrt(sum_partials_256(m, part_in, sh));
  if (blockIdx.x == 0 && threadIdx.x == 0) *beta_out = b;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    vprev[i] = v[i];
    v[i] = w[i] / b;
  }
}

__global__ void k_sqrt(int n, const double* __restrict__ in, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sqrt(in[i]);
}

// vprev = v; v = w / b with b = *b_ptr
__global__ void k_lanczos_shift(int n, const double* __restrict__ b_ptr, const double* __restrict__ w,
                                double* __restrict__ v, double* __restrict__ vprev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vprev[i] = v[i];
  v[i] = w[i] / *b_ptr;
}

__global__ void k_sqrt_scalar(double* p) { p[0] = sqrt(p[0]); }

__global__ void k_fill(long long n, double v, double* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

__global__ void k_copy(long long n, const double* __restrict__ in, double* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// ---- tentative-prolongator transfers ----------------------------------------
// P-hat is block diagonal over the aggregates: fine row r holds the len_r <= 6
// consecutive columns c0_r.. of its aggregate. Rows store c0 * 8 + len and the
// values column-major (ell[k * n + r]), so a thread per row reads coalesced.
// Two rows per thread, h apart (h = half the rows, rounded up), so that every
// load instruction stays coalesced; both rows' loads issue before either sum.
template <typename T>
__global__ void __launch_bounds__(256) k_tent_prolong(int n, int h, const int* __restrict__ c0len,
                                                      const T* __restrict__ ell, const T* __restrict__ xc,
                                                      T* __restrict__ y) {
  const int r0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (r0 >= h) return;
  const int r1 = r0 + h;
  const bool two = r1 < n;
  const int e0 = c0len[r0], e1 = two ? c0len[r1] : 0;
  const int c00 = e0 >> 3, len0 = e0 & 7, c01 = e1 >> 3, len1 = e1 & 7;
  T a0[6], a1[6], x0[6], x1[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    a0[k] = k < len0 ? ell[(size_t)k * n + r0] : T(0);
    x0[k] = k < len0 ? xc[c00 + k] : T(0);
    a1[k] = k < len1 ? ell[(size_t)k * n + r1] : T(0);
    x1[k] = k < len1 ? xc[c01 + k] : T(0);
  }
  T acc0 = 0, acc1 = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (k < len0) acc0 += a0[k] * x0[k];
    if (k < len1) acc1 += a1[k] * x1[k];
  }
  y[r0] = acc0;
  if (two) y[r1] = acc1;
}

// xc = P-hat^T y: 32 lanes per aggregate over its fine rows, then a reduction
// across the lanes for each of its columns. ella holds the values in the
// aggregates' row order (ella[k * n + i] for member i), so the lanes read them
// contiguously and only y is gathered.
// With start.dinv set, the coarse level's zero-guess smoother start follows
// in the epilogue: d = dinv xc / theta, x2 = d (k_cheb_first_zero's rounding).
template <typename T>
struct TentStart {
  const T* dinv = nullptr;
  T* d = nullptr;
  T* x2 = nullptr;
  double theta = 1.0;
};
template <typename T>
__global__ void __launch_bounds__(256) k_tent_restrict(int naggs, const int* __restrict__ agg_ptr,
                                                       const int* __restrict__ agg_rows,
                                                       const int* __restrict__ agg_c0len, int n,
                                                       const T* __restrict__ ella, const T* __restrict__ y,
                                                       T* __restrict__ xc, TentStart<T> st) {
  const int a = (blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int lane = threadIdx.x & 31;
  if (a >= naggs) return;  // whole 32-lane groups leave together
  const int e = agg_c0len[a], c0 = e >> 3, len = e & 7;
  T acc[6] = {0, 0, 0, 0, 0, 0};
  // four of the lane's rows at a time: their indices, then y and the values
  // load before any is added; the lane still adds its rows in ascending order
  const int i1 = agg_ptr[a + 1];
  for (int i0 = agg_ptr[a] + lane; i0 < i1; i0 += 4 * 32) {
    int rows[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) rows[u] = i0 + 32 * u < i1 ? agg_rows[i0 + 32 * u] : 0;
    T yr[4], ev[4][6];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = i0 + 32 * u < i1;
      yr[u] = in ? y[rows[u]] : T(0);
#pragma unroll
      for (int k = 0; k < 6; ++k) ev[u][k] = (in && k < len) ? ella[(size_t)k * n + i0 + 32 * u] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (i0 + 32 * u >= i1) break;
#pragma unroll
      for (int k = 0; k < 6; ++k)
        if (k < len) acc[k] += ev[u][k] * yr[u];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, 32);
  }
  if (lane < len) {
    T v = acc[0];
#pragma unroll
    for (int k = 1; k < 6; ++k)
      if (lane == k) v = acc[k];
    xc[c0 + lane] = v;
    if (st.dinv) {
      const double di = (double)st.dinv[c0 + lane] * (double)v / st.theta;
      st.d[c0 + lane] = (T)di;
      st.x2[c0 + lane] = (T)(0.0 + di);
    }
  }
}

// ---- dense inverse of the coarsest operator --------------------------------
// Blocked in-place Gauss-Jordan on the row-major n x n matrix, one pivot block
// K of kGjB rows at a time (the sweep of K):
//   P = A_KK^-1,  A_KJ <- P A_KJ,  A_IJ <- A_IJ - A_IK A_KJ,  A_IK <- -A_IK P,  A_KK <- P
// for I, J outside K. After the last block A holds A^-1. The scalar pivots met
// inside the diagonal blocks are the D of the unpivoted LDL^T, which the caller
// checks for a singular coarse operator (Multigrid.cpp:766-781). The operator
// is symmetric positive definite, so no pivoting is needed.
constexpr int kGjB = 32;   // pivot block
constexpr int kGjT = 64;   // update tile

// row i of the dense matrix: zero it, then scatter the CSR row
__global__ void __launch_bounds__(256) k_dense_from_csr(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                                                       const double* __restrict__ val, double* __restrict__ a) {
  const int i = blockIdx.x;
  double* row = a + (size_t)i * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) row[j] = 0.0;
  __syncthreads();
  for (int k = ptr[i] + threadIdx.x; k < ptr[i + 1]; k += blockDim.x) row[col[k]] = val[k];
}

// Every block inverts A_KK in LDS (redundantly), then updates the row panel
// A_KJ for its kGjT columns and saves the old column panel A_IK of its kGjT
// rows into cbuf[n][kGjB]. Block 0 also writes P and the pivots.
__global__ void __launch_bounds__(256) k_gj_panel(int n, int k0, double* __restrict__ a, double* __restrict__ pbuf,
                                                 double* __restrict__ cbuf, double* __restrict__ piv) {
  __shared__ double P[kGjB][kGjB + 1];
  __shared__ double R[kGjB][kGjT + 1];
  const int m = min(kGjB, n - k0);
  const int tid = threadIdx.x;
  for (int e = tid; e < kGjB * kGjB; e += 256) {
    const int r = e / kGjB, q = e % kGjB;
    P[r][q] = (r < m && q < m) ? a[(size_t)(k0 + r) * n + k0 + q] : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    const double pk = P[k][k];
    const double p = 1.0 / pk;
    double nv[(kGjB * kGjB) / 256];
#pragma unroll
    for (int t = 0; t < (kGjB * kGjB) / 256; ++t) {
      const int e = tid + 256 * t, r = e / kGjB, q = e % kGjB;
      double v = P[r][q];
      if (r < m && q < m) {
        if (r == k) v = q == k ? p : v * p;
        else if (q == k) v = -P[r][k] * p;
        else v = v - P[r][k] * (P[k][q] * p);
      }
      nv[t] = v;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < (kGjB * kGjB) / 256; ++t) {
      const int e = tid + 256 * t;
      P[e / kGjB][e % kGjB] = nv[t];
    }
    if (blockIdx.x == 0 && tid == 0) piv[k0 + k] = pk;
    __syncthreads();
  }
  if (blockIdx.x == 0)
    for (int e = tid; e < kGjB * kGjB; e += 256) pbuf[e] = P[e / kGjB][e % kGjB];
  // row panel: A_KJ <- P A_KJ over this block's columns outside K
  const int j0 = blockIdx.x * kGjT;
  for (int e = tid; e < kGjB * kGjT; e += 256) {
    const int t = e / kGjT, q = e % kGjT, j = j0 + q;
    R[t][q] = (t < m && j < n) ? a[(size_t)(k0 + t) * n + j] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < kGjB * kGjT; e += 256) {
    const int r = e / kGjT, q = e % kGjT, j = j0 + q;
    if (r >= m || j >= n || (j >= k0 && j < k0 + m)) continue;
    double s = 0.0;
    for (int t = 0; t < m; ++t) s = fma(P[r][t], R[t][q], s);
    a[(size_t)(k0 + r) * n + j] = s;
  }
  // old column panel of this block's rows (zero inside K)
  for (int e = tid; e < kGjT * kGjB; e += 256) {
    const int r = e / kGjB, t = e % kGjB, i = j0 + r;
    if (i >= n) continue;
    const bool inK = i >= k0 && i < k0 + m;
    cbuf[(size_t)i * kGjB + t] = (!inK && t < m) ? a[(size_t)i * n + k0 + t] : 0.0;
  }
}

// A_IJ <- A_IJ - C_I R_J with C the saved column panel and R the new row panel
// (P in the columns of K, which gives A_IK <- -A_IK P); A_KK <- P.
__global__ void __launch_bounds__(256) k_gj_update(int n, int k0, double* __restrict__ a,
                                                  const double* __restrict__ pbuf, const double* __restrict__ cbuf) {
  __shared__ double Cs[kGjT][kGjB + 1];
  __shared__ double Rs[kGjB][kGjT + 1];
  const int m = min(kGjB, n - k0);
  const int i0 = blockIdx.y * kGjT, j0 = blockIdx.x * kGjT;
  const int tid = threadIdx.x;
  for (int e = tid; e < kGjT * kGjB; e += 256) {
    const int r = e / kGjB, t = e % kGjB, i = i0 + r;
    Cs[r][t] = (i < n && t < m) ? cbuf[(size_t)i * kGjB + t] : 0.0;
  }
  for (int e = tid; e < kGjB * kGjT; e += 256) {
    const int t = e / kGjT, q = e % kGjT, j = j0 + q;
    double v = 0.0;
    if (t < m && j < n) v = (j >= k0 && j < k0 + m) ? pbuf[t * kGjB + (j - k0)] : a[(size_t)(k0 + t) * n + j];
    Rs[t][q] = v;
  }
  __syncthreads();
  const int tx = tid % 16, ty = tid / 16;
  double acc[4][4] = {};
  for (int t = 0; t < m; ++t) {
    double cv[4], rv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cv[u] = Cs[ty + 16 * u][t];
      rv[u] = Rs[t][tx + 16 * u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = fma(cv[u], rv[v], acc[u][v]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int i = i0 + ty + 16 * u;
    if (i >= n) continue;
    const bool iK = i >= k0 && i < k0 + m;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = j0 + tx + 16 * v;
      if (j >= n) continue;
      const bool jK = j >= k0 && j < k0 + m;
      double* x = a + (size_t)i * n + j;
      if (iK) {
        if (jK) *x = pbuf[(i - k0) * kGjB + (j - k0)];
      } else {
        *x = jK ? -acc[u][v] : *x - acc[u][v];
      }
    }
  }
}

// Node-block sums of squares for the strength of connection (Multigrid.cpp:
// 457-490): for node I (rows [node_ptr[I], node_ptr[I+1]), one shared pattern)
// and each run of its first row (the columns of one node J), the sum over the
// node's rows in order, and over the run's entries in column order, of v^2 --
// the order of the host's accumulation, so the sums are the same bits.
__global__ void k_node_block_sums(int nn, const int* __restrict__ node_ptr, const int* __restrict__ node_of,
                                  const int* __restrict__ ptr, const int* __restrict__ col,
                                  const double* __restrict__ val, const int* __restrict__ run_ptr,
                                  double* __restrict__ sums) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nn) return;
  const int r0 = node_ptr[I], dn = node_ptr[I + 1] - r0;
  const int k0 = ptr[r0], k1 = ptr[r0 + 1];
  int t = run_ptr[I];
  for (int k = k0; k < k1;) {
    const int J = node_of[col[k]];
    const int dj = node_ptr[J + 1] - node_ptr[J];
    const int koff = k - k0;
    double acc = 0.0;
    for (int i = 0; i < dn; ++i) {
      const double* v = val + ptr[r0 + i] + koff;
      for (int q = 0; q < dj; ++q) acc += v[q] * v[q];
    }
    sums[t++] = acc;
    k += dj;
  }
}

// max |v| over the values (block maxima, then one block)
__global__ void __launch_bounds__(256) k_absmax(long long n, const double* __restrict__ v, double* __restrict__ out) {
  __shared__ double sh[256];
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    m = fmax(m, fabs(v[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Symmetry of a CSR matrix with sorted rows (Multigrid.cpp:968-991): entry
// (i, j) against (j, i), found by bisection in row j (0 when absent); a
// difference above tol * max|v| lowers *bad to (i << 32) | j.
__global__ void k_symmetry(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                           const double* __restrict__ val, const double* __restrict__ maxabs, double tol,
                           unsigned long long* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lim = tol * maxabs[0];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const int j = col[k];
    int lo = ptr[j], hi = ptr[j + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (col[mid] < i) lo = mid + 1; else hi = mid;
    }
    const double other = (lo < ptr[j + 1] && col[lo] == i) ? val[lo] : 0.0;
    if (fabs(other - val[k]) > lim) {
      atomicMin(bad, ((unsigned long long)i << 32) | (unsigned)j);
      return;
    }
  }
}

}  // namespace

void node_block_sums(int nn, const int* node_ptr, const int* node_of, const int* ptr, const int* col,
                     const double* val, const int* run_ptr, double* sums, hipStream_t s) {
  if (nn) k_node_block_sums<<<nblk(nn, 256), 256, 0, s>>>(nn, node_ptr, node_of, ptr, col, val, run_ptr, sums);
}

void symmetry_check(const DCsr& A, double tol, double* scratch, unsigned long long* bad, hipStream_t s) {
  // scratch: 1025 doubles (block maxima, then the maximum)
  const int nb = std::max(1, std::min(1024, nblk(A.nnz, 256)));
  k_absmax<<<nb, 256, 0, s>>>(A.nnz, A.val.p, scratch + 1);
  k_absmax<<<1, 256, 0, s>>>(nb, scratch + 1, scratch);
  if (A.rows) k_symmetry<<<nblk(A.rows, 256), 256, 0, s>>>(A.rows, A.ptr.p, A.col.p, A.val.p, scratch, tol, bad);
}

void block_product(const BlockProductArgs& a, bool smooth, hipStream_t s) {
  if (a.nitems == 0) return;
  if (a.wmax < 1 || a.wmax > kWave || a.dmax < 1 || a.dmax > 6)
    fail(SLX_SOLVER_ERROR, "multigrid: block product shape outside the device kernel's range");
  const long long waves = (a.nitems + kWave / a.wmax - 1) / (kWave / a.wmax);
  const int grid = (int)((waves + 3) / 4 + 7) / 8 * 8;
  if (!smooth) fail(SLX_SOLVER_ERROR, "multigrid: plain block products run as row products");
  if (a.dmax <= 3) k_block_product<1, 3><<<grid, 256, 0, s>>>(a); else k_block_product<1, 6><<<grid, 256, 0, s>>>(a);
}

void row_product(const RowProductArgs& a, hipStream_t s) {
  if (a.nnodes == 0) return;
  if (a.wmax < 1 || a.wmax > kWave || a.dmax < 1 || a.dmax > 6 || a.jmax < 1 || a.jmax > 6 || a.wpb < 1 || a.wpb > 4)
    fail(SLX_SOLVER_ERROR, "multigrid: row product shape outside the device kernel's range");
  const int dm = a.dmax <= 3 && a.jmax <= 3 ? 3 : 6;
  const int jm = a.jmax <= 3 ? 3 : 6;
  if (a.sa && dm != 3) fail(SLX_SOLVER_ERROR, "multigrid: row product A staging needs at most 3 rows per node");
  const size_t lds = row_product_lds(dm, a.lenc_max, a.pre ? a.emax : 0, a.sa ? a.lena_max : 0) * (size_t)a.wpb;
  if (lds > 64 * 1024) fail(SLX_SOLVER_ERROR, "multigrid: a row product row is too long for the device kernel");
  const int grid = (int)(((long long)a.nnodes + a.wpb - 1) / a.wpb + 7) / 8 * 8;
#define SLX_ROW_PRODUCT(D, J, P, S)                                                                               \
  k_row_product<D, J, P, S><<<grid, 64 * a.wpb, lds, s>>>(a.nnodes, a.wpb, a.lenc_max, a.emax, a.lena_max, a.wmax, \
                                                          a.order,                                                 \
                                                       a.node_ptr, a.aptr, a.aval, a.run_ptr, a.run,              \
                                                       a.entry, a.bval, a.cptr, a.cval)
  if (dm == 3) {
    if (a.sa) {
      if (a.pre) SLX_ROW_PRODUCT(3, 3, true, true); else SLX_ROW_PRODUCT(3, 3, false, true);
    } else {
      if (a.pre) SLX_ROW_PRODUCT(3, 3, true, false); else SLX_ROW_PRODUCT(3, 3, false, false);
    }
  } else if (jm == 3) {
    if (a.pre) SLX_ROW_PRODUCT(6, 3, true, false); else SLX_ROW_PRODUCT(6, 3, false, false);
  } else {
    if (a.pre) SLX_ROW_PRODUCT(6, 6, true, false); else SLX_ROW_PRODUCT(6, 6, false, false);
  }
#undef SLX_ROW_PRODUCT
}

void spgemm_numeric(const DCsr& A, const DCsr& B, DCsr& C, int* error_flag, hipStream_t s) {
  if (C.rows == 0) return;
  const int cap = std::max(C.max_row, 1);
  const size_t lds = (size_t)cap * (sizeof(double) + sizeof(int));
  if (lds > 60 * 1024) fail(SLX_SOLVER_ERROR, "multigrid: a Galerkin product row is too long for the device kernel");
  const int grid = std::min(C.rows, 256 * 64);
  k_spgemm_numeric<0><<<grid, kWave, lds, s>>>(C.rows, A.ptr.p, A.col.p, A.val.p, B.ptr.p, B.col.p, B.val.p, C.ptr.p,
                                               C.col.p, C.val.p, cap, nullptr, 0.0, error_flag);
}

void smooth_prolongator(const DCsr& A, const DCsr& Ptent, DCsr& P, const double* dinv, double omega, int* error_flag,
                        hipStream_t s) {
  if (P.rows == 0) return;
  const int cap = std::max(P.max_row, 1);
  const size_t lds = (size_t)cap * (sizeof(double) + sizeof(int));
  if (lds > 60 * 1024) fail(SLX_SOLVER_ERROR, "multigrid: a prolongator row is too long for the device kernel");
  const int grid = std::min(P.rows, 256 * 64);
  k_spgemm_numeric<1><<<grid, kWave, lds, s>>>(P.rows, A.ptr.p, A.col.p, A.val.p, Ptent.ptr.p, Ptent.col.p,
                                               Ptent.val.p, P.ptr.p, P.col.p, P.val.p, cap, dinv, omega, error_flag);
}

void gather(long long n, const int* src, const double* in, double* out, hipStream_t s) {
  if (n) k_gather<<<nblk(n, 256), 256, 0, s>>>(n, src, in, out);
}

void gather_nodes(int nnodes, const int* cnode, const int* rptr, const int* src, const double* in, double* out,
                  hipStream_t s) {
  if (nnodes > 0) k_gather_nodes<<<nnodes, 128, 0, s>>>(cnode, rptr, src, in, out);
}

void symmetrize_values(long long n, const int* ia, const int* ib, const double* raw, double* out, hipStream_t s) {
  if (n) k_symmetrize<<<nblk(n, 256), 256, 0, s>>>(n, ia, ib, raw, out);
}

void inv_diag(const DCsr& A, double* dinv, int* bad_row, hipStream_t s) {
  if (A.rows) k_inv_diag<<<nblk(A.rows, 256), 256, 0, s>>>(A.rows, A.ptr.p, A.col.p, A.val.p, dinv, bad_row);
}

void vsqrt(int n, const double* in, double* out, hipStream_t s) {
  if (n) k_sqrt<<<nblk(n, 256), 256, 0, s>>>(n, in, out);
}

int lanczos_blocks(int n) { return std::max(1, std::min(1024, nblk((long long)n * 16, 256))); }

void lanczos_apply(const DCsr& A, const double* sq, const double* v, const double* vprev, const double* beta,
                   double* w, double* partials, hipStream_t s) {
  if (A.rows)
    k_lanczos_apply<<<lanczos_blocks(A.rows), 256, 0, s>>>(A.rows, A.ptr.p, A.cbase.p ? A.ncol.p : A.col.p,
                                                           A.cbase.p, A.val.p, sq, v, vprev, beta, w, partials);
}

int axmy_blocks(int n) { return std::max(1, std::min(1024, nblk(n, 256))); }


void axmy_dot_sum(int n, int m, const double* part_in, double* alpha, const double* v, double* w, double* part_out,
                  hipStream_t s) {
  if (n) k_axmy_dot_sum<<<axmy_blocks(n), 256, 0, s>>>(n, m, part_in, alpha, v, w, part_out);
}

void lanczos_shift_sum(int n, int m, const double* part_in, double* beta, const double* w, double* v, double* vprev,
                       hipStream_t s) {
  if (n) k_lanczos_shift_sum<<<axmy_blocks(n), 256, 0, s>>>(n, m, part_in, beta, w, v, vprev);
}

void lanczos_shift(int n, const double* b, const double* w, double* v, double* vprev, hipStream_t s) {
  if (n) k_lanczos_shift<<<nblk(n, 256), 256, 0, s>>>(n, b, w, v, vprev);
}

void sqrt_scalar(double* p, hipStream_t s) { k_sqrt_scalar<<<1, 1, 0, s>>>(p); }

void fill(long long n, double v, double* out, hipStream_t s) {
  if (n) k_fill<<<nblk(n, 256), 256, 0, s>>>(n, v, out);
}

void copy(long long n, const double* in, double* out, hipStream_t s) {
  if (n) k_copy<<<nblk(n, 256), 256, 0, s>>>(n, in, out);
}

int dense_inverse_work(int n) { return kGjB * kGjB + n * kGjB; }

void dense_inverse(int n, const int* ptr, const int* col, const double* val, double* a, double* piv, double* work,
                   hipStream_t s) {
  if (n <= 0) return;
  k_dense_from_csr<<<n, 256, 0, s>>>(n, ptr, col, val, a);
  double* pbuf = work;
  double* cbuf = work + kGjB * kGjB;
  const int tiles = (n + kGjT - 1) / kGjT;
  for (int k0 = 0; k0 < n; k0 += kGjB) {
    k_gj_panel<<<tiles, 256, 0, s>>>(n, k0, a, pbuf, cbuf, piv);
    k_gj_update<<<dim3(tiles, tiles), 256, 0, s>>>(n, k0, a, pbuf, cbuf);
  }
}

template <typename T>
void tent_prolong(const TentBlocks& t, const T* ell, const T* xc, T* y, hipStream_t s) {
  const int h = (t.n + 1) / 2;
  if (t.n) k_tent_prolong<T><<<nblk(h, 256), 256, 0, s>>>(t.n, h, t.c0len, ell, xc, y);
}
template <typename T>
void tent_restrict(const TentBlocks& t, const T* ella, const T* y, T* xc, hipStream_t s, const T* dinv, double theta,
                   T* d, T* x2) {
  TentStart<T> st;
  st.dinv = dinv;
  st.d = d;
  st.x2 = x2;
  st.theta = theta;
  if (t.naggs) k_tent_restrict<T><<<nblk((long long)t.naggs * 32, 256), 256, 0, s>>>(t.naggs, t.agg_ptr, t.agg_rows,
                                                                                    t.agg_c0len, t.n, ella, y, xc, st);
}
template void tent_prolong<double>(const TentBlocks&, const double*, const double*, double*, hipStream_t);
template void tent_prolong<float>(const TentBlocks&, const float*, const float*, float*, hipStream_t);
template void tent_restrict<double>(const TentBlocks&, const double*, const double*, double*, hipStream_t,
                                    const double*, double, double*, double*);
template void tent_restrict<float>(const TentBlocks&, const float*, const float*, float*, hipStream_t, const float*,
                                   double, float*, float*);

}  // namespace k
}  // namespace slx
