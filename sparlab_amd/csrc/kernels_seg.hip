// Kernels of the segmented (batched) CG: independent systems on the diagonal
// blocks of one matrix, one launch for all of them. Rows are cut into chunks
// that never cross a segment boundary; chunk c covers rows
// [chunk_begin[c], chunk_end[c]) of segment chunk_seg[c], and the chunks of
// segment s are [seg_chunk[s], seg_chunk[s+1]). Per-segment scalars live in
// device arrays of length nseg; a segment whose flag active[s] is 0 keeps its
// iterate and residual (its step length is zero).
#include <hip/hip_runtime.h>

#include "slx_seg.hpp"

namespace slx {
namespace kseg {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w];
  return s;
}

// partials[c] = sum over chunk c of a.b
__global__ void __launch_bounds__(kBlock) k_chunk_dot(Chunks ch, const double* __restrict__ a,
                                                       const double* __restrict__ b, double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int i = ch.begin[c] + threadIdx.x; i < ch.end[c]; i += kBlock) s += a[i] * b[i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) partials[c] = t;
}

// out[s] = sum of the partials of segment s's chunks, in chunk order
__global__ void k_seg_reduce(Chunks ch, int nseg, const double* __restrict__ partials, double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  double t = 0.0;
  for (int c = ch.seg_chunk[s]; c < ch.seg_chunk[s + 1]; ++c) t += partials[c];
  out[s] = t;
}

// active segments: alpha_s = rz_s / pq_s, x += alpha p, r -= alpha q; partials of r.r
__global__ void __launch_bounds__(kBlock) k_seg_update(Chunks ch, double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        const double* __restrict__ rz, const double* __restrict__ pq,
                                                        const int* __restrict__ active,
                                                        double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  const int c = blockIdx.x;
  const int sg = ch.seg[c];
  double s = 0.0;
  if (active[sg]) {
    const double alpha = rz[sg] / pq[sg];
    for (int i = ch.begin[c] + threadIdx.x; i < ch.end[c]; i += kBlock) {
      x[i] += alpha * p[i];
      const double ri = r[i] - alpha * q[i];
      r[i] = ri;
      s += ri * ri;
    }
  } else {
    for (int i = ch.begin[c] + threadIdx.x; i < ch.end[c]; i += kBlock) s += r[i] * r[i];
  }
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) partials[c] = t;
}

// p = z + beta_s p
__global__ void __launch_bounds__(kBlock) k_seg_direction(Chunks ch, double* __restrict__ p,
                                                           const double* __restrict__ z,
                                                           const double* __restrict__ beta) {
  const int c = blockIdx.x;
  const double bt = beta[ch.seg[c]];
  for (int i = ch.begin[c] + threadIdx.x; i < ch.end[c]; i += kBlock) p[i] = z[i] + bt * p[i];
}

// beta_s = (rz_new_s - rz_cross_s) / rz_old_s (flexible) or rz_new_s / rz_old_s; rz_old = rz_new
__global__ void k_seg_beta(int nseg, const double* __restrict__ rzn, const double* __restrict__ rzx,
                           double* __restrict__ rz, double* __restrict__ beta) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const double den = rz[s];
  beta[s] = den != 0.0 ? (rzx ? (rzn[s] - rzx[s]) / den : rzn[s] / den) : 0.0;
  rz[s] = rzn[s];
}

// flag[0] |= 1 when a row of segment s has a column outside the segment
__global__ void k_check_blocks(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                               const int* __restrict__ row_seg, const int* __restrict__ seg_ptr,
                               int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = row_seg[i];
  const int lo = seg_ptr[s], hi = seg_ptr[s + 1];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k)
    if (col[k] < lo || col[k] >= hi) {
      atomicOr(flag, 1);
      return;
    }
}

}  // namespace

void dot(const Chunks& ch, int nseg, const double* a, const double* b, double* partials, double* out,
         hipStream_t s) {
  if (ch.count == 0) return;
  k_chunk_dot<<<ch.count, kBlock, 0, s>>>(ch, a, b, partials);
  k_seg_reduce<<<(nseg + 63) / 64, 64, 0, s>>>(ch, nseg, partials, out);
}

void update(const Chunks& ch, int nseg, double* x, double* r, const double* p, const double* q, const double* rz,
            const double* pq, const int* active, double* partials, double* rr, hipStream_t s) {
  if (ch.count == 0) return;
  k_seg_update<<<ch.count, kBlock, 0, s>>>(ch, x, r, p, q, rz, pq, active, partials);
  k_seg_reduce<<<(nseg + 63) / 64, 64, 0, s>>>(ch, nseg, partials, rr);
}

void direction(const Chunks& ch, double* p, const double* z, const double* beta, hipStream_t s) {
  if (ch.count) k_seg_direction<<<ch.count, kBlock, 0, s>>>(ch, p, z, beta);
}

void beta(int nseg, const double* rzn, const double* rzx, double* rz, double* b, hipStream_t s) {
  k_seg_beta<<<(nseg + 63) / 64, 64, 0, s>>>(nseg, rzn, rzx, rz, b);
}

void check_blocks(int n, const int* ptr, const int* col, const int* row_seg, const int* seg_ptr, int* flag,
                  hipStream_t s) {
  if (n) k_check_blocks<<<(n + 255) / 256, 256, 0, s>>>(n, ptr, col, row_seg, seg_ptr, flag);
}

}  // namespace kseg
}  // namespace slx
