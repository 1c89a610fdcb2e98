// HIP kernels of the SIMP solve path for gfx950 (MI355X): 64-wide wavefronts,
// operator products in CSR with a group of lanes per row, deterministic
// two-stage reductions (fixed block -> partial map, partials summed in a fixed
// tree), structured-grid gather assembly without atomics, element kernels with
// the element matrix staged in LDS.
//
// Floating-point contraction is off for this file (build flag), so every
// expression rounds as written; that keeps assembly, filter and SIMP arithmetic
// bit-identical to the CPU restatement where the tests ask for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "slx_kernels.hpp"

namespace slx {
namespace k {
namespace {

constexpr int kBlock = 256;
constexpr int kReduceBlocks = 1024;  // fixed partial count of every reduction

__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  // Workgroups are dealt round-robin to the 8 XCDs; map them so each XCD sweeps
  // one contiguous range of rows (its own L2 then holds that range's x). The
  // grid is a multiple of 8 (grid8), which makes the map a permutation.
  return (b % 8) * (nblocks / 8) + b / 8;
}

template <int LANES>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LANES);
  return v;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  // wave reduce then cross-wave through LDS; fixed order
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) t += sh[i];
  }
  return t;  // valid on thread 0
}

// matrix values: FP64, FP32 or bfloat16 bits (widened exactly)
__device__ __forceinline__ double vload(const double* v, int k) { return v[k]; }
__device__ __forceinline__ double vload(const float* v, int k) { return (double)v[k]; }
__device__ __forceinline__ double vload(const uint16_t* v, int k) {
  return (double)__uint_as_float((uint32_t)v[k] << 16);
}

// One lane's share of a row's products: the lane takes every LANES-th entry;
// four entries are loaded before any is used so that each lane keeps eight
// independent loads (then four gathers) in flight.
template <int LANES, typename VAL, typename VT>
__device__ __forceinline__ double row_dot(int s, int e, int lane, const int* __restrict__ col,
                                          const VAL* __restrict__ val, const VT* __restrict__ x) {
  double acc = 0.0;
  int k = s + lane;
  for (; k + 3 * LANES < e; k += 4 * LANES) {
    const int c0 = col[k], c1 = col[k + LANES], c2 = col[k + 2 * LANES], c3 = col[k + 3 * LANES];
    const double v0 = vload(val, k), v1 = vload(val, k + LANES), v2 = vload(val, k + 2 * LANES),
                 v3 = vload(val, k + 3 * LANES);
    const VT x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
    acc += v0 * (double)x0;
    acc += v1 * (double)x1;
    acc += v2 * (double)x2;
    acc += v3 * (double)x3;
  }
  for (; k < e; k += LANES) acc += vload(val, k) * (double)x[col[k]];
  return acc;
}

// The smoother start fused into a product's epilogue (modes 4 and 5 of
// k_spmv): the output rows' D^-1, the start's d and iterate, and theta.
template <typename VT>
struct StartArgs {
  const VT* dinv = nullptr;
  VT* d = nullptr;
  VT* x2 = nullptr;
  double theta = 1.0;
};

// MODE 0: y = A x; 1: y = b - A x; 2: y += A x; 3: y = A x and partial x.y;
// 4: y = A x, then the zero-guess smoother start on y: d = dinv y / theta,
//    x2 = d (a restriction that starts the coarse level's pre-smoothing);
// 5: y = r = b - A x, then the smoother start from x: d = dinv r / theta,
//    x2 = x + d (post-smoothing; x2 is another buffer than x, which other
//    rows still read).
// Modes 4 and 5 round r and d to VT before using them, as the separate
// k_cheb_first_zero / k_cheb_first passes do, so the results are the same bits.
// cbase (may be null): rows of one node share one column list, entry k of row
// i reads its column at col[cbase[i] + k]
template <int LANES, int MODE, typename VT, typename VAL = VT>
__global__ void __launch_bounds__(kBlock) k_spmv(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                                                  const int* __restrict__ cbase,
                                                  const VAL* __restrict__ val, const VT* __restrict__ x,
                                                  const VT* __restrict__ b, VT* __restrict__ y,
                                                  double* __restrict__ partials, long long xs, long long ys,
                                                  StartArgs<VT> st) {
  __shared__ double sh[kBlock / 64];
  // load case blockIdx.y of a block of right-hand sides (case-major vectors, strides xs and ys)
  const long long vc = blockIdx.y;
  x += vc * xs;
  y += vc * ys;
  if (MODE == 1 || MODE == 5) b += vc * ys;
  if (MODE == 3) partials += vc * gridDim.x;
  if (MODE == 4 || MODE == 5) {
    st.d += vc * ys;
    st.x2 += vc * ys;
  }
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int rows_per_block = kBlock / LANES;
  const int row = blk * rows_per_block + threadIdx.x / LANES;
  const int lane = threadIdx.x % LANES;
  double prod = 0.0;
  if (row < n) {
    const int* colr = cbase ? col + cbase[row] : col;
    const double acc0 = row_dot<LANES, VAL, VT>(ptr[row], ptr[row + 1], lane, colr, val, x);
    const double acc = group_sum<LANES>(acc0);
    if (lane == 0) {
      if (MODE == 0) y[row] = (VT)acc;
      if (MODE == 1) y[row] = (VT)((double)b[row] - acc);
      if (MODE == 2) y[row] = (VT)((double)y[row] + acc);
      if (MODE == 3) {
        y[row] = (VT)acc;
        prod = (double)x[row] * acc;
      }
      if (MODE == 4) {
        const VT yr = (VT)acc;
        y[row] = yr;
        const double di = (double)st.dinv[row] * (double)yr / st.theta;
        st.d[row] = (VT)di;
        st.x2[row] = (VT)(0.0 + di);
      }
      if (MODE == 5) {
        const VT rr = (VT)((double)b[row] - acc);
        y[row] = rr;
        const double di = (double)st.dinv[row] * (double)rr / st.theta;
        st.d[row] = (VT)di;
        st.x2[row] = (VT)((double)x[row] + di);
      }
    }
  }
  if (MODE == 3) {
    const double t = block_sum(prod, sh);
    if (threadIdx.x == 0) partials[blk] = t;
  }
}

// One thread per row, products summed in column order: the CPU restatement's
// arithmetic exactly (used where results must match bit for bit).
__global__ void __launch_bounds__(kBlock) k_spmv_serial(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                                                        const double* __restrict__ val, const double* __restrict__ x,
                                                        double* __restrict__ y) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double s = 0.0;
  for (int k = ptr[row]; k < ptr[row + 1]; ++k) s += val[k] * x[col[k]];
  y[row] = s;
}

// Products staged through LDS, rows summed in column order by one thread each:
// the column-order arithmetic of k_spmv_serial with coalesced loads. A block
// handles SR rows whose entries fit in the stage (the host checks). MODE 0:
// y = A x; MODE 1: partial sums of vol_e * clamp(y_e, 0, 1) (OC volume).
constexpr int kStage = 4096;
template <int SR, int MODE>
__global__ void __launch_bounds__(kBlock) k_spmv_staged(int n, const int* __restrict__ ptr,
                                                         const int* __restrict__ col, const double* __restrict__ val,
                                                         const double* __restrict__ x, double* __restrict__ y,
                                                         const double* __restrict__ vol, double* __restrict__ partials) {
  __shared__ double prod[kStage];
  __shared__ double sh[kBlock / 64];
  const int r0 = blockIdx.x * SR;
  const int r1 = min(n, r0 + SR);
  const int k0 = ptr[r0], k1 = ptr[r1];
  for (int k = k0 + (int)threadIdx.x; k < k1; k += kBlock) prod[k - k0] = val[k] * x[col[k]];
  __syncthreads();
  double part = 0.0;
  if ((int)threadIdx.x < r1 - r0) {
    const int row = r0 + threadIdx.x;
    double acc = 0.0;
    for (int k = ptr[row]; k < ptr[row + 1]; ++k) acc += prod[k - k0];
    if (MODE == 0) y[row] = acc;
    if (MODE == 1) part = fmin(fmax(acc, 0.0), 1.0) * vol[row];
  }
  if (MODE == 1) {
    const double t = block_sum(part, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
  }
}

// Chebyshev step (src/fem/Multigrid.cpp:833-853), fused:
//   r_i <- r_i - (A d_in)_i;  d_out_i = c1 d_in_i + c2 dinv_i r_i;  x_i += d_out_i
template <int LANES, typename VT, typename VAL = VT>
__global__ void __launch_bounds__(kBlock) k_cheb_step(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                                                       const int* __restrict__ cbase,
                                                       const VAL* __restrict__ val, const VT* __restrict__ dinv,
                                                       const VT* __restrict__ r_in, VT* __restrict__ r_out,
                                                       const VT* __restrict__ d_in, VT* __restrict__ d_out,
                                                       VT* __restrict__ x, double c1, double c2, long long vs) {
  const long long vc = blockIdx.y;  // load case of a block of right-hand sides
  r_in += vc * vs;
  r_out += vc * vs;
  d_in += vc * vs;
  d_out += vc * vs;
  x += vc * vs;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int row = blk * (kBlock / LANES) + threadIdx.x / LANES;
  const int lane = threadIdx.x % LANES;
  if (row >= n) return;
  const int* colr = cbase ? col + cbase[row] : col;
  const double acc = group_sum<LANES>(row_dot<LANES, VAL, VT>(ptr[row], ptr[row + 1], lane, colr, val, d_in));
  if (lane == 0) {
    const double r = (double)r_in[row] - acc;
    const double dn = c1 * (double)d_in[row] + c2 * (double)dinv[row] * r;
    r_out[row] = (VT)r;
    d_out[row] = (VT)dn;
    x[row] = (VT)((double)x[row] + dn);
  }
}

// Zero-guess start of the smoother: d = dinv b / theta, x = d.
template <typename VT>
__global__ void k_cheb_first_zero(int n, const VT* __restrict__ dinv, const VT* __restrict__ b, double theta,
                                  VT* __restrict__ x, VT* __restrict__ d, long long vs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long o = blockIdx.y * vs;  // load case blockIdx.y
  b += o;
  x += o;
  d += o;
  const double di = (double)dinv[i] * (double)b[i] / theta;
  d[i] = (VT)di;
  x[i] = (VT)(0.0 + di);
}

// Non-zero-guess start: d = dinv r / theta, x += d (r = b - A x computed before).
template <typename VT>
__global__ void k_cheb_first(int n, const VT* __restrict__ dinv, const VT* __restrict__ r, double theta,
                             VT* __restrict__ x, VT* __restrict__ d, long long vs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long o = blockIdx.y * vs;  // load case blockIdx.y
  r += o;
  x += o;
  d += o;
  const double di = (double)dinv[i] * (double)r[i] / theta;
  d[i] = (VT)di;
  x[i] = (VT)((double)x[i] + di);
}

// y = M b for a dense row-major n x n M: one wave per row.
template <typename VT>
__global__ void __launch_bounds__(kBlock) k_dense_gemv(int n, const VT* __restrict__ M, const VT* __restrict__ b,
                                                        VT* __restrict__ y, long long vs) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  b += blockIdx.y * vs;  // load case blockIdx.y
  y += blockIdx.y * vs;
  double acc = 0.0;
  const VT* m = M + (size_t)row * n;
  for (int j = lane; j < n; j += 64) acc += (double)m[j] * (double)b[j];
  acc = group_sum<64>(acc);
  if (lane == 0) y[row] = (VT)acc;
}

__global__ void __launch_bounds__(kBlock) k_dot_partials(int n, const double* __restrict__ a,
                                                          const double* __restrict__ b, double* __restrict__ partials,
                                                          long long vs) {
  __shared__ double sh[kBlock / 64];
  a += blockIdx.y * vs;  // load case blockIdx.y: its partials follow the previous case's
  b += blockIdx.y * vs;
  partials += (long long)blockIdx.y * gridDim.x;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += a[i] * b[i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// partials of sum rho (1 - rho), the grey level's sum
__global__ void __launch_bounds__(kBlock) k_grey_partials(int n, const double* __restrict__ rho,
                                                           double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += rho[i] * (1.0 - rho[i]);
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Sums m partials in a fixed order into out[0]; optional sqrt.
__global__ void __launch_bounds__(1024) k_reduce_partials(int m, const double* __restrict__ partials,
                                                          double* __restrict__ out, int take_sqrt, int out_stride) {
  __shared__ double sh[16];
  partials += (long long)blockIdx.x * m;  // one block per load case
  out += (long long)blockIdx.x * out_stride;
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) s += partials[i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(t) : t;
}

// PCG update (src/fem/Multigrid.cpp:1194-1203): alpha = rz / pq;
// x += alpha p; r -= alpha q; partial sums of r.r
__global__ void __launch_bounds__(kBlock) k_pcg_update(int n, double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        const double* __restrict__ scal, int i_rz, int i_pq,
                                                        double* __restrict__ partials, long long vs,
                                                        const int* __restrict__ active) {
  __shared__ double sh[kBlock / 64];
  // load case blockIdx.y: vectors at stride vs, 16 scalars per case, partials per case;
  // a case whose active flag is 0 keeps its iterate and residual
  const long long o = blockIdx.y * vs;
  x += o;
  r += o;
  p += o;
  q += o;
  scal += 16 * blockIdx.y;
  partials += (long long)blockIdx.y * gridDim.x;
  const bool live = !active || active[blockIdx.y];
  const double alpha = live ? scal[i_rz] / scal[i_pq] : 0.0;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (live) {
      x[i] += alpha * p[i];
      const double ri = r[i] - alpha * q[i];
      r[i] = ri;
      s += ri * ri;
    } else {
      s += r[i] * r[i];
    }
  }
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// End of the preconditioner step (src/fem/Multigrid.cpp:1205-1215), first
// pass: z = the cycle's result (zc, widened; null: z already holds it) and the
// block partials of r.z and, for flexible CG, of r.z_old with the previous z.
template <typename VT>
__global__ void __launch_bounds__(kBlock) k_precond_partials(int n, const VT* __restrict__ zc,
                                                              const double* __restrict__ r, double* __restrict__ z,
                                                              int flexible, double* __restrict__ partials,
                                                              long long vs) {
  __shared__ double sh[kBlock / 64];
  const long long o = blockIdx.y * vs;  // load case blockIdx.y
  if (zc) zc += o;
  r += o;
  z += o;
  partials += 2LL * blockIdx.y * gridDim.x;
  double s1 = 0.0, s2 = 0.0;
  // four of the thread's elements at a time: every load before any sum, the
  // sums in the thread's ascending order
  const int stride = gridDim.x * blockDim.x;
  for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
    double zo[4], zn[4], ri[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * stride;
      const bool in = i < n;
      zo[u] = in ? z[i] : 0.0;
      zn[u] = in ? (zc ? (double)zc[i] : zo[u]) : 0.0;
      ri[u] = in ? r[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * stride;
      if (i >= n) break;
      if (zc) z[i] = zn[u];
      s1 += ri[u] * zn[u];
      s2 += ri[u] * zo[u];
    }
  }
  const double t1 = block_sum(s1, sh);
  __syncthreads();
  const double t2 = flexible ? block_sum(s2, sh) : 0.0;
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t1;
    partials[gridDim.x + blockIdx.x] = t2;
  }
}

// Second pass (one block): rz_new and r.z_old summed in a fixed order, then
// beta = rz_new / rz_old (flexible: (rz_new - r.z_old) / rz_old) into
// scal[i_beta], and scal[i_rzn] = scal[i_rz] = rz_new.
__global__ void __launch_bounds__(1024) k_precond_finish(int m, const double* __restrict__ partials, double* scal,
                                                         int i_rz, int i_rzn, int i_beta, int flexible) {
  __shared__ double sh[16];
  partials += 2LL * blockIdx.x * m;  // one block per load case
  scal += 16 * blockIdx.x;
  double v = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) v += partials[i];
  const double rzn = block_sum(v, sh);
  __syncthreads();
  double w = 0.0;
  if (flexible)
    for (int i = threadIdx.x; i < m; i += blockDim.x) w += partials[m + i];
  const double rzx = flexible ? block_sum(w, sh) : 0.0;
  if (threadIdx.x == 0) {
    const double rz = scal[i_rz];
    scal[i_beta] = flexible ? (rzn - rzx) / rz : rzn / rz;
    scal[i_rzn] = rzn;
    scal[i_rz] = rzn;
  }
}

// Start of the FP32 cycle: b = (float) r, then the zero-guess smoother start
// x = d = dinv b / theta.
__global__ void k_cheb_first_zero_from64(int n, const float* __restrict__ dinv, const double* __restrict__ r,
                                         double theta, float* __restrict__ b, float* __restrict__ x,
                                         float* __restrict__ d, long long vs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long o = blockIdx.y * vs;  // load case blockIdx.y
  r += o;
  b += o;
  x += o;
  d += o;
  const float bi = (float)r[i];
  b[i] = bi;
  const double di = (double)dinv[i] * (double)bi / theta;
  d[i] = (float)di;
  x[i] = (float)(0.0 + di);
}

// p = z + beta p with beta = rz_new / rz_old (src/fem/Multigrid.cpp:1209-1215)
__global__ void k_pcg_direction(int n, double* __restrict__ p, const double* __restrict__ z,
                                const double* __restrict__ scal, int i_new, int i_old, long long vs) {
  p += blockIdx.y * vs;  // load case blockIdx.y
  z += blockIdx.y * vs;
  scal += 16 * blockIdx.y;
  const double beta = i_old >= 0 ? scal[i_new] / scal[i_old] : scal[i_new];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

__global__ void k_copy_scalar(double* scal, int from, int to) { scal[to] = scal[from]; }

// Polak-Ribiere beta of flexible CG: (r.z_new - r.z_old) / rz_old
__global__ void k_flexible_beta(double* scal, int i_new, int i_cross, int i_old, int i_out) {
  scal[i_out] = (scal[i_new] - scal[i_cross]) / scal[i_old];
}

// double -> bfloat16 bits via FP32, round to nearest even
__global__ void k_to_bf16(long long n, const double* __restrict__ in, uint16_t* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t u = __float_as_uint((float)in[i]);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    out[i] = (uint16_t)((u >> 16) | 0x40u);
    return;
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  out[i] = (uint16_t)(u >> 16);
}

template <typename A, typename B>
__global__ void k_convert(int n, const A* __restrict__ in, B* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (B)in[i];
}

__global__ void k_axpy(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// ---------------------------------------------------------------------------
// Structured-grid gather assembly: every K_ff entry sums s_e Ke0 over the
// cells shared by its two nodes in ascending element order, which is the order
// the reference's element-loop scatter adds them (src/fem/Assembler.cpp:182-214).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int hex_corner(int lx, int ly, int lz) {
  // VTK hexahedron corner of the local offsets (StructuredMesh.cpp:224-234)
  return lz * 4 + (ly ? (lx ? 2 : 3) : (lx ? 1 : 0));
}

// A free DOF's grid node and component: from the packed table (x | y << 10 |
// z << 20 | c << 30, built once per model when the grid fits) or by division.
template <bool PACKED>
__device__ __forceinline__ void dof_coords3(int f, const int* __restrict__ free_dofs,
                                            const unsigned* __restrict__ xyzc, int nnx, int nny, int& x, int& y, int& z,
                                            int& c) {
  if (PACKED) {
    const unsigned p = xyzc[f];
    x = (int)(p & 1023u), y = (int)((p >> 10) & 1023u), z = (int)((p >> 20) & 1023u), c = (int)(p >> 30);
  } else {
    const int g = free_dofs[f];
    const int I = g / 3;
    c = g % 3;
    x = I % nnx, y = (I / nnx) % nny, z = I / (nnx * nny);
  }
}

template <int LANES, bool PACKED>
__global__ void __launch_bounds__(kBlock) k_assemble_hex8(int nrows, const int* __restrict__ ptr,
                                                           const int* __restrict__ col, double* __restrict__ val,
                                                           const int* __restrict__ free_dofs,
                                                           const unsigned* __restrict__ xyzc,
                                                           const double* __restrict__ scale,
                                                           const double* __restrict__ ke0, int nx, int ny, int nz) {
  __shared__ double ke[576];
  for (int t = threadIdx.x; t < 576; t += blockDim.x) ke[t] = ke0[t];
  __syncthreads();
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int row = blk * (kBlock / LANES) + threadIdx.x / LANES;
  const int lane = threadIdx.x % LANES;
  if (row >= nrows) return;
  const int nnx = nx + 1, nny = ny + 1;
  int ix, iy, iz, ci;
  dof_coords3<PACKED>(row, free_dofs, xyzc, nnx, nny, ix, iy, iz, ci);
  for (int k = ptr[row] + lane; k < ptr[row + 1]; k += LANES) {
    int jx, jy, jz, cj;
    dof_coords3<PACKED>(col[k], free_dofs, xyzc, nnx, nny, jx, jy, jz, cj);
    const int x0 = max(max(ix, jx) - 1, 0), x1 = min(min(ix, jx), nx - 1);
    const int y0 = max(max(iy, jy) - 1, 0), y1 = min(min(iy, jy), ny - 1);
    const int z0 = max(max(iz, jz) - 1, 0), z1 = min(min(iz, jz), nz - 1);
    double v = 0.0;
    for (int cz = z0; cz <= z1; ++cz)
      for (int cy = y0; cy <= y1; ++cy)
        for (int cx = x0; cx <= x1; ++cx) {
          const int e = (cz * ny + cy) * nx + cx;
          const int a = hex_corner(jx - cx, jy - cy, jz - cz);
          const int b = hex_corner(ix - cx, iy - cy, iz - cz);
          v += scale[e] * ke[(a * 3 + cj) * 24 + b * 3 + ci];
        }
    val[k] = v;
  }
}

// Node-block assembly of a Hex8 K_ff: 32 lanes per node, lane q < 27 the
// neighbour at offset (q % 3, q / 3 % 3, q / 9) - 1 (ascending node order). A
// node's rows share one pattern, the free DOFs of its neighbours in ascending
// node order, so the lane's block starts at the number of free components of the
// earlier neighbours in every row of the node. The lane sums the 3 x 3 block over
// the cells the two nodes share in ascending element order, each entry with the
// products and sums of k_assemble_hex8 (the same bits), and writes its free
// entries. It reads no column indices.
__global__ void __launch_bounds__(kBlock) k_assemble_hex8_nodes(int nnx, int nny, int nnz,
                                                                 const int* __restrict__ nodemap,
                                                                 const int* __restrict__ ptr,
                                                                 const double* __restrict__ scale,
                                                                 const double* __restrict__ ke0,
                                                                 double* __restrict__ val) {
  __shared__ double ke[576];
  __shared__ int cnt[kBlock];
  for (int t = threadIdx.x; t < 576; t += blockDim.x) ke[t] = ke0[t];
  const int nx = nnx - 1, ny = nny - 1, nz = nnz - 1;  // cells per direction
  const long long nn = (long long)nnx * nny * nnz;
  const int q = threadIdx.x & 31, g0 = threadIdx.x & ~31;
  const long long node = (long long)xcd_block(blockIdx.x, gridDim.x) * (kBlock / 32) + (threadIdx.x >> 5);
  const bool valid = node < nn && q < 27;
  int ix = 0, iy = 0, iz = 0, jx = 0, jy = 0, jz = 0, mi = 0, mj = 0;
  if (valid) {
    ix = (int)(node % nnx), iy = (int)((node / nnx) % nny), iz = (int)(node / ((long long)nnx * nny));
    jx = ix + q % 3 - 1, jy = iy + (q / 3) % 3 - 1, jz = iz + q / 9 - 1;
    mi = nodemap[node];
    if (jx >= 0 && jx < nnx && jy >= 0 && jy < nny && jz >= 0 && jz < nnz)
      mj = nodemap[((long long)jz * nny + jy) * nnx + jx];
  }
  cnt[threadIdx.x] = __popc(mj & 7);
  __syncthreads();
  if ((mi & 7) == 0 || (mj & 7) == 0) return;
  int off = 0;
  for (int p = 0; p < q; ++p) off += cnt[g0 + p];
  const int x0 = max(max(ix, jx) - 1, 0), x1 = min(min(ix, jx), nx - 1);
  const int y0 = max(max(iy, jy) - 1, 0), y1 = min(min(iy, jy), ny - 1);
  const int z0 = max(max(iz, jz) - 1, 0), z1 = min(min(iz, jz), nz - 1);
  double v[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int cz = z0; cz <= z1; ++cz)
    for (int cy = y0; cy <= y1; ++cy)
      for (int cx = x0; cx <= x1; ++cx) {
        const double se = scale[((long long)cz * ny + cy) * nx + cx];
        const int a = hex_corner(jx - cx, jy - cy, jz - cz);
        const int b = hex_corner(ix - cx, iy - cy, iz - cz);
#pragma unroll
        for (int ci = 0; ci < 3; ++ci)
#pragma unroll
          for (int cj = 0; cj < 3; ++cj) v[ci][cj] += se * ke[(a * 3 + cj) * 24 + b * 3 + ci];
      }
  int row = mi >> 3;
#pragma unroll
  for (int ci = 0; ci < 3; ++ci) {
    if (!((mi >> ci) & 1)) continue;
    double* out = val + ptr[row] + off;
    int k = 0;
#pragma unroll
    for (int cj = 0; cj < 3; ++cj)
      if ((mj >> cj) & 1) out[k++] = v[ci][cj];
    ++row;
  }
}

// x | y << 15 | c << 30 for Q4
template <bool PACKED>
__device__ __forceinline__ void dof_coords2(int f, const int* __restrict__ free_dofs,
                                            const unsigned* __restrict__ xyzc, int nnx, int& x, int& y, int& c) {
  if (PACKED) {
    const unsigned p = xyzc[f];
    x = (int)(p & 32767u), y = (int)((p >> 15) & 32767u), c = (int)(p >> 30);
  } else {
    const int g = free_dofs[f];
    const int I = g / 2;
    c = g % 2;
    x = I % nnx, y = I / nnx;
  }
}

template <int LANES, bool PACKED>
__global__ void __launch_bounds__(kBlock) k_assemble_quad4(int nrows, const int* __restrict__ ptr,
                                                            const int* __restrict__ col, double* __restrict__ val,
                                                            const int* __restrict__ free_dofs,
                                                            const unsigned* __restrict__ xyzc,
                                                            const double* __restrict__ scale,
                                                            const double* __restrict__ ke0, int nx, int ny) {
  __shared__ double ke[64];
  for (int t = threadIdx.x; t < 64; t += blockDim.x) ke[t] = ke0[t];
  __syncthreads();
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int row = blk * (kBlock / LANES) + threadIdx.x / LANES;
  const int lane = threadIdx.x % LANES;
  if (row >= nrows) return;
  const int nnx = nx + 1;
  int ix, iy, ci;
  dof_coords2<PACKED>(row, free_dofs, xyzc, nnx, ix, iy, ci);
  for (int k = ptr[row] + lane; k < ptr[row + 1]; k += LANES) {
    int jx, jy, cj;
    dof_coords2<PACKED>(col[k], free_dofs, xyzc, nnx, jx, jy, cj);
    const int x0 = max(max(ix, jx) - 1, 0), x1 = min(min(ix, jx), nx - 1);
    const int y0 = max(max(iy, jy) - 1, 0), y1 = min(min(iy, jy), ny - 1);
    double v = 0.0;
    for (int cy = y0; cy <= y1; ++cy)
      for (int cx = x0; cx <= x1; ++cx) {
        const int e = cy * nx + cx;
        const int a = hex_corner(jx - cx, jy - cy, 0);
        const int b = hex_corner(ix - cx, iy - cy, 0);
        v += scale[e] * ke[(a * 2 + cj) * 8 + b * 2 + ci];
      }
    val[k] = v;
  }
}

// ---------------------------------------------------------------------------
// Element loop of the compliance sensitivity (src/topopt/Sensitivity.cpp:182-189):
// q_e = u_e^T Ke0 u_e per load case; energy_e += w 0.5 s_e q_e; dc_e -= w ds_e q_e.
// One thread per element, Ke0 in LDS (every lane reads the same entry: broadcast).
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_sensitivity(int nelem, int nx, int ny, const double* __restrict__ ke0,
                                                         const int* __restrict__ reduced, const double* __restrict__ u,
                                                         long long nfree, int ncases, const double* __restrict__ w,
                                                         const double* __restrict__ fac,
                                                         const double* __restrict__ dfac,
                                                         double* __restrict__ energy, double* __restrict__ dc) {
  constexpr int NPE = DIM == 3 ? 8 : 4;
  constexpr int ND = NPE * DIM;
  // Ke0 is read with wave-uniform addresses (scalar loads); both loops below are
  // unrolled so that the element vector stays in registers.
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int nnx = nx + 1, nny = ny + 1;
  int nodes[NPE];
  if (DIM == 3) {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    const int n0 = (ez * nny + ey) * nnx + ex, lay = nnx * nny;
    nodes[0] = n0; nodes[1] = n0 + 1; nodes[2] = n0 + nnx + 1; nodes[3] = n0 + nnx;
    nodes[4] = n0 + lay; nodes[5] = n0 + lay + 1; nodes[6] = n0 + lay + nnx + 1; nodes[7] = n0 + lay + nnx;
  } else {
    const int ex = e % nx, ey = e / nx;
    const int n0 = ey * nnx + ex;
    nodes[0] = n0; nodes[1] = n0 + 1; nodes[2] = n0 + nnx + 1; nodes[3] = n0 + nnx;
  }
  int idx[ND];
#pragma unroll
  for (int a = 0; a < NPE; ++a)
#pragma unroll
    for (int c = 0; c < DIM; ++c) idx[a * DIM + c] = reduced[nodes[a] * DIM + c];
  double en = 0.0, d = 0.0;
  const double fe = fac[e], dfe = dfac[e];
  for (int l = 0; l < ncases; ++l) {
    const double* ul = u + (size_t)l * nfree;
    double ue[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) ue[i] = idx[i] >= 0 ? ul[idx[i]] : 0.0;
    double quad = 0.0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < ND; ++j) s += ke0[i * ND + j] * ue[j];
      quad += ue[i] * s;
    }
    en += w[l] * 0.5 * fe * quad;
    d -= w[l] * dfe * quad;
  }
  energy[e] = en;
  dc[e] = d;
}

// physical = clamp(y, 0, 1); fac = SIMP factor; dfac = its derivative
// (src/topopt/SimpInterpolation.cpp:59-70)
__global__ void k_simp(int n, const double* __restrict__ filtered, double penalty, double emin,
                       double* __restrict__ phys, double* __restrict__ fac, double* __restrict__ dfac) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double r = fmin(fmax(filtered[e], 0.0), 1.0);
  phys[e] = r;
  fac[e] = emin + (1.0 - emin) * pow(r, penalty);
  dfac[e] = penalty * (1.0 - emin) * pow(r, penalty - 1.0);
}

// Heaviside projection of the clamped filtered density (src/topopt/Projection.cpp:58-70);
// in place allowed. dproj (may be null) = its derivative with respect to the clamped value.
__global__ void k_project(int n, const double* in, double beta, double eta, double* out, double* dproj) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double r = fmin(fmax(in[e], 0.0), 1.0);
  const double a = tanh(beta * eta);
  const double denom = a + tanh(beta * (1.0 - eta));
  const double t = tanh(beta * (r - eta));
  out[e] = (a + t) / denom;
  if (dproj) dproj[e] = beta * (1.0 - t * t) / denom;
}

__global__ void k_mul(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) c[e] = a[e] * b[e];
}

// OC candidate (src/topopt/OptimalityCriteria.cpp:14-36)
__global__ void k_oc_candidate(int n, const double* __restrict__ x, const double* __restrict__ dc,
                               const double* __restrict__ dv, const double* __restrict__ lower,
                               const double* __restrict__ upper, const unsigned char* __restrict__ passive,
                               double lambda, double move, double damping, double* __restrict__ xn) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double lo = fmax(lower[e], x[e] - move);
  const double hi = fmin(upper[e], x[e] + move);
  if (passive[e]) {
    xn[e] = lower[e];
    return;
  }
  const double denom = lambda * dv[e];
  double b = 0.0;
  if (denom > 0.0) b = fmax(0.0, -dc[e] / denom);
  const double target = x[e] * pow(b, damping);
  xn[e] = fmin(fmax(target, lo), hi);
}

// One trial of the OC bisection on the device. lambda = sqrt(lo hi) from the
// state (or lam when explicit); writes the candidate and the per-block partial
// sums of w.x_new, w = H^T vol (the filtered volume of a design within [0, 1]).
// A finished bisection (state[5] != 0) makes further trials no-ops.
// state: lo, hi, volume, lambda, trials, done, converged
__global__ void __launch_bounds__(kBlock) k_oc_trial(int n, const double* __restrict__ x, const double* __restrict__ dc,
                                                      const double* __restrict__ dv, const double* __restrict__ lower,
                                                      const double* __restrict__ upper,
                                                      const unsigned char* __restrict__ passive, double move,
                                                      double damping, const double* __restrict__ w,
                                                      const double* __restrict__ state, int explicit_lambda, double lam,
                                                      double* __restrict__ xn, double* __restrict__ partials) {
  if (!explicit_lambda && state[5] != 0.0) return;
  __shared__ double sh[kBlock / 64];
  const double lambda = explicit_lambda ? lam : sqrt(state[0] * state[1]);
  double acc = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    double v;
    if (passive[e]) {
      v = lower[e];
    } else {
      const double lo = fmax(lower[e], x[e] - move);
      const double hi = fmin(upper[e], x[e] + move);
      const double denom = lambda * dv[e];
      double b = 0.0;
      if (denom > 0.0) b = fmax(0.0, -dc[e] / denom);
      v = fmin(fmax(x[e] * pow(b, damping), lo), hi);
    }
    xn[e] = v;
    acc += w[e] * v;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Sums a trial's partials into its volume and applies the bisection rule of
// optimality_criteria_update (src/topopt/OptimalityCriteria.cpp:96-139).
__global__ void __launch_bounds__(1024) k_oc_step(int m, const double* __restrict__ partials, double* __restrict__ state,
                                                  int explicit_lambda, double target, double vol_tol,
                                                  int max_bisections) {
  if (!explicit_lambda && state[5] != 0.0) return;
  __shared__ double sh[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) s += partials[i];
  const double V = block_sum(s, sh);
  if (threadIdx.x != 0) return;
  state[2] = V;
  if (explicit_lambda) return;
  double lo = state[0], hi = state[1];
  const double mid = sqrt(lo * hi);
  const double it = state[4] + 1.0;
  state[4] = it;
  state[3] = mid;
  if (fabs(V - target) <= vol_tol * target) {
    state[5] = 1.0;
    state[6] = 1.0;
    return;
  }
  if (V > target) lo = mid; else hi = mid;
  state[0] = lo;
  state[1] = hi;
  if (hi / lo - 1.0 < 1.0e-15) {
    state[5] = 1.0;
    state[6] = fabs(V - target) <= 1.0e-6 * fmax(target, 1.0e-30) ? 1.0 : 0.0;
    return;
  }
  if (it >= (double)max_bisections) state[5] = 1.0;
}

// partial sums of vol_e * clamp((H x)_e, 0, 1); one thread per row, column order
__global__ void __launch_bounds__(kBlock) k_filter_volume(int n, const int* __restrict__ ptr,
                                                           const int* __restrict__ col, const double* __restrict__ val,
                                                           const double* __restrict__ x, const double* __restrict__ vol,
                                                           double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  double acc = 0.0;
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = ptr[row]; k < ptr[row + 1]; ++k) s += val[k] * x[col[k]];
    acc += fmin(fmax(s, 0.0), 1.0) * vol[row];
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// partial maxima of |a - b|
__global__ void __launch_bounds__(kBlock) k_max_abs_diff(int n, const double* __restrict__ a,
                                                          const double* __restrict__ b, double* __restrict__ partials) {
  __shared__ double sh[kBlock];
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, fabs(a[i] - b[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_reduce_max(int m, const double* __restrict__ partials,
                                                        double* __restrict__ out) {
  __shared__ double sh[kBlock];
  double v = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) v = fmax(v, partials[i]);
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void k_clamp(int n, double* __restrict__ x, const double* __restrict__ lo, const double* __restrict__ hi) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) x[e] = fmin(fmax(x[e], lo[e]), hi[e]);
}

int blocks_for(long long n, int per_block) { return (int)((n + per_block - 1) / per_block); }
int grid8(long long n, int per_block) { return (blocks_for(n, per_block) + 7) / 8 * 8; }

int lanes_for(const DCsr& A) {
  // average row length decides the group width
  const double avg = A.rows ? double(A.nnz) / A.rows : 0.0;
  static const int forced = [] {
    const char* e = std::getenv("SPARLAB_SPMV_LANES");
    return e ? std::atoi(e) : 0;
  }();
  auto valid = [](int l) { return l == 1 || l == 4 || l == 8 || l == 16 || l == 64; };
  if (valid(forced)) return forced;
  if (valid(A.lanes)) return A.lanes;
  return avg > 256 ? 64 : avg > 48 ? 16 : avg > 12 ? 8 : 4;
}

// nv load cases (gridDim.y), case-major vectors: x at stride A.cols, b and y at stride A.rows
template <int MODE, typename VT, typename VAL = VT>
void launch_spmv(const DCsr& A, const VAL* val, const VT* x, const VT* b, VT* y, double* partials, hipStream_t s,
                 int nv = 1, StartArgs<VT> st = StartArgs<VT>()) {
  const int lanes = lanes_for(A);
  const dim3 nb(grid8(A.rows, kBlock / lanes), nv);
  if (A.rows == 0) return;
  const int* cols = A.cbase.p ? A.ncol.p : A.col.p;
  const long long xs = A.cols, ys = A.rows;
  switch (lanes) {
    case 64: k_spmv<64, MODE, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, x, b, y, partials, xs, ys, st); break;
    case 16: k_spmv<16, MODE, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, x, b, y, partials, xs, ys, st); break;
    case 8: k_spmv<8, MODE, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, x, b, y, partials, xs, ys, st); break;
    case 1: k_spmv<1, MODE, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, x, b, y, partials, xs, ys, st); break;
    default: k_spmv<4, MODE, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, x, b, y, partials, xs, ys, st); break;
  }
}

// FP32-vector products read the bfloat16 copy of the values when there is one
template <int MODE>
void launch_spmv32(const DCsr& A, const float* x, const float* b, float* y, hipStream_t s, int nv,
                   StartArgs<float> st = StartArgs<float>()) {
  if (A.val16.p) launch_spmv<MODE, float, uint16_t>(A, A.val16.p, x, b, y, nullptr, s, nv, st);
  else launch_spmv<MODE, float>(A, A.val32.p, x, b, y, nullptr, s, nv, st);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void spmv(const DCsr& A, const double* x, double* y, hipStream_t s, int nv) {
  launch_spmv<0, double>(A, A.val.p, x, nullptr, y, nullptr, s, nv);
}
void spmv_add(const DCsr& A, const double* x, double* y, hipStream_t s, int nv) {
  launch_spmv<2, double>(A, A.val.p, x, nullptr, y, nullptr, s, nv);
}
void residual(const DCsr& A, const double* x, const double* b, double* r, hipStream_t s, int nv) {
  launch_spmv<1, double>(A, A.val.p, x, b, r, nullptr, s, nv);
}
void spmv32(const DCsr& A, const float* x, float* y, hipStream_t s, int nv) {
  launch_spmv32<0>(A, x, nullptr, y, s, nv);
}
void spmv_add32(const DCsr& A, const float* x, float* y, hipStream_t s, int nv) {
  launch_spmv32<2>(A, x, nullptr, y, s, nv);
}
void residual32(const DCsr& A, const float* x, const float* b, float* r, hipStream_t s, int nv) {
  launch_spmv32<1>(A, x, b, r, s, nv);
}
void spmv_start(const DCsr& A, const double* x, double* y, const double* dinv, double theta, double* d, double* x2,
                hipStream_t s, int nv) {
  launch_spmv<4, double>(A, A.val.p, x, nullptr, y, nullptr, s, nv, StartArgs<double>{dinv, d, x2, theta});
}
void residual_first(const DCsr& A, const double* x, const double* b, double* r, const double* dinv, double theta,
                    double* d, double* x2, hipStream_t s, int nv) {
  launch_spmv<5, double>(A, A.val.p, x, b, r, nullptr, s, nv, StartArgs<double>{dinv, d, x2, theta});
}
void spmv_start32(const DCsr& A, const float* x, float* y, const float* dinv, double theta, float* d, float* x2,
                  hipStream_t s, int nv) {
  launch_spmv32<4>(A, x, nullptr, y, s, nv, StartArgs<float>{dinv, d, x2, theta});
}
void residual_first32(const DCsr& A, const float* x, const float* b, float* r, const float* dinv, double theta,
                      float* d, float* x2, hipStream_t s, int nv) {
  launch_spmv32<5>(A, x, b, r, s, nv, StartArgs<float>{dinv, d, x2, theta});
}
int spmv_dot_blocks(const DCsr& A) { return grid8(A.rows, kBlock / lanes_for(A)); }
void spmv_dot(const DCsr& A, const double* x, double* y, double* partials, hipStream_t s, int nv) {
  launch_spmv<3, double>(A, A.val.p, x, nullptr, y, partials, s, nv);
}
int staged_rows(const DCsr& A) {
  // rows per block such that every block's entries fit the stage
  return A.max_row * 256 <= kStage ? 256 : A.max_row * 128 <= kStage ? 128 : A.max_row * 64 <= kStage ? 64 : 0;
}
void spmv_serial(const DCsr& A, const double* x, double* y, hipStream_t s) {
  if (!A.rows) return;
  switch (staged_rows(A)) {
    case 256: k_spmv_staged<256, 0><<<blocks_for(A.rows, 256), kBlock, 0, s>>>(A.rows, A.ptr.p, A.col.p, A.val.p, x, y, nullptr, nullptr); break;
    case 128: k_spmv_staged<128, 0><<<blocks_for(A.rows, 128), kBlock, 0, s>>>(A.rows, A.ptr.p, A.col.p, A.val.p, x, y, nullptr, nullptr); break;
    case 64: k_spmv_staged<64, 0><<<blocks_for(A.rows, 64), kBlock, 0, s>>>(A.rows, A.ptr.p, A.col.p, A.val.p, x, y, nullptr, nullptr); break;
    default: k_spmv_serial<<<blocks_for(A.rows, kBlock), kBlock, 0, s>>>(A.rows, A.ptr.p, A.col.p, A.val.p, x, y); break;
  }
}

template <typename VT, typename VAL = VT>
static void cheb_step_t(const DCsr& A, const VAL* val, const VT* dinv, const VT* r_in, VT* r_out, const VT* d_in,
                        VT* d_out, VT* x, double c1, double c2, hipStream_t s, int nv = 1) {
  const int lanes = lanes_for(A);
  const dim3 nb(grid8(A.rows, kBlock / lanes), nv);
  const long long vs = A.rows;
  if (A.rows == 0) return;
  const int* cols = A.cbase.p ? A.ncol.p : A.col.p;
  switch (lanes) {
    case 64: k_cheb_step<64, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, dinv, r_in, r_out, d_in, d_out, x, c1, c2, vs); break;
    case 16: k_cheb_step<16, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, dinv, r_in, r_out, d_in, d_out, x, c1, c2, vs); break;
    case 8: k_cheb_step<8, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, dinv, r_in, r_out, d_in, d_out, x, c1, c2, vs); break;
    case 1: k_cheb_step<1, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, dinv, r_in, r_out, d_in, d_out, x, c1, c2, vs); break;
    default: k_cheb_step<4, VT, VAL><<<nb, kBlock, 0, s>>>(A.rows, A.ptr.p, cols, A.cbase.p, val, dinv, r_in, r_out, d_in, d_out, x, c1, c2, vs); break;
  }
}
void cheb_step(const DCsr& A, const double* dinv, const double* r_in, double* r_out, const double* d_in, double* d_out,
               double* x, double c1, double c2, hipStream_t s, int nv) {
  cheb_step_t<double>(A, A.val.p, dinv, r_in, r_out, d_in, d_out, x, c1, c2, s, nv);
}
void cheb_step32(const DCsr& A, const float* dinv, const float* r_in, float* r_out, const float* d_in, float* d_out,
                 float* x, double c1, double c2, hipStream_t s, int nv) {
  if (A.val16.p)
    cheb_step_t<float, uint16_t>(A, A.val16.p, dinv, r_in, r_out, d_in, d_out, x, c1, c2, s, nv);
  else
    cheb_step_t<float>(A, A.val32.p, dinv, r_in, r_out, d_in, d_out, x, c1, c2, s, nv);
}
void cheb_first_zero(int n, const double* dinv, const double* b, double theta, double* x, double* d, hipStream_t s,
                     int nv) {
  if (n) k_cheb_first_zero<double><<<dim3(blocks_for(n, kBlock), nv), kBlock, 0, s>>>(n, dinv, b, theta, x, d, n);
}
void cheb_first(int n, const double* dinv, const double* r, double theta, double* x, double* d, hipStream_t s,
                int nv) {
  if (n) k_cheb_first<double><<<dim3(blocks_for(n, kBlock), nv), kBlock, 0, s>>>(n, dinv, r, theta, x, d, n);
}
void cheb_first_zero32(int n, const float* dinv, const float* b, double theta, float* x, float* d, hipStream_t s,
                       int nv) {
  if (n) k_cheb_first_zero<float><<<dim3(blocks_for(n, kBlock), nv), kBlock, 0, s>>>(n, dinv, b, theta, x, d, n);
}
void cheb_first32(int n, const float* dinv, const float* r, double theta, float* x, float* d, hipStream_t s,
                  int nv) {
  if (n) k_cheb_first<float><<<dim3(blocks_for(n, kBlock), nv), kBlock, 0, s>>>(n, dinv, r, theta, x, d, n);
}
void dense_gemv(int n, const double* M, const double* b, double* y, hipStream_t s, int nv) {
  if (n) k_dense_gemv<double><<<dim3(blocks_for((long long)n * 64, kBlock), nv), kBlock, 0, s>>>(n, M, b, y, n);
}
void dense_gemv32(int n, const float* M, const float* b, float* y, hipStream_t s, int nv) {
  if (n) k_dense_gemv<float><<<dim3(blocks_for((long long)n * 64, kBlock), nv), kBlock, 0, s>>>(n, M, b, y, n);
}

int reduce_blocks(long long n) { return (int)std::max<long long>(1, std::min<long long>(kReduceBlocks, (n + kBlock - 1) / kBlock)); }

void dot(long long n, const double* a, const double* b, double* partials, double* out, hipStream_t s, int nv,
         int out_stride) {
  const int nb = reduce_blocks(n);
  k_dot_partials<<<dim3(nb, nv), kBlock, 0, s>>>((int)n, a, b, partials, n);
  k_reduce_partials<<<nv, 1024, 0, s>>>(nb, partials, out, 0, out_stride);
}
void grey_sum(long long n, const double* rho, double* partials, double* out, hipStream_t s) {
  const int nb = reduce_blocks(n);
  k_grey_partials<<<nb, kBlock, 0, s>>>((int)n, rho, partials);
  k_reduce_partials<<<1, 1024, 0, s>>>(nb, partials, out, 0, 0);
}
void reduce(int m, const double* partials, double* out, int take_sqrt, hipStream_t s, int nv, int out_stride) {
  k_reduce_partials<<<nv, 1024, 0, s>>>(m, partials, out, take_sqrt, out_stride);
}
void pcg_update(long long n, double* x, double* r, const double* p, const double* q, const double* scal, int i_rz,
                int i_pq, double* partials, double* out, hipStream_t s, int nv, const int* active) {
  const int nb = reduce_blocks(n);
  k_pcg_update<<<dim3(nb, nv), kBlock, 0, s>>>((int)n, x, r, p, q, scal, i_rz, i_pq, partials, n, active);
  k_reduce_partials<<<nv, 1024, 0, s>>>(nb, partials, out, 0, 16);
}
void precond_tail(long long n, const float* zc, const double* r, double* z, double* scal, int i_rz, int i_rzn,
                  int i_beta, bool flexible, double* partials, hipStream_t s, int nv) {
  const int nb = reduce_blocks(n);
  k_precond_partials<float><<<dim3(nb, nv), kBlock, 0, s>>>((int)n, zc, r, z, flexible ? 1 : 0, partials, n);
  k_precond_finish<<<nv, 1024, 0, s>>>(nb, partials, scal, i_rz, i_rzn, i_beta, flexible ? 1 : 0);
}
void cheb_first_zero_from64(int n, const float* dinv, const double* r, double theta, float* b, float* x, float* d,
                            hipStream_t s, int nv) {
  if (n) k_cheb_first_zero_from64<<<dim3(blocks_for(n, kBlock), nv), kBlock, 0, s>>>(n, dinv, r, theta, b, x, d, n);
}
void pcg_direction(long long n, double* p, const double* z, const double* scal, int i_new, int i_old, hipStream_t s,
                   int nv) {
  k_pcg_direction<<<dim3(reduce_blocks(n), nv), kBlock, 0, s>>>((int)n, p, z, scal, i_new, i_old, n);
}
void copy_scalar(double* scal, int from, int to, hipStream_t s) { k_copy_scalar<<<1, 1, 0, s>>>(scal, from, to); }
void flexible_beta(double* scal, int i_new, int i_cross, int i_old, int i_out, hipStream_t s) {
  k_flexible_beta<<<1, 1, 0, s>>>(scal, i_new, i_cross, i_old, i_out);
}
void to_float(long long n, const double* in, float* out, hipStream_t s) {
  if (n) k_convert<double, float><<<blocks_for(n, kBlock), kBlock, 0, s>>>((int)n, in, out);
}
void to_bf16(long long n, const double* in, uint16_t* out, hipStream_t s) {
  if (n) k_to_bf16<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, in, out);
}
void to_double(long long n, const float* in, double* out, hipStream_t s) {
  if (n) k_convert<float, double><<<blocks_for(n, kBlock), kBlock, 0, s>>>((int)n, in, out);
}
void axpy(long long n, double a, const double* x, double* y, hipStream_t s) {
  if (n) k_axpy<<<blocks_for(n, kBlock), kBlock, 0, s>>>((int)n, a, x, y);
}

void assemble_structured(int dim, const DCsr& K, const int* free_dofs, const double* scale, const double* ke0, int nx,
                         int ny, int nz, hipStream_t s, const unsigned* xyzc, const int* nodemap) {
  constexpr int LANES = 16;
  const int nb = grid8(K.rows, kBlock / LANES);
  if (K.rows == 0) return;
  const char* rows_env = std::getenv("SPARLAB_ROW_ASSEMBLY");
  const bool by_rows = rows_env && rows_env[0] == '1';
  if (dim == 3 && nodemap && !by_rows) {
    const long long nn = (long long)(nx + 1) * (ny + 1) * (nz + 1);
    k_assemble_hex8_nodes<<<grid8(nn, kBlock / 32), kBlock, 0, s>>>(nx + 1, ny + 1, nz + 1, nodemap, K.ptr.p, scale,
                                                                     ke0, K.val.p);
    return;
  }
  if (dim == 3) {
    if (xyzc)
      k_assemble_hex8<LANES, true><<<nb, kBlock, 0, s>>>(K.rows, K.ptr.p, K.col.p, K.val.p, free_dofs, xyzc, scale, ke0,
                                                         nx, ny, nz);
    else
      k_assemble_hex8<LANES, false><<<nb, kBlock, 0, s>>>(K.rows, K.ptr.p, K.col.p, K.val.p, free_dofs, xyzc, scale,
                                                          ke0, nx, ny, nz);
  } else if (xyzc) {
    k_assemble_quad4<LANES, true><<<nb, kBlock, 0, s>>>(K.rows, K.ptr.p, K.col.p, K.val.p, free_dofs, xyzc, scale, ke0,
                                                        nx, ny);
  } else {
    k_assemble_quad4<LANES, false><<<nb, kBlock, 0, s>>>(K.rows, K.ptr.p, K.col.p, K.val.p, free_dofs, xyzc, scale,
                                                         ke0, nx, ny);
  }
}

void sensitivity(int dim, int nelem, int nx, int ny, const double* ke0, const int* reduced, const double* u,
                 long long nfree, int ncases, const double* w, const double* fac, const double* dfac, double* energy,
                 double* dc, hipStream_t s) {
  const int nb = blocks_for(nelem, kBlock);
  if (!nb) return;
  if (dim == 3)
    k_sensitivity<3><<<nb, kBlock, 0, s>>>(nelem, nx, ny, ke0, reduced, u, nfree, ncases, w, fac, dfac, energy, dc);
  else
    k_sensitivity<2><<<nb, kBlock, 0, s>>>(nelem, nx, ny, ke0, reduced, u, nfree, ncases, w, fac, dfac, energy, dc);
}

void simp(int n, const double* filtered, double penalty, double emin, double* phys, double* fac, double* dfac,
          hipStream_t s) {
  if (n) k_simp<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, filtered, penalty, emin, phys, fac, dfac);
}
void project(int n, const double* in, double beta, double eta, double* out, double* dproj, hipStream_t s) {
  if (n) k_project<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, in, beta, eta, out, dproj);
}
void mul(int n, const double* a, const double* b, double* c, hipStream_t s) {
  if (n) k_mul<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, a, b, c);
}
void oc_candidate(int n, const double* x, const double* dc, const double* dv, const double* lower, const double* upper,
                  const unsigned char* passive, double lambda, double move, double damping, double* xn,
                  hipStream_t s) {
  if (n) k_oc_candidate<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, x, dc, dv, lower, upper, passive, lambda, move,
                                                                   damping, xn);
}
int oc_trial_blocks(int n) { return reduce_blocks(n); }
void oc_trial(int n, const double* x, const double* dc, const double* dv, const double* lower, const double* upper,
              const unsigned char* passive, double move, double damping, const double* w, double* state,
              int explicit_lambda, double lam, double target, double vol_tol, int max_bisections, double* xn,
              double* partials, hipStream_t s) {
  if (!n) return;
  const int nb = reduce_blocks(n);
  k_oc_trial<<<nb, kBlock, 0, s>>>(n, x, dc, dv, lower, upper, passive, move, damping, w, state, explicit_lambda, lam,
                                   xn, partials);
  k_oc_step<<<1, 1024, 0, s>>>(nb, partials, state, explicit_lambda, target, vol_tol, max_bisections);
}
void filter_volume(const DCsr& H, const double* x, const double* vol, double* partials, double* out, hipStream_t s) {
  const int sr = staged_rows(H);
  if (sr == 128 || sr == 256) {
    // 128 rows per block keeps the partial count within the reduction buffer for any n up to 2^19 blocks
    const int nb = blocks_for(H.rows, 128);
    k_spmv_staged<128, 1><<<nb, kBlock, 0, s>>>(H.rows, H.ptr.p, H.col.p, H.val.p, x, nullptr, vol, partials);
    k_reduce_partials<<<1, 1024, 0, s>>>(nb, partials, out, 0, 0);
    return;
  }
  const int nb = reduce_blocks(H.rows);
  k_filter_volume<<<nb, kBlock, 0, s>>>(H.rows, H.ptr.p, H.col.p, H.val.p, x, vol, partials);
  k_reduce_partials<<<1, 1024, 0, s>>>(nb, partials, out, 0, 0);
}
int filter_volume_blocks(const DCsr& H) { return blocks_for(H.rows, 128); }
void max_abs_diff(int n, const double* a, const double* b, double* partials, double* out, hipStream_t s) {
  const int nb = reduce_blocks(n);
  k_max_abs_diff<<<nb, kBlock, 0, s>>>(n, a, b, partials);
  k_reduce_max<<<1, kBlock, 0, s>>>(nb, partials, out);
}
void clamp(int n, double* x, const double* lo, const double* hi, hipStream_t s) {
  if (n) k_clamp<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, x, lo, hi);
}

}  // namespace k
}  // namespace slx
