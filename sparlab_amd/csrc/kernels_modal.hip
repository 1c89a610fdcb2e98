// HIP kernels of the free-vibration (modal) analysis on structured grids:
// the mass product of the free DOFs and its diagonal, the diagonal of an
// assembled CSR matrix, the q x q Gram products of two n x q column-major
// blocks, the block combination Y = X R and the eigenpair residual norms.
// None of them is on the SIMP step: they run q times (or once) per subspace
// iteration, next to the q AMG-PCG solves that dominate it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "slx_kernels.hpp"

namespace slx {
namespace k {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxQ = 64;  // widest block (subspace size) the combination supports

// 1-D consistent mass of a linear two-node element, as a fraction of its
// length: 1/3 between a node and itself, 1/6 between the two nodes. The mass of
// a box element (Q4 rectangle, Hex8 brick) is the product of its axes' factors
// times rho * volume.
__device__ __forceinline__ double mass_factor(int offset) { return offset == 0 ? (1.0 / 3.0) : (1.0 / 6.0); }

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  return t;
}

// One thread per grid node: y(free DOFs of the node) = rho V sum over its
// elements e of s_e Me(node, m) x(m) (consistent) or s_e / npe x(node)
// (lumped, row-sum). diag != 0: the diagonal of that matrix instead (x unused).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_mass(int nx, int ny, int nz, double rho_v, int lumped, int diag,
                                                 const double* __restrict__ scale, const int* __restrict__ reduced,
                                                 const double* __restrict__ x, double* __restrict__ y) {
  const int nnx = nx + 1, nny = ny + 1, nnz = DIM == 3 ? nz + 1 : 1;
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= (long long)nnx * nny * nnz) return;
  const int i = (int)(node % nnx), j = (int)((node / nnx) % nny), kk = DIM == 3 ? (int)(node / ((long long)nnx * nny)) : 0;
  int out[DIM];
  bool any = false;
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    out[c] = reduced[node * DIM + c];
    any = any || out[c] >= 0;
  }
  if (!any) return;
  constexpr int npe = DIM == 3 ? 8 : 4;
  double acc[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) acc[c] = 0.0;
  for (int ez = (DIM == 3 ? kk - 1 : 0); ez <= (DIM == 3 ? kk : 0); ++ez) {
    if (DIM == 3 && (ez < 0 || ez >= nz)) continue;
    for (int ey = j - 1; ey <= j; ++ey) {
      if (ey < 0 || ey >= ny) continue;
      for (int ex = i - 1; ex <= i; ++ex) {
        if (ex < 0 || ex >= nx) continue;
        const long long e = ((long long)ez * ny + ey) * nx + ex;
        const double s = scale ? scale[e] : 1.0;
        if (s == 0.0) continue;
        if (diag || lumped) {
          double w = lumped ? 1.0 / npe : (DIM == 3 ? 1.0 / 27.0 : 1.0 / 9.0);
          w *= s;
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            if (diag) {
              acc[c] += w;
            } else if (out[c] >= 0) {
              acc[c] += w * x[out[c]];
            }
          }
          continue;
        }
        // the element's corners m, each weighted by the product of the axis factors
        for (int cz = 0; cz < (DIM == 3 ? 2 : 1); ++cz)
          for (int cy = 0; cy < 2; ++cy)
            for (int cx = 0; cx < 2; ++cx) {
              const int mi = ex + cx, mj = ey + cy, mk = DIM == 3 ? ez + cz : 0;
              double w = s * mass_factor(mi - i) * mass_factor(mj - j);
              if (DIM == 3) w *= mass_factor(mk - kk);
              const long long m = ((long long)mk * nny + mj) * nnx + mi;
#pragma unroll
              for (int c = 0; c < DIM; ++c) {
                const int r = reduced[m * DIM + c];
                if (r >= 0) acc[c] += w * x[r];
              }
            }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c)
    if (out[c] >= 0) y[out[c]] = rho_v * acc[c];
}

__global__ void __launch_bounds__(kBlock) k_csr_diag(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                                                     const double* __restrict__ val, double* __restrict__ d) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double v = 0.0;
  for (int p = ptr[r]; p < ptr[r + 1]; ++p)
    if (col[p] == r) v += val[p];
  d[r] = v;
}

// partials[pair * nb + b], pair = i * q + j: block b's share of A_i . B_j
__global__ void __launch_bounds__(kBlock) k_gram_partials(int n, int q, const double* __restrict__ A,
                                                          const double* __restrict__ B, double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  const int pair = blockIdx.y, i = pair / q, j = pair % q;
  const double* a = A + (size_t)i * n;
  const double* b = B + (size_t)j * n;
  double v = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) v += a[r] * b[r];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) partials[(size_t)pair * gridDim.x + blockIdx.x] = v;
}

// out[pair] = sum of that pair's nb partials, in index order (one thread per pair)
__global__ void k_sum_pairs(int npairs, int nb, const double* __restrict__ partials, double* __restrict__ out) {
  const int pair = blockIdx.x * blockDim.x + threadIdx.x;
  if (pair >= npairs) return;
  double v = 0.0;
  for (int b = 0; b < nb; ++b) v += partials[(size_t)pair * nb + b];
  out[pair] = v;
}

// Y[:, j] = sum_i X[:, i] R[i, j] for j < m (R column-major, q x m), one thread per row
__global__ void __launch_bounds__(kBlock) k_block_combine(int n, int q, int m, const double* __restrict__ X,
                                                          const double* __restrict__ R, double* __restrict__ Y) {
  __shared__ double rs[kMaxQ * kMaxQ];
  for (int t = threadIdx.x; t < q * m; t += blockDim.x) rs[t] = R[t];
  __syncthreads();
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double acc[kMaxQ];
#pragma unroll
  for (int jj = 0; jj < kMaxQ; ++jj) acc[jj] = 0.0;
  for (int i = 0; i < q; ++i) {
    const double xi = X[(size_t)i * n + r];
#pragma unroll
    for (int jj = 0; jj < kMaxQ; ++jj)
      if (jj < m) acc[jj] += xi * rs[jj * q + i];
  }
#pragma unroll
  for (int jj = 0; jj < kMaxQ; ++jj)
    if (jj < m) Y[(size_t)jj * n + r] = acc[jj];
}

// partials[(2 i) * nb + b] = block b's share of |KX_i - mu_i MX_i|^2,
// partials[(2 i + 1) * nb + b] of |mu_i MX_i|^2
__global__ void __launch_bounds__(kBlock) k_eig_residual_partials(int n, const double* __restrict__ KX,
                                                                  const double* __restrict__ MX,
                                                                  const double* __restrict__ mu,
                                                                  double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.y;
  const double l = mu[i];
  const double* kx = KX + (size_t)i * n;
  const double* mx = MX + (size_t)i * n;
  double v = 0.0, w = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const double lm = l * mx[r];
    const double d = kx[r] - lm;
    v += d * d;
    w += lm * lm;
  }
  v = block_sum(v, sh);
  __syncthreads();
  w = block_sum(w, sh);
  if (threadIdx.x == 0) {
    partials[(size_t)(2 * i) * gridDim.x + blockIdx.x] = v;
    partials[(size_t)(2 * i + 1) * gridDim.x + blockIdx.x] = w;
  }
}

__global__ void __launch_bounds__(kBlock) k_scale_copy(int n, double a, const double* __restrict__ x,
                                                       double* __restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) y[r] = a * x[r];
}

int gram_blocks(int n) { return std::max(1, std::min(64, (n + kBlock - 1) / kBlock)); }

}  // namespace

int modal_partials(int n, int q) { return gram_blocks(n) * std::max(q * q, 2 * q); }

void mass_apply(int dim, int nx, int ny, int nz, double rho_v, bool lumped, bool diagonal, const double* scale,
                const int* reduced, const double* x, double* y, hipStream_t s) {
  const long long nn = (long long)(nx + 1) * (ny + 1) * (dim == 3 ? nz + 1 : 1);
  const int nb = (int)((nn + kBlock - 1) / kBlock);
  if (dim == 3)
    k_mass<3><<<nb, kBlock, 0, s>>>(nx, ny, nz, rho_v, lumped ? 1 : 0, diagonal ? 1 : 0, scale, reduced, x, y);
  else
    k_mass<2><<<nb, kBlock, 0, s>>>(nx, ny, nz, rho_v, lumped ? 1 : 0, diagonal ? 1 : 0, scale, reduced, x, y);
}

void csr_diagonal(const DCsr& A, double* d, hipStream_t s) {
  if (A.rows) k_csr_diag<<<(A.rows + kBlock - 1) / kBlock, kBlock, 0, s>>>(A.rows, A.ptr.p, A.col.p, A.val.p, d);
}

void gram(int n, int q, const double* A, const double* B, double* partials, double* out, hipStream_t s) {
  const int nb = gram_blocks(n);
  k_gram_partials<<<dim3(nb, q * q), kBlock, 0, s>>>(n, q, A, B, partials);
  k_sum_pairs<<<(q * q + 63) / 64, 64, 0, s>>>(q * q, nb, partials, out);
}

void block_combine(int n, int q, int m, const double* X, const double* R, double* Y, hipStream_t s) {
  if (q > kMaxQ || m > kMaxQ) throw Error(SLX_CONFIG_ERROR, "block_combine: at most 64 columns");
  k_block_combine<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, q, m, X, R, Y);
}

void eig_residuals(int n, int m, const double* KX, const double* MX, const double* mu, double* partials, double* out,
                   hipStream_t s) {
  const int nb = gram_blocks(n);
  k_eig_residual_partials<<<dim3(nb, m), kBlock, 0, s>>>(n, KX, MX, mu, partials);
  k_sum_pairs<<<(2 * m + 63) / 64, 64, 0, s>>>(2 * m, nb, partials, out);
}

void scale_copy(int n, double a, const double* x, double* y, hipStream_t s) {
  if (n) k_scale_copy<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, a, x, y);
}

}  // namespace k
}  // namespace slx
