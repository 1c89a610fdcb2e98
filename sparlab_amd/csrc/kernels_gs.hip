// Symmetric Gauss-Seidel smoothing (src/fem/Multigrid.cpp:788-806) by
// dependency levels. The reference sweeps the rows in order (forward before
// the coarse correction, backward after it), each row reading the unknowns of
// the rows already updated in that sweep. Grouping the rows into levels where
// no row reads an unknown of its own level (amg.cpp gs_schedule) lets every
// row of a level run at once and still see exactly the values the sequential
// sweep gives it: the rows it depends on sit in lower levels and are done, the
// others in higher levels and are not yet touched. One launch per level.
#include <hip/hip_runtime.h>

#include "slx_kernels.hpp"

namespace slx {
namespace k {
namespace {

__device__ __forceinline__ double value(double v) { return v; }
__device__ __forceinline__ float value(float v) { return v; }
__device__ __forceinline__ float value(uint16_t v) { return __uint_as_float((unsigned)v << 16); }

// x_i += (b_i - sum_k a_ik x_k) dinv_i for the rows of one level, summed in the
// row's order with separate products and sums (the restatement's rounding; this
// file builds without floating-point contraction); nv cases side by side
template <typename V, typename T>
__global__ void __launch_bounds__(256) k_gs_level(int count, const int* __restrict__ rows, const int* __restrict__ ptr,
                                                  const int* __restrict__ col, const V* __restrict__ val,
                                                  const T* __restrict__ dinv, const T* __restrict__ b, T* x,
                                                  long long n, int nv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  const int k0 = ptr[i], k1 = ptr[i + 1];
  for (int v = 0; v < nv; ++v) {
    const T* bv = b + (long long)v * n;
    T* xv = x + (long long)v * n;
    T s = bv[i];
    for (int k = k0; k < k1; ++k) s -= (T)value(val[k]) * xv[col[k]];
    xv[i] += s * dinv[i];
  }
}

__global__ void __launch_bounds__(256) k_fill32(long long n, float v, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = v;
}

}  // namespace

void fill32(long long n, float v, float* out, hipStream_t s) {
  if (n > 0) k_fill32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, v, out);
}

void gs_sweep(const DCsr& A, const int* rows, const std::vector<int>& level_ptr, const double* dinv, const double* b,
              double* x, hipStream_t s, int nv) {
  for (size_t L = 0; L + 1 < level_ptr.size(); ++L) {
    const int cnt = level_ptr[L + 1] - level_ptr[L];
    if (cnt <= 0) continue;
    k_gs_level<double, double><<<(cnt + 255) / 256, 256, 0, s>>>(cnt, rows + level_ptr[L], A.ptr.p, A.col.p, A.val.p,
                                                                  dinv, b, x, (long long)A.rows, nv);
  }
}

void gs_sweep32(const DCsr& A, const int* rows, const std::vector<int>& level_ptr, const float* dinv, const float* b,
                float* x, hipStream_t s, int nv) {
  for (size_t L = 0; L + 1 < level_ptr.size(); ++L) {
    const int cnt = level_ptr[L + 1] - level_ptr[L];
    if (cnt <= 0) continue;
    const int g = (cnt + 255) / 256;
    if (A.val16.p)
      k_gs_level<uint16_t, float><<<g, 256, 0, s>>>(cnt, rows + level_ptr[L], A.ptr.p, A.col.p, A.val16.p, dinv, b, x,
                                                    (long long)A.rows, nv);
    else
      k_gs_level<float, float><<<g, 256, 0, s>>>(cnt, rows + level_ptr[L], A.ptr.p, A.col.p, A.val32.p, dinv, b, x,
                                                 (long long)A.rows, nv);
  }
}

}  // namespace k
}  // namespace slx
