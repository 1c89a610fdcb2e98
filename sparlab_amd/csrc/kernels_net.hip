// Kernels of the warm-start network: a 3-level 3-D convolutional encoder-decoder
// on the node grid of a structured mesh, BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
//
// Tensors are channels-last, [voxel][channel], voxel = (z * Y + y) * X + x, with
// channel counts that are multiples of 16. A 3x3x3 convolution is an implicit
// GEMM: M = 16 voxels per wave, N = 16 output channels per tile, K = 27 taps x
// C_in input channels, logical K index L = tap * C_in + c with
// tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1). Each lane group of the MFMA
// supplies 8 consecutive logical K values for both operands, so the product
// does not depend on how the instruction orders K internally. Out-of-grid
// neighbours read as zero (zero padding); the output has the input's size.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "slx_net.hpp"

namespace slx {
namespace knet {
namespace {

constexpr int kWave = 64;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                             // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }

__device__ __forceinline__ bool neighbour(int x, int y, int z, int tap, int X, int Y, int Z, long long& nv) {
  const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
  if (xx < 0 || xx >= X || yy < 0 || yy >= Y || zz < 0 || zz >= Z) return false;
  nv = ((long long)zz * Y + yy) * X + xx;
  return true;
}

inline int blocks_for(long long n, int per) { return (int)((n + per - 1) / per); }

// ---------------------------------------------------------------------------
// convolution: out[v][co] = epi(sum_{tap, c} in[v + tap][c] * w[co][tap * cin + c])
// epi: + bias[co], + add_f32[v][co], + add_bf16[v][co], relu; fp32 and/or bf16 store.
// One wave per 16 voxels, all output channels (NT tiles of 16).
// ---------------------------------------------------------------------------
template <int NT>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int v0 = (blockIdx.x * 4 + wave) * 16;
  if (v0 >= a.V) return;  // whole wave
  const int m = lane & 15, g = lane >> 4;
  const int v = v0 + m;
  const bool vin = v < a.V;
  int x = 0, y = 0, z = 0;
  if (vin) {
    x = v % a.X;
    const int t = v / a.X;
    y = t % a.Y;
    z = t / a.Y;
  }
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int K = 27 * a.cin;
  const int chunks = (K + 31) / 32;
  for (int q = 0; q < chunks; ++q) {
    const int L = q * 32 + 8 * g;
    bf16x8 av = zero8();
    bf16x8 bv[NT];
    if (L < K) {
      const int tap = L / a.cin, ch = L - tap * a.cin;
      long long nv;
      if (vin && neighbour(x, y, z, tap, a.X, a.Y, a.Z, nv)) av = load8(a.in + nv * a.cin + ch);
#pragma unroll
      for (int t = 0; t < NT; ++t) bv[t] = load8(a.w + (size_t)(t * 16 + m) * K + L);
    } else {
#pragma unroll
      for (int t = 0; t < NT; ++t) bv[t] = zero8();
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[t], acc[t], 0, 0, 0);
  }
  // lane holds rows 4g + i (voxels) of column m (output channel) of each tile
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = t * 16 + m;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int vv = v0 + 4 * g + i;
      if (vv >= a.V) continue;
      const size_t o = (size_t)vv * a.cout + co;
      float r = acc[t][i];
      if (a.bias) r += a.bias[co];
      if (a.add_f32) r += a.add_f32[o];
      if (a.add_bf16) r += bf2f(a.add_bf16[o]);
      if (a.relu) r = r > 0.f ? r : 0.f;
      if (a.out_f32) a.out_f32[o] = r;
      if (a.out_bf16) a.out_bf16[o] = f2bf(r);
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradient of one tap: part[c][co][tap][ci] = sum over the voxels of chunk c
// of dz[v][co] * in[v + tap][ci]. M = output channels, N = input channels,
// K = voxels. Each wave stages 64 voxels of dz and of the shifted input in LDS,
// transposed to [channel][voxel], so an operand is one 16-byte LDS read.
// ---------------------------------------------------------------------------
constexpr int kStride = 72;  // LDS row stride in bf16 (64 voxels + pad), 16-byte aligned rows

template <int COT, int CIT>
__global__ void __launch_bounds__(256) k_conv_wgrad(WgradArgs a) {
  extern __shared__ uint16_t s_net[];
  constexpr int cout = 16 * COT, cin = 16 * CIT;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int m = lane & 15, g = lane >> 4;
  const int c = blockIdx.x, tap = blockIdx.y;
  uint16_t* s_dz = s_net + (size_t)wave * (cout + cin) * kStride;
  uint16_t* s_x = s_dz + cout * kStride;
  f32x4 acc[COT][CIT];
#pragma unroll
  for (int i = 0; i < COT; ++i)
#pragma unroll
    for (int j = 0; j < CIT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int begin = c * a.chunk;
  const int end = min(a.V, begin + a.chunk);
  const int tiles = (a.chunk + 63) / 64;
  for (int tb = 0; tb < tiles; tb += 4) {
    const int v = begin + (tb + wave) * 64 + lane;
    const bool vin = (tb + wave) < tiles && v < end;
    // stage dz[v][:] and in[v + tap][:] (zero outside the grid or the chunk)
    long long nv = -1;
    if (vin) {
      const int x = v % a.X, t = v / a.X;
      if (!neighbour(x, t % a.Y, t / a.Y, tap, a.X, a.Y, a.Z, nv)) nv = -1;
    }
#pragma unroll
    for (int c8 = 0; c8 < cout; c8 += 8) {
      u32x4 d = u32x4{0u, 0u, 0u, 0u};
      if (vin) d = *reinterpret_cast<const u32x4*>(a.dz + (size_t)v * cout + c8);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        s_dz[(c8 + 2 * j) * kStride + lane] = (uint16_t)(d[j] & 0xffffu);
        s_dz[(c8 + 2 * j + 1) * kStride + lane] = (uint16_t)(d[j] >> 16);
      }
    }
#pragma unroll
    for (int c8 = 0; c8 < cin; c8 += 8) {
      u32x4 d = u32x4{0u, 0u, 0u, 0u};
      if (nv >= 0) d = *reinterpret_cast<const u32x4*>(a.in + nv * cin + c8);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        s_x[(c8 + 2 * j) * kStride + lane] = (uint16_t)(d[j] & 0xffffu);
        s_x[(c8 + 2 * j + 1) * kStride + lane] = (uint16_t)(d[j] >> 16);
      }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      bf16x8 A[COT], B[CIT];
#pragma unroll
      for (int i = 0; i < COT; ++i) A[i] = load8(s_dz + (i * 16 + m) * kStride + h * 32 + 8 * g);
#pragma unroll
      for (int j = 0; j < CIT; ++j) B[j] = load8(s_x + (j * 16 + m) * kStride + h * 32 + 8 * g);
#pragma unroll
      for (int i = 0; i < COT; ++i)
#pragma unroll
        for (int j = 0; j < CIT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // sum the four waves in order through LDS, then write this chunk's partial
  float* s_red = reinterpret_cast<float*>(s_net);
  constexpr int per_wave = COT * CIT * 256;
#pragma unroll
  for (int i = 0; i < COT; ++i)
#pragma unroll
    for (int j = 0; j < CIT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) s_red[wave * per_wave + ((i * CIT + j) * 4 + r) * 64 + lane] = acc[i][j][r];
  __syncthreads();
  for (int e = threadIdx.x; e < per_wave; e += blockDim.x) {
    const float s = ((s_red[e] + s_red[per_wave + e]) + s_red[2 * per_wave + e]) + s_red[3 * per_wave + e];
    const int ln = e % 64, r = (e / 64) % 4, ij = e / 256;
    const int i = ij / CIT, j = ij % CIT;
    const int co = i * 16 + 4 * (ln >> 4) + r, ci = j * 16 + (ln & 15);
    a.part[(((size_t)c * cout + co) * 27 + tap) * cin + ci] = s;
  }
}

// out[i] = sum_c part[c][i], c ascending
__global__ void k_sum_chunks(int chunks, long long n, const float* __restrict__ part, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += part[(size_t)c * n + i];
  out[i] = s;
}

// per-channel sums of g[V][C] (bias gradients): partials per block of 256 voxels, then ordered
__global__ void k_colsum_partials(int V, int C, const float* __restrict__ g, double* __restrict__ part) {
  const int c = threadIdx.x % C, lanes = blockDim.x / C, sub = threadIdx.x / C;
  __shared__ double s[256];
  double acc = 0.0;
  const int v0 = blockIdx.x * 256;
  for (int v = v0 + sub; v < min(V, v0 + 256); v += lanes) acc += (double)g[(size_t)v * C + c];
  s[threadIdx.x] = acc;
  __syncthreads();
  if (sub == 0) {
    double t = 0.0;
    for (int k = 0; k < lanes; ++k) t += s[k * C + c];
    part[(size_t)blockIdx.x * C + c] = t;
  }
}

// out[c] = sum_b part[b][c]: one block of 256 per channel, fixed summation order
__global__ void k_colsum_final(int blocks, int C, const double* __restrict__ part, float* __restrict__ out) {
  __shared__ double s[256];
  const int c = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < blocks; b += 256) a += part[(size_t)b * C + c];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = (float)s[0];
}

// ---------------------------------------------------------------------------
// resolution changes: 2x2x2 average pooling (partial windows at odd edges) and
// nearest upsampling, with their adjoints. Sums run dz, dy, dx ascending.
// ---------------------------------------------------------------------------
__global__ void k_pool(int X, int Y, int Z, int C, const uint16_t* __restrict__ in, uint16_t* __restrict__ out) {
  const int X2 = (X + 1) / 2, Y2 = (Y + 1) / 2, Z2 = (Z + 1) / 2;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)X2 * Y2 * Z2 * C) return;
  const int c = (int)(idx % C);
  const long long cv = idx / C;
  const int cx = (int)(cv % X2), cy = (int)((cv / X2) % Y2), cz = (int)(cv / ((long long)X2 * Y2));
  float s = 0.f;
  int cnt = 0;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int x = 2 * cx + dx, y = 2 * cy + dy, z = 2 * cz + dz;
        if (x >= X || y >= Y || z >= Z) continue;
        s += bf2f(in[(((long long)z * Y + y) * X + x) * C + c]);
        ++cnt;
      }
  out[idx] = f2bf(s / (float)cnt);
}

// gf[v][c] = gc[parent][c] / count(parent) (+ add[v][c])
__global__ void k_pool_adjoint(int X, int Y, int Z, int C, const float* __restrict__ gc, const float* __restrict__ add,
                               float* __restrict__ gf) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)X * Y * Z * C) return;
  const int c = (int)(idx % C);
  const long long v = idx / C;
  const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((long long)X * Y));
  const int X2 = (X + 1) / 2, Y2 = (Y + 1) / 2;
  const int cx = x / 2, cy = y / 2, cz = z / 2;
  const int cnt = (min(X, 2 * cx + 2) - 2 * cx) * (min(Y, 2 * cy + 2) - 2 * cy) * (min(Z, 2 * cz + 2) - 2 * cz);
  float r = gc[(((long long)cz * Y2 + cy) * X2 + cx) * C + c] / (float)cnt;
  if (add) r += add[idx];
  gf[idx] = r;
}

// fine[v][c] = coarse[parent][c]
__global__ void k_upsample(int X, int Y, int Z, int C, const uint16_t* __restrict__ coarse, uint16_t* __restrict__ fine) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)X * Y * Z * C) return;
  const int c = (int)(idx % C);
  const long long v = idx / C;
  const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((long long)X * Y));
  const int X2 = (X + 1) / 2, Y2 = (Y + 1) / 2;
  fine[idx] = coarse[((((long long)(z / 2)) * Y2 + y / 2) * X2 + x / 2) * C + c];
}

// gc[cv][c] = sum of gf over the children of cv
__global__ void k_upsample_adjoint(int X, int Y, int Z, int C, const float* __restrict__ gf, float* __restrict__ gc) {
  const int X2 = (X + 1) / 2, Y2 = (Y + 1) / 2, Z2 = (Z + 1) / 2;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)X2 * Y2 * Z2 * C) return;
  const int c = (int)(idx % C);
  const long long cv = idx / C;
  const int cx = (int)(cv % X2), cy = (int)((cv / X2) % Y2), cz = (int)(cv / ((long long)X2 * Y2));
  float s = 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int x = 2 * cx + dx, y = 2 * cy + dy, z = 2 * cz + dz;
        if (x >= X || y >= Y || z >= Z) continue;
        s += gf[(((long long)z * Y + y) * X + x) * C + c];
      }
  gc[idx] = s;
}

// dz = g * (act > 0); fp32 and bf16 copies
__global__ void k_relu_adjoint(long long n, const float* __restrict__ g, const uint16_t* __restrict__ act,
                               float* __restrict__ dz, uint16_t* __restrict__ dz16) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r = bf2f(act[i]) > 0.f ? g[i] : 0.f;
  dz[i] = r;
  dz16[i] = f2bf(r);
}

// weighted mean-squared error over the first nout channels,
// loss = sum_v w_v sum_c (y - t)^2 / (W nout) with W = sum_v w_v (w = 1, W = V without weights):
// gy = 2 w_v (y - t) / (W nout); per-block partial sums of w_v (y - t)^2
__global__ void k_mse(int V, int C, int nout, const float* __restrict__ y, const float* __restrict__ t,
                      const float* __restrict__ w, const double* __restrict__ wsum, float* __restrict__ gy,
                      uint16_t* __restrict__ gy16, double* __restrict__ part) {
  __shared__ double s[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (i < (long long)V * C) {
    const int c = (int)(i % C);
    const long long v = i / C;
    float gr = 0.f;
    if (c < nout) {
      const float d = y[i] - t[v * nout + c];
      const float wv = w ? w[v] : 1.f;
      const float W = w ? (float)*wsum : (float)V;
      e2 = (double)wv * ((double)d * (double)d);
      gr = 2.f * wv * d / (W * (float)nout);
    }
    gy[i] = gr;
    gy16[i] = f2bf(gr);
  }
  s[threadIdx.x] = e2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// *out = scale * sum(part) / (den ? *den : 1); one block of 256, fixed summation order
__global__ void k_sum_double(int n, const double* __restrict__ part, double scale, const double* __restrict__ den,
                             double* __restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) a += part[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0] * scale / (den ? *den : 1.0);
}

// per-block partial sums of a float vector (256 per block)
__global__ void k_sum_float_partials(int n, const float* __restrict__ a, double* __restrict__ part) {
  __shared__ double s[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  s[threadIdx.x] = i < n ? (double)a[i] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Adam (Kingma & Ba 2015) with bias correction, FP32 master weights
__global__ void k_adam(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 operands of one layer: w[co][tap][ci] (forward) and wt[ci][tap'][co] = w[co][26 - tap'][ci]
// (the adjoint convolution is a convolution with flipped taps and swapped channels)
__global__ void k_prep_weights(int cout, int cin, const float* __restrict__ w32, uint16_t* __restrict__ w,
                               uint16_t* __restrict__ wt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)cout * 27 * cin) return;
  const int ci = (int)(i % cin), tap = (int)((i / cin) % 27), co = (int)(i / (27LL * cin));
  const uint16_t h = f2bf(w32[i]);
  w[i] = h;
  wt[((size_t)ci * 27 + (26 - tap)) * cout + co] = h;
}

// fp32 [V][cin_used] -> bf16 [V][C] (zero-padded channels)
__global__ void k_pack_input(int V, int used, int C, const float* __restrict__ in, uint16_t* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)V * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  out[i] = c < used ? f2bf(in[v * used + c]) : (uint16_t)0;
}

// ---------------------------------------------------------------------------
// structured-mesh features and results (node grid X = nx + 1, ...; element
// e = (k * ny + j) * nx + i; DOF = node * dim + component)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double node_mean(const double* ev, int i, int j, int k, int nx, int ny, int nz) {
  double s = 0.0;
  int cnt = 0;
  for (int kk = max(k - 1, 0); kk <= min(k, max(nz - 1, 0)); ++kk)
    for (int jj = max(j - 1, 0); jj <= min(j, ny - 1); ++jj)
      for (int ii = max(i - 1, 0); ii <= min(i, nx - 1); ++ii) {
        s += ev[((size_t)kk * ny + jj) * nx + ii];
        ++cnt;
      }
  return s / cnt;
}

// channels: 0 node density, 1 its change, then dim each: u_prev / s0, r / s1 and the
// previous displacement change / s2 (zero when dprev is null);
// weight[v]: mean stiffness factor of the node's elements (the loss weight)
// out[i] = sqrt(sumsq[i] / n): root-mean-square scales from sums of squares on the device
__global__ void k_rms(int count, const double* __restrict__ sumsq, double n, double* __restrict__ out) {
  const int i = threadIdx.x;
  if (i < count) out[i] = sqrt(sumsq[i] / n);
}

__global__ void k_features(int dim, int nx, int ny, int nz, const double* __restrict__ phys,
                           const double* __restrict__ phys_prev, const double* __restrict__ fac,
                           const int* __restrict__ reduced, const double* __restrict__ u,
                           const double* __restrict__ r, const double* __restrict__ dprev,
                           const double* __restrict__ scales, int C, uint16_t* __restrict__ out,
                           float* __restrict__ weight) {
  const int X = nx + 1, Y = ny + 1, Z = dim == 3 ? nz + 1 : 1;
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (long long)X * Y * Z) return;
  const int i = (int)(v % X), j = (int)((v / X) % Y), k = (int)(v / ((long long)X * Y));
  const int nzz = dim == 3 ? nz : 1;
  const double rho = node_mean(phys, i, j, k, nx, ny, nzz);
  const double drho = rho - node_mean(phys_prev, i, j, k, nx, ny, nzz);
  weight[v] = (float)node_mean(fac, i, j, k, nx, ny, nzz);
  uint16_t* o = out + v * C;
  o[0] = f2bf((float)rho);
  o[1] = f2bf((float)drho);
  const double su = scales[0] > 0.0 ? 1.0 / scales[0] : 0.0;
  const double sr = scales[1] > 0.0 ? 1.0 / scales[1] : 0.0;
  const double sd = scales[2] > 0.0 ? 1.0 / scales[2] : 0.0;
  for (int d = 0; d < dim; ++d) {
    const int f = reduced[v * dim + d];
    o[2 + d] = f2bf(f >= 0 ? (float)(u[f] * su) : 0.f);
    o[2 + dim + d] = f2bf(f >= 0 ? (float)(r[f] * sr) : 0.f);
    o[2 + 2 * dim + d] = f2bf(f >= 0 && dprev ? (float)(dprev[f] * sd) : 0.f);
  }
  for (int c = 2 + 3 * dim; c < C; ++c) o[c] = 0;
}

// d[f] = y[node][k] for every free DOF f = reduced[node * dim + k]
__global__ void k_prediction(int dim, long long nodes, int C, const int* __restrict__ reduced,
                             const float* __restrict__ y, double* __restrict__ d) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes * dim) return;
  const int f = reduced[i];
  if (f < 0) return;
  d[f] = (double)y[(i / dim) * C + i % dim];
}

// t[node][d] = delta[f] / *scale (0 at prescribed DOFs)
__global__ void k_target(int dim, long long nodes, const int* __restrict__ reduced, const double* __restrict__ delta,
                         const double* __restrict__ scale, float* __restrict__ t) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes * dim) return;
  const int f = reduced[i];
  const double sc = *scale > 0.0 ? 1.0 / *scale : 0.0;
  t[i] = f >= 0 ? (float)(delta[f] * sc) : 0.f;
}

}  // namespace

// ---------------------------------------------------------------------------
void conv(const ConvArgs& a, hipStream_t s) {
  if (a.V <= 0) return;
  if (a.cin % 16 || a.cout % 16 || a.cout > 64 || a.cin > 64) fail(SLX_SOLVER_ERROR, "net: unsupported channel count");
  const int grid = blocks_for(a.V, 64);
  switch (a.cout / 16) {
    case 1: k_conv<1><<<grid, 256, 0, s>>>(a); break;
    case 2: k_conv<2><<<grid, 256, 0, s>>>(a); break;
    case 3: k_conv<3><<<grid, 256, 0, s>>>(a); break;
    default: k_conv<4><<<grid, 256, 0, s>>>(a); break;
  }
}

int wgrad_chunks(int V, int chunk) { return blocks_for(V, chunk); }

void conv_wgrad(const WgradArgs& a, int cout, int cin, hipStream_t s) {
  if (a.V <= 0) return;
  const dim3 grid((unsigned)wgrad_chunks(a.V, a.chunk), 27);
  const size_t stage = (size_t)4 * (cout + cin) * kStride * sizeof(uint16_t);
  const size_t red = (size_t)4 * (cout / 16) * (cin / 16) * 256 * sizeof(float);
  const size_t lds = stage > red ? stage : red;
#define SLX_WG(CO, CI) \
  if (cout == 16 * CO && cin == 16 * CI) return (void)k_conv_wgrad<CO, CI><<<grid, 256, lds, s>>>(a);
  SLX_WG(1, 1) SLX_WG(2, 1) SLX_WG(1, 2) SLX_WG(4, 2) SLX_WG(2, 4) SLX_WG(2, 2) SLX_WG(4, 4) SLX_WG(1, 4) SLX_WG(4, 1)
#undef SLX_WG
  fail(SLX_SOLVER_ERROR, "net: unsupported channel pair for the weight gradient");
}

void sum_chunks(int chunks, long long n, const float* part, float* out, hipStream_t s) {
  if (n) k_sum_chunks<<<blocks_for(n, 256), 256, 0, s>>>(chunks, n, part, out);
}

int colsum_blocks(int V) { return blocks_for(V, 256); }

void colsum(int V, int C, const float* g, double* part, float* out, hipStream_t s) {
  if (C > 256 || 256 % C) fail(SLX_SOLVER_ERROR, "net: unsupported channel count for a bias gradient");
  const int nb = colsum_blocks(V);
  k_colsum_partials<<<nb, 256, 0, s>>>(V, C, g, part);
  k_colsum_final<<<C, 256, 0, s>>>(nb, C, part, out);
}

void pool(int X, int Y, int Z, int C, const uint16_t* in, uint16_t* out, hipStream_t s) {
  const long long n = (long long)((X + 1) / 2) * ((Y + 1) / 2) * ((Z + 1) / 2) * C;
  k_pool<<<blocks_for(n, 256), 256, 0, s>>>(X, Y, Z, C, in, out);
}

void pool_adjoint(int X, int Y, int Z, int C, const float* gc, const float* add, float* gf, hipStream_t s) {
  const long long n = (long long)X * Y * Z * C;
  k_pool_adjoint<<<blocks_for(n, 256), 256, 0, s>>>(X, Y, Z, C, gc, add, gf);
}

void upsample(int X, int Y, int Z, int C, const uint16_t* coarse, uint16_t* fine, hipStream_t s) {
  const long long n = (long long)X * Y * Z * C;
  k_upsample<<<blocks_for(n, 256), 256, 0, s>>>(X, Y, Z, C, coarse, fine);
}

void upsample_adjoint(int X, int Y, int Z, int C, const float* gf, float* gc, hipStream_t s) {
  const long long n = (long long)((X + 1) / 2) * ((Y + 1) / 2) * ((Z + 1) / 2) * C;
  k_upsample_adjoint<<<blocks_for(n, 256), 256, 0, s>>>(X, Y, Z, C, gf, gc);
}

void relu_adjoint(long long n, const float* g, const uint16_t* act, float* dz, uint16_t* dz16, hipStream_t s) {
  if (n) k_relu_adjoint<<<blocks_for(n, 256), 256, 0, s>>>(n, g, act, dz, dz16);
}

int mse_blocks(int V, int C) { return blocks_for((long long)V * C, 256); }

void mse(int V, int C, int nout, const float* y, const float* t, const float* w, const double* wsum, float* gy,
         uint16_t* gy16, double* part, double* loss, hipStream_t s) {
  const int nb = mse_blocks(V, C);
  k_mse<<<nb, 256, 0, s>>>(V, C, nout, y, t, w, wsum, gy, gy16, part);
  k_sum_double<<<1, 256, 0, s>>>(nb, part, w ? 1.0 / nout : 1.0 / ((double)V * nout), w ? wsum : nullptr, loss);
}

void sum_float(int n, const float* a, double* part, double* out, hipStream_t s) {
  const int nb = blocks_for(n, 256);
  k_sum_float_partials<<<nb, 256, 0, s>>>(n, a, part);
  k_sum_double<<<1, 256, 0, s>>>(nb, part, 1.0, nullptr, out);
}

void adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps, int step,
          hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  if (n) k_adam<<<blocks_for(n, 256), 256, 0, s>>>(n, p, g, m, v, lr, b1, b2, eps, c1, c2);
}

void prep_weights(int cout, int cin, const float* w32, uint16_t* w, uint16_t* wt, hipStream_t s) {
  const long long n = (long long)cout * 27 * cin;
  k_prep_weights<<<blocks_for(n, 256), 256, 0, s>>>(cout, cin, w32, w, wt);
}

void pack_input(int V, int used, int C, const float* in, uint16_t* out, hipStream_t s) {
  k_pack_input<<<blocks_for((long long)V * C, 256), 256, 0, s>>>(V, used, C, in, out);
}

void features(int dim, int nx, int ny, int nz, const double* phys, const double* phys_prev, const double* fac,
              const int* reduced, const double* u, const double* r, const double* dprev, const double* scales, int C,
              uint16_t* out, float* weight, hipStream_t s) {
  const long long V = (long long)(nx + 1) * (ny + 1) * (dim == 3 ? nz + 1 : 1);
  k_features<<<blocks_for(V, 256), 256, 0, s>>>(dim, nx, ny, nz, phys, phys_prev, fac, reduced, u, r, dprev, scales,
                                                C, out, weight);
}

void rms(int count, const double* sumsq, double n, double* out, hipStream_t s) {
  if (count > 0) k_rms<<<1, 64, 0, s>>>(count, sumsq, n, out);
}

void prediction(int dim, long long nodes, int C, const int* reduced, const float* y, double* d, hipStream_t s) {
  k_prediction<<<blocks_for(nodes * dim, 256), 256, 0, s>>>(dim, nodes, C, reduced, y, d);
}

void target(int dim, long long nodes, const int* reduced, const double* delta, const double* scale, float* t,
            hipStream_t s) {
  k_target<<<blocks_for(nodes * dim, 256), 256, 0, s>>>(dim, nodes, reduced, delta, scale, t);
}

}  // namespace knet
}  // namespace slx
