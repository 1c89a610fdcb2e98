// Matrix-free K_ff products of a structured Hex8 model:
//   (K x)_i = sum over the elements e around node i of s_e (Ke0 x_e) restricted to i's rows,
// the same operator the assembled K_ff holds (K_ff = sum_e s_e Ke0 scattered,
// src/fem/Assembler.cpp:182-214), applied without reading it. One thread per
// node; a block stages the input of a 256-node tile (8 x 8 x 4 by default, or
// one of the other shapes of Tiles) plus a one-node halo and the factors of the
// tile's elements in LDS. For each of the 8 elements around a node the node's
// local corner is fixed by the element's offset, so the Ke0 coefficients a wave
// needs are uniform across it. A node's sums do not depend on the tile shape.
//
// Epilogues fuse what the callers do with the product: y = Kx with the partial
// sums of x.y (CG), r = b - Kx (residuals), the Chebyshev step, the Lanczos
// step on D^-1/2 K D^-1/2, and the smoothing halves of the level-0 transfers
// P = (I - omega D^-1 K) P-hat and R = P^T = P-hat^T (I - omega K D^-1). Vectors are in the free-DOF layout; prescribed DOFs
// read as zero and are never written.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "slx_mf.hpp"

namespace slx {
namespace kmf {
namespace {

constexpr int kThreads = 256;

// a tile of SX x SY x SZ nodes, one per thread
template <int SX, int SY, int SZ>
struct Tile {
  static_assert(SX * SY * SZ == kThreads, "one node per thread");
  static constexpr int TX = SX, TY = SY, TZ = SZ;
  static constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;  // node halo
  static constexpr int EX = TX + 1, EY = TY + 1, EZ = TZ + 1;  // elements touching the tile
};
// the shapes MfArgs::tile selects (kTileShapes of them)
using Tile0 = Tile<8, 8, 4>;
using Tile1 = Tile<16, 4, 4>;
using Tile2 = Tile<32, 4, 2>;
using Tile3 = Tile<16, 8, 2>;
static_assert(kTileShapes == 4, "one Tile per shape");

template <typename S>
int tile_blocks(int X, int Y, int Z) {
  const int n = ((X + S::TX - 1) / S::TX) * ((Y + S::TY - 1) / S::TY) * ((Z + S::TZ - 1) / S::TZ);
  return (n + 7) / 8 * 8;
}

// VTK hexahedron corner of local offsets (StructuredMesh.cpp:224-234) and its inverse
__device__ __forceinline__ constexpr int corner(int lx, int ly, int lz) {
  return lz * 4 + (ly ? (lx ? 2 : 3) : (lx ? 1 : 0));
}
__device__ __forceinline__ constexpr int corner_x(int b) { return (b & 3) == 1 || (b & 3) == 2; }
__device__ __forceinline__ constexpr int corner_y(int b) { return (b & 3) >= 2; }
__device__ __forceinline__ constexpr int corner_z(int b) { return b >= 4; }

// bijective block -> tile remap: consecutive tiles (which share halos) on one XCD
__device__ __forceinline__ int xcd_tile(int b, int nb) { return (b % 8) * (nb / 8) + b / 8; }

template <typename T>
__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) s += sh[w];
  return s;
}

// four components per node in LDS (the fourth is padding) so that one vector
// read fetches a node's three
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kElemPad = 416;  // the tile's 405 elements in 26 column tiles of 16

// MX (FP32 only): the element products Y_e = Ke0 u_e of the tile's elements run on
// the matrix cores as Y = Ke0 U (v_mfma_f32_16x16x32_bf16, FP32 accumulation),
// each operand split into bfloat16 high and low parts (Ke0_h U_h + Ke0_h U_l +
// Ke0_l U_h: about 16 significant bits, the error of the dropped low x low term);
// the scaled products go to LDS by (element row, element) and each node sums its
// 8 elements' rows. Lane layout of the instruction: the first operand's lane
// (m, g) = (lane % 16, lane / 16) holds row m, columns 8g..8g+7; the second
// operand's holds rows 8g..8g+7 of column m; the result's lane holds rows
// 4g..4g+3 of column m.
// A node's free DOFs are consecutive in the free numbering, so the packed map
// nodemap[node] = (first free index << 3) | (mask of its free components) locates
// all three with one load. Component c sits at base + popcount(mask below c); a
// prescribed one reads entry base (or 0) and is zeroed, so no branch splits the loads.
template <typename T>
__device__ __forceinline__ void load3(const T* __restrict__ v, int m, T& a0, T& a1, T& a2) {
  const int base = m >> 3, mask = m & 7;
  const int i1 = base + (mask & 1), i2 = i1 + ((mask >> 1) & 1);
  const T v0 = v[base], v1 = v[(mask & 2) ? i1 : base], v2 = v[(mask & 4) ? i2 : base];
  a0 = (mask & 1) ? v0 : T(0);
  a1 = (mask & 2) ? v1 : T(0);
  a2 = (mask & 4) ? v2 : T(0);
}
// stores the free components; a node with all three free takes one 3-wide store
template <typename T>
__device__ __forceinline__ void store3(T* __restrict__ v, int m, const T (&o)[3]) {
  const int base = m >> 3, mask = m & 7;
  if (mask == 7) {
    T* p = v + base;
    p[0] = o[0];
    p[1] = o[1];
    p[2] = o[2];
  } else {
    int i = base;
    if (mask & 1) v[i++] = o[0];
    if (mask & 2) v[i++] = o[1];
    if (mask & 4) v[i] = o[2];
  }
}

// FP32 vector-FMA variants are held to 80 registers (6 waves per SIMD): the step is
// bound by how many tiles' staging and products overlap, and the few registers
// that spill cost less than the lost occupancy
template <typename T, int MODE, bool MX, typename S>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && !MX ? 6 : 1)))
k_mf_hex8(MfArgs<T> a) {
  static_assert(!MX || sizeof(T) == 4, "the matrix-core path is FP32");
  constexpr int TX = S::TX, TY = S::TY, TZ = S::TZ, HX = S::HX, HY = S::HY, HZ = S::HZ;
  constexpr int EX = S::EX, EY = S::EY, EZ = S::EZ;
  static_assert(!MX || EX * EY * EZ <= kElemPad, "the matrix-core path pads the tile's elements to kElemPad");
  using V4 = typename Vec4<T>::type;
  __shared__ V4 s_u[HZ * HY * HX];
  __shared__ T s_s[EZ * EY * EX];
  __shared__ double s_red[kThreads / 64];
  __shared__ float s_y[MX ? 24 * kElemPad : 1];
  const int ntx = (a.X + TX - 1) / TX, nty = (a.Y + TY - 1) / TY, ntz = (a.Z + TZ - 1) / TZ;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= ntx * nty * ntz) {
    if ((MODE == 1 || (MODE == 4 && a.partials)) && threadIdx.x == 0) a.partials[blockIdx.x] = 0.0;
    return;
  }
  const int x0 = (tile % ntx) * TX, y0 = ((tile / ntx) % nty) * TY, z0 = (tile / (ntx * nty)) * TZ;
  const int tx = threadIdx.x % TX, ty = (threadIdx.x / TX) % TY, tz = threadIdx.x / (TX * TY);
  const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
  const bool inside = gx < a.X && gy < a.Y && gz < a.Z;
  // the node's map first: the epilogue's operands load before the element
  // products, which hide their latency
  const int me = inside ? a.nodemap[((long long)gz * a.Y + gy) * a.X + gx] : 0;
  // Stage the input (scaled for the Lanczos mode) over the tile and its halo,
  // and the factors of the tile's elements. Every load is issued before any
  // is used (indices, then values): entries outside the mesh load node 0 and
  // are masked, so no branch splits the loads.
  constexpr int kHalo = HZ * HY * HX, kHaloIt = (kHalo + kThreads - 1) / kThreads;
  constexpr int kElem = EZ * EY * EX, kElemIt = (kElem + kThreads - 1) / kThreads;
  const int nex = a.X - 1, ney = a.Y - 1, nez = a.Z - 1;  // elements per direction
  int hm[kHaloIt];
#pragma unroll
  for (int it = 0; it < kHaloIt; ++it) {
    const int h = threadIdx.x + it * kThreads;
    const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
    const bool ok = h < kHalo && gx >= 0 && gx < a.X && gy >= 0 && gy < a.Y && gz >= 0 && gz < a.Z;
    const int m = a.nodemap[ok ? (((long long)gz * a.Y + gy) * a.X + gx) : 0];
    hm[it] = ok ? m : 0;
  }
  T sv[kElemIt];
#pragma unroll
  for (int it = 0; it < kElemIt; ++it) {
    const int h = threadIdx.x + it * kThreads;
    const int ex = h % EX, ey = (h / EX) % EY, ez = h / (EX * EY);
    const int gx = x0 - 1 + ex, gy = y0 - 1 + ey, gz = z0 - 1 + ez;
    const bool ok = h < kElem && gx >= 0 && gx < nex && gy >= 0 && gy < ney && gz >= 0 && gz < nez;
    const T v = a.scale[ok ? ((long long)gz * ney + gy) * nex + gx : 0];
    sv[it] = ok ? v : T(0);
  }
  T hu[kHaloIt][3];
#pragma unroll
  for (int it = 0; it < kHaloIt; ++it) {
    load3(a.x, hm[it], hu[it][0], hu[it][1], hu[it][2]);
    if (MODE == 4 || MODE == 5) {
      T s0, s1, s2;
      load3(a.in_scale, hm[it], s0, s1, s2);
      hu[it][0] *= s0;
      hu[it][1] *= s1;
      hu[it][2] *= s2;
    }
  }
#pragma unroll
  for (int it = 0; it < kHaloIt; ++it) {
    const int h = threadIdx.x + it * kThreads;
    if (h < kHalo) {
      V4 v;
      v.x = hu[it][0];
      v.y = hu[it][1];
      v.z = hu[it][2];
      v.w = 0;
      s_u[h] = v;
    }
  }
#pragma unroll
  for (int it = 0; it < kElemIt; ++it) {
    const int h = threadIdx.x + it * kThreads;
    if (h < kElem) s_s[h] = sv[it];
  }
  // (prescribed components read as zero: no branches around the loads)
  T px[3] = {0, 0, 0}, pb[3] = {0, 0, 0}, pd[3] = {0, 0, 0}, ps[3] = {0, 0, 0};
  if (MODE == 1 || MODE == 3 || MODE == 4 || MODE == 5 || MODE == 6 || MODE == 7) load3(a.x, me, px[0], px[1], px[2]);
  if (MODE == 2 || MODE == 3 || MODE == 7) load3(a.b, me, pb[0], pb[1], pb[2]);
  if (MODE == 3 || MODE == 6 || MODE == 7) load3(a.dinv, me, pd[0], pd[1], pd[2]);
  if (MODE == 3 || MODE == 6) load3(a.xs, me, ps[0], ps[1], ps[2]);
  if (MODE == 4) {
    load3(a.in_scale, me, pd[0], pd[1], pd[2]);
    load3(a.vprev, me, ps[0], ps[1], ps[2]);
  }
  __syncthreads();
  T acc0 = 0, acc1 = 0, acc2 = 0;
  if constexpr (MX) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    // Ke0 rows mt * 16 + m, columns 8g..8g+7 (zero past 24), high and low parts
    bf16x8 ah[2], al[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int row = mt * 16 + m;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = 8 * g + q;
        const bool in = row < 24 && k < 24;
        const float v = in ? (float)a.ke[in ? row * 24 + k : 0] : 0.f;
        const __bf16 h = (__bf16)v;
        ah[mt][q] = h;
        al[mt][q] = (__bf16)(v - (float)h);
      }
    }
    const float* su = reinterpret_cast<const float*>(s_u);
    for (int nt = wave; nt < kElemPad / 16; nt += kThreads / 64) {
      const int e = nt * 16 + m;  // the tile's element (column of U)
      bf16x8 bh, bl;
      const bool live = e < kElem && g < 3;
      const int ex = e % EX, ey = (e / EX) % EY, ez = e / (EX * EY);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        // element DOF k = 8g + q: corner k / 3, component k % 3
        const int k = 8 * g + q, b = k / 3, c = k - 3 * (k / 3);
        // padded columns (e >= 405) and rows (k >= 24) read node 0 and are masked
        const int node = live ? ((ez + corner_z(b)) * HY + (ey + corner_y(b))) * HX + (ex + corner_x(b)) : 0;
        const float v = live ? su[4 * node + c] : 0.f;
        const __bf16 h = (__bf16)v;
        bh[q] = h;
        bl[q] = (__bf16)(v - (float)h);
      }
      const float se = e < kElem ? (float)s_s[e] : 0.f;
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mt], bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt], bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt], bh, acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = mt * 16 + 4 * g + v;
          if (row < 24) s_y[row * kElemPad + e] = acc[v] * se;
        }
      }
    }
    __syncthreads();
    if (inside) {
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int ox = (o & 1) - 1, oy = ((o >> 1) & 1) - 1, oz = (o >> 2) - 1;
        const int e = ((tz + oz + 1) * EY + (ty + oy + 1)) * EX + (tx + ox + 1);
        const int ca = corner(-ox, -oy, -oz);
        acc0 += s_y[(3 * ca) * kElemPad + e];
        acc1 += s_y[(3 * ca + 1) * kElemPad + e];
        acc2 += s_y[(3 * ca + 2) * kElemPad + e];
      }
    }
  } else if (inside) {
    // The node's 3 x 3 x 3 neighbourhood is read from LDS once, a z layer at a
    // time, and the 8 element products take their corners from registers
    // (27 reads instead of 64). The products run in the same order as a
    // corner-by-corner loop, so the sums are the same.
    // two z layers live at a time: [0] the lower, [1] the upper of the
    // element layer being summed
    T nx[2][3][3], ny[2][3][3], nz[2][3][3];
    auto load_layer = [&](int slot, int dz) {
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const V4 u = s_u[((tz + dz) * HY + (ty + dy)) * HX + (tx + dx)];
          nx[slot][dy][dx] = u.x;
          ny[slot][dy][dx] = u.y;
          nz[slot][dy][dx] = u.z;
        }
    };
    load_layer(0, 0);
    load_layer(1, 1);
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = (o & 1) - 1, oy = ((o >> 1) & 1) - 1, oz = (o >> 2) - 1;
      if (o == 4) {
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
          for (int dx = 0; dx < 3; ++dx) {
            nx[0][dy][dx] = nx[1][dy][dx];
            ny[0][dy][dx] = ny[1][dy][dx];
            nz[0][dy][dx] = nz[1][dy][dx];
          }
        load_layer(1, 2);
      }
      // element at (node + offset); tile-local element and halo coordinates coincide
      const int ex = tx + ox + 1, ey = ty + oy + 1, ez = tz + oz + 1;
      const T se = s_s[(ez * EY + ey) * EX + ex];
      if (se == T(0)) continue;  // outside the mesh
      const int ca = corner(-ox, -oy, -oz);
      const T* kr = a.ke + (3 * ca) * 24;
      T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        // fused multiply-adds: the element product differs from the assembled
        // operator's sum in rounding only
#pragma clang fp contract(fast)
        const int dz = corner_z(b), dy = oy + 1 + corner_y(b), dx = ox + 1 + corner_x(b);
        const T ux = nx[dz][dy][dx], uy = ny[dz][dy][dx], uz = nz[dz][dy][dx];
        const int cb = 3 * b;
        t0 += kr[cb] * ux;
        t0 += kr[cb + 1] * uy;
        t0 += kr[cb + 2] * uz;
        t1 += kr[24 + cb] * ux;
        t1 += kr[25 + cb] * uy;
        t1 += kr[26 + cb] * uz;
        t2 += kr[48 + cb] * ux;
        t2 += kr[49 + cb] * uy;
        t2 += kr[50 + cb] * uz;
      }
      acc0 += se * t0;
      acc1 += se * t1;
      acc2 += se * t2;
    }
  }
  double dot = 0.0;
  {
    const T acc[3] = {acc0, acc1, acc2};
    const int mask = me & 7;
    T o1[3], o2[3], o3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool free_c = (mask >> c) & 1;
      const T y = acc[c];
      if (MODE == 0) {
        o1[c] = y;
      } else if (MODE == 1) {
        o1[c] = y;
        if (free_c) dot += (double)px[c] * (double)y;
      } else if (MODE == 2) {
        o1[c] = pb[c] - y;
      } else if (MODE == 3) {
        // r_out = r_in - K d_in, d_out = c1 d_in + c2 D^-1 r_out, x += d_out (x input = d_in)
        const double r = (double)pb[c] - (double)y;
        const double dn = a.c1 * (double)px[c] + a.c2 * (double)pd[c] * r;
        o1[c] = (T)r;
        o2[c] = (T)dn;
        o3[c] = (T)((double)ps[c] + dn);
      } else if (MODE == 4) {
        // Lanczos: w = s (K (s v)) - beta vprev, with the partials of w.v
        const double beta = a.beta ? *a.beta : 0.0;
        const double wv = (double)y * (double)pd[c] - beta * (double)ps[c];
        o1[c] = (T)wv;
        if (free_c) dot += wv * (double)px[c];
      } else if (MODE == 7) {
        // post-smoothing residual and the smoother's start from x:
        // r = b - K x, d = D^-1 r / theta, x2 = x + d (k_cheb_first's rounding)
        const T r = pb[c] - y;
        const double di = (double)pd[c] * (double)r / a.c1;
        o1[c] = r;
        o2[c] = (T)di;
        o3[c] = (T)((double)px[c] + di);
      } else if (MODE == 5) {
        // smoothed restriction, first half: y = (I - omega K D^-1) x
        o1[c] = (T)((double)px[c] - a.c1 * (double)y);
      } else {
        // smoothed prolongation, second half: xs += (I - omega D^-1 K) x
        o3[c] = (T)((double)ps[c] + (double)px[c] - a.c1 * (double)pd[c] * (double)y);
      }
    }
    if (MODE != 6) store3(a.y, me, o1);
    if (MODE == 3 || MODE == 7) store3(a.d_out, me, o2);
    if (MODE == 3 || MODE == 6 || MODE == 7) store3(a.xs, me, o3);
  }
  if (MODE == 1 || (MODE == 4 && a.partials)) {
    const double s = block_sum<T>(dot, s_red);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
  }
}

template <typename T, int MODE, bool MX = false, typename S = Tile0>
void launch(const MfArgs<T>& a, hipStream_t s) {
  k_mf_hex8<T, MODE, MX, S><<<tile_blocks<S>(a.X, a.Y, a.Z), kThreads, 0, s>>>(a);
}

// modes without block partials, in the shape a.tile names
template <typename T, int MODE>
void launch_tiled(const MfArgs<T>& a, hipStream_t s) {
  switch (a.tile) {
    case 1: launch<T, MODE, false, Tile1>(a, s); break;
    case 2: launch<T, MODE, false, Tile2>(a, s); break;
    case 3: launch<T, MODE, false, Tile3>(a, s); break;
    default: launch<T, MODE, false, Tile0>(a, s); break;
  }
}

// inv_diag of K_ff from the element factors: for each free DOF (node, k), 1 /
// sum over the node's elements, in ascending element order, of s_e Ke0(a,k; a,k)
// (a the node's corner in e) -- the assembled diagonal entry bit for bit, since
// the assembly sums the same products in the same order. One thread per node.
__global__ void __launch_bounds__(256) k_mf_inv_diag(int X, int Y, int Z, const int* __restrict__ reduced,
                                                     const double* __restrict__ scale, const double* __restrict__ ke,
                                                     double* __restrict__ dinv, int* __restrict__ bad_row) {
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)X * Y * Z;
  if (node >= nn) return;
  const int i = (int)(node % X), j = (int)((node / X) % Y), k = (int)(node / ((long long)X * Y));
  const int ex = X - 1, ey = Y - 1, ez = Z - 1;  // elements per direction
  // VTK corner of offset (dx, dy, dz) from the element's first node
  constexpr int corner[2][2][2] = {{{0, 1}, {3, 2}}, {{4, 5}, {7, 6}}};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int row = reduced[node * 3 + c];
    if (row < 0) continue;
    double d = 0.0;
    for (int zz = k - 1; zz <= k; ++zz) {
      if (zz < 0 || zz >= ez) continue;
      for (int yy = j - 1; yy <= j; ++yy) {
        if (yy < 0 || yy >= ey) continue;
        for (int xx = i - 1; xx <= i; ++xx) {
          if (xx < 0 || xx >= ex) continue;
          const long long e = ((long long)zz * ey + yy) * ex + xx;
          const int a = corner[k - zz][j - yy][i - xx];
          d += scale[e] * ke[(a * 3 + c) * 24 + a * 3 + c];
        }
      }
    }
    if (!(d > 0.0)) atomicMin(bad_row, row);
    dinv[row] = 1.0 / d;
  }
}

}  // namespace

void inv_diag(int X, int Y, int Z, const int* reduced, const double* scale, const double* ke, double* dinv,
              int* bad_row, hipStream_t s) {
  const long long nn = (long long)X * Y * Z;
  if (nn) k_mf_inv_diag<<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(X, Y, Z, reduced, scale, ke, dinv, bad_row);
}

bool node_map(const std::vector<int>& reduced, std::vector<int>& map) {
  const size_t nn = reduced.size() / 3;
  map.assign(nn, 0);
  for (size_t n = 0; n < nn; ++n) {
    int base = -1, mask = 0;
    for (int k = 0; k < 3; ++k) {
      const int r = reduced[3 * n + (size_t)k];
      if (r < 0) continue;
      if (base < 0) base = r;
      mask |= 1 << k;
    }
    if (mask == 0) continue;
    if (base >= (1 << 28)) return false;
    map[n] = (base << 3) | mask;
  }
  return true;
}

int blocks(int X, int Y, int Z) { return tile_blocks<Tile0>(X, Y, Z); }

void tile_shape(int tile, int* dims) {
  switch (tile) {
    case 1: dims[0] = Tile1::TX, dims[1] = Tile1::TY, dims[2] = Tile1::TZ; break;
    case 2: dims[0] = Tile2::TX, dims[1] = Tile2::TY, dims[2] = Tile2::TZ; break;
    case 3: dims[0] = Tile3::TX, dims[1] = Tile3::TY, dims[2] = Tile3::TZ; break;
    default: dims[0] = Tile0::TX, dims[1] = Tile0::TY, dims[2] = Tile0::TZ; break;
  }
}

// opt-in (SPARLAB_MF_MFMA=1): at 128 x 64 x 64 the matrix-core variant's LDS (51 KB
// per block) leaves 3 waves per SIMD, and the step, bound by its staging latency,
// ran slower than the vector-FMA kernel
bool matrix_cores() {
  const char* e = std::getenv("SPARLAB_MF_MFMA");
  return e && e[0] == '1';
}

template <typename T>
void apply(const MfArgs<T>& a, int mode, hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    // the FP32 V-cycle's modes take the matrix-core element products
    if (matrix_cores() && (mode == 2 || mode == 3 || mode == 5 || mode == 6)) {
      switch (mode) {
        case 2: launch<T, 2, true>(a, s); break;
        case 3: launch<T, 3, true>(a, s); break;
        case 5: launch<T, 5, true>(a, s); break;
        default: launch<T, 6, true>(a, s); break;
      }
      return;
    }
  }
  switch (mode) {
    case 0: launch<T, 0>(a, s); break;
    case 1: launch<T, 1>(a, s); break;
    case 2: launch_tiled<T, 2>(a, s); break;
    case 3: launch_tiled<T, 3>(a, s); break;
    case 4: launch<T, 4>(a, s); break;
    case 5: launch_tiled<T, 5>(a, s); break;
    case 6: launch_tiled<T, 6>(a, s); break;
    default: launch_tiled<T, 7>(a, s); break;
  }
}

template void apply<double>(const MfArgs<double>&, int, hipStream_t);
template void apply<float>(const MfArgs<float>&, int, hipStream_t);

}  // namespace kmf
}  // namespace slx
