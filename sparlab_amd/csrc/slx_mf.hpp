// Matrix-free K_ff products of structured Hex8 models (kernels_mf.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace slx {
namespace kmf {

// One launch computes y = K_ff x in one of these modes:
//   0  y = K x
//   1  y = K x and partials[block] = sum of x.y over the block (blocks() partials)
//   2  y = b - K x
//   3  Chebyshev step with d_in = x, r_in = b: y = r_out = r_in - K d_in,
//      d_out = c1 d_in + c2 dinv r_out, xs += d_out
//   4  Lanczos step: y = in_scale (K (in_scale x)) - (*beta) vprev (beta null: 0),
//      and when partials is set partials[block] = sum of y.x over the block
//   5  restriction smoothing: y = x - c1 K (in_scale x)   (in_scale = D^-1, c1 = omega)
//   6  prolongation smoothing: xs += x - c1 dinv (K x)
//   7  residual and the smoother's start: y = r = b - K x, d_out = dinv r / c1
//      (c1 = theta), xs = x + d_out (xs another buffer than x)
// X, Y, Z are the node counts per direction; nodemap[node] = (free index of the
// node's first free DOF << 3) | (bit k set when DOF node * 3 + k is free), 0 for a
// node with every DOF prescribed (node_map() builds it); scale holds the element
// factors s_e and ke the 24 x 24 element matrix (node-major DOF order, VTK corner
// numbering). tile picks the launch's tile shape (tile_shape()) for modes 2, 3, 5,
// 6 and 7; modes 0, 1 and 4 and the matrix-core variant take shape 0, so the
// blocks() partials keep one grouping. Every shape gives the same values bit for bit.
constexpr int kTileShapes = 4;
template <typename T>
struct MfArgs {
  int X, Y, Z;
  int tile = 0;
  const int* nodemap;
  const T* scale;
  const T* ke;
  const T* x;
  T* y;
  const T* b = nullptr;
  const T* dinv = nullptr;
  T* d_out = nullptr;
  T* xs = nullptr;
  double c1 = 0.0, c2 = 0.0;
  const T* in_scale = nullptr;
  const T* vprev = nullptr;
  const double* beta = nullptr;
  double* partials = nullptr;
};

int blocks(int X, int Y, int Z);
// nodes per direction of tile shape t (0: 8 x 8 x 4, 1: 16 x 4 x 4, 2: 32 x 4 x 2, 3: 16 x 8 x 2)
void tile_shape(int t, int* dims);
// the packed node map from DOF -> free index (-1 prescribed), 3 DOFs per node;
// false when a free index does not fit in 28 bits
bool node_map(const std::vector<int>& reduced, std::vector<int>& map);
// 1 / diag(K_ff) from the element factors (the assembled diagonal bit for bit);
// a non-positive entry lowers *bad_row to its row
void inv_diag(int X, int Y, int Z, const int* reduced, const double* scale, const double* ke, double* dinv,
              int* bad_row, hipStream_t s);
template <typename T>
void apply(const MfArgs<T>& a, int mode, hipStream_t s);

}  // namespace kmf
}  // namespace slx
