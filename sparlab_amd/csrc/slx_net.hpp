// Warm-start network of the sparlab_hip library: launchers of kernels_net.hip
// and the host object that owns the network's parameters, activations and
// training state on the device.
//
// Architecture (node grid of a structured mesh, X x Y x Z, Z = 1 for Q4):
//   e1 = relu(conv(x, 16 -> 16))                 level 1
//   e2 = relu(conv(pool(e1), 16 -> 32))          level 2 (ceil(X/2) x ...)
//   e3 = relu(conv(pool(e2), 32 -> 64))          level 3
//   d2 = relu(conv(up(e3), 64 -> 32) + e2)       level 2
//   d1 = relu(conv(up(d2), 32 -> 16) + e1)       level 1
//   y  = conv(d1, 16 -> 16)                      level 1, first nout channels used
// 3x3x3 convolutions with bias; BF16 weights and activations, FP32
// accumulation, FP32 master weights and gradients, Adam.
#pragma once

#include <cstdint>
#include <random>
#include <vector>

#include "slx_internal.hpp"

namespace slx {
namespace knet {

struct ConvArgs {
  int X, Y, Z, V;
  int cin, cout;             // multiples of 16, at most 64
  const uint16_t* in;        // [V][cin] bf16
  const uint16_t* w;         // [cout][27][cin] bf16
  const float* bias;         // [cout] or null
  const float* add_f32;      // [V][cout] or null
  const uint16_t* add_bf16;  // [V][cout] or null
  int relu;
  float* out_f32;            // [V][cout] or null
  uint16_t* out_bf16;        // [V][cout] or null
};

struct WgradArgs {
  int X, Y, Z, V;
  int chunk;             // voxels per block (multiple of 64)
  const uint16_t* dz;    // [V][cout] bf16
  const uint16_t* in;    // [V][cin] bf16
  float* part;           // [chunks][cout][27][cin]
};

void conv(const ConvArgs& a, hipStream_t s);
int wgrad_chunks(int V, int chunk);
void conv_wgrad(const WgradArgs& a, int cout, int cin, hipStream_t s);
void sum_chunks(int chunks, long long n, const float* part, float* out, hipStream_t s);
int colsum_blocks(int V);
void colsum(int V, int C, const float* g, double* part, float* out, hipStream_t s);
void pool(int X, int Y, int Z, int C, const uint16_t* in, uint16_t* out, hipStream_t s);
void pool_adjoint(int X, int Y, int Z, int C, const float* gc, const float* add, float* gf, hipStream_t s);
void upsample(int X, int Y, int Z, int C, const uint16_t* coarse, uint16_t* fine, hipStream_t s);
void upsample_adjoint(int X, int Y, int Z, int C, const float* gf, float* gc, hipStream_t s);
void relu_adjoint(long long n, const float* g, const uint16_t* act, float* dz, uint16_t* dz16, hipStream_t s);
int mse_blocks(int V, int C);
void mse(int V, int C, int nout, const float* y, const float* t, const float* w, const double* wsum, float* gy,
         uint16_t* gy16, double* part, double* loss, hipStream_t s);
void sum_float(int n, const float* a, double* part, double* out, hipStream_t s);
void adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps, int step,
          hipStream_t s);
void prep_weights(int cout, int cin, const float* w32, uint16_t* w, uint16_t* wt, hipStream_t s);
void pack_input(int V, int used, int C, const float* in, uint16_t* out, hipStream_t s);
void features(int dim, int nx, int ny, int nz, const double* phys, const double* phys_prev, const double* fac,
              const int* reduced, const double* u, const double* r, const double* dprev, const double* scales, int C,
              uint16_t* out, float* weight, hipStream_t s);
void prediction(int dim, long long nodes, int C, const int* reduced, const float* y, double* d, hipStream_t s);
// out[i] = sqrt(sumsq[i] / n) for i < count (count <= 64)
void rms(int count, const double* sumsq, double n, double* out, hipStream_t s);
void target(int dim, long long nodes, const int* reduced, const double* delta, const double* scale, float* t,
            hipStream_t s);

}  // namespace knet

class WarmNet {
 public:
  static constexpr int kLayers = 6;
  static constexpr int kIn = 16;  // padded input channels
  struct Layer {
    int cin, cout, level;
    size_t w_off, b_off;  // offsets in the flat parameter vector
  };

  // (re)build for a node grid; parameters drawn from a fixed-seed generator
  void configure(int X, int Y, int Z, int nin, int nout, unsigned long long seed, hipStream_t s);
  bool ready() const { return V1 > 0; }
  int voxels() const { return V1; }
  int nin() const { return nin_; }
  int nout() const { return nout_; }
  size_t parameter_count() const { return nparams; }
  const std::vector<Layer>& layers() const { return layers_; }

  void set_parameters(const float* host, hipStream_t s);
  void get_parameters(float* host, hipStream_t s) const;
  void reset_optimizer();

  // x16 (the packed input, [V][16] bf16) -> y ([V][16] fp32)
  uint16_t* input() { return x16.p; }
  const float* output() const { return y.p; }
  void forward(hipStream_t s);
  // loss and gradients for target t ([V][nout] fp32, device) and optional voxel
  // weights w ([V], device; the loss is then normalised by their sum); the loss
  // stays on the device
  void backward(const float* t, const float* w, hipStream_t s);
  const float* gradients() const { return grad.p; }
  const double* loss_device() const { return loss.p; }
  void adam_step(float lr, hipStream_t s);

 private:
  void prep(hipStream_t s);
  int lvl_X[3] = {0, 0, 0}, lvl_Y[3] = {0, 0, 0}, lvl_Z[3] = {0, 0, 0}, lvl_V[3] = {0, 0, 0};
  int V1 = 0, nin_ = 0, nout_ = 0, step_ = 0;
  size_t nparams = 0;
  std::vector<Layer> layers_;
  DBuf<float> params, grad, adam_m, adam_v;
  DBuf<uint16_t> w16, wt16;  // per-layer bf16 operands, at the layer's w_off
  // activations (bf16) and outputs
  DBuf<uint16_t> x16, e1, p1, e2, p2, e3, u3, d2, u2, d1;
  DBuf<float> y;
  // backward buffers
  DBuf<float> gy, g_d1, g_z5, g_u2, g_d2, g_z4, g_u3, g_e3, g_z3, g_p2, g_e2, g_z2, g_p1, g_e1, g_z1;
  DBuf<uint16_t> gy16, gz5_16, gz4_16, gz3_16, gz2_16, gz1_16;
  DBuf<float> wpart;
  DBuf<double> dpart, loss, wsum;
};

// Learned warm start of the compliance solves: the state the context keeps.
struct WarmStart {
  slx_warmstart_options opt{};
  slx_warmstart_stats stats{};
  WarmNet net;
  bool model_grid = false;  // net configured for the context's node grid
  DBuf<double> phys_prev;
  bool have_prev = false;
  DBuf<double> scales;  // rms of u_prev, r_plain, the previous change, this change
  DBuf<double> uold, r, dir, kdir, ksec, x0;
  double gain_net = 0.0, gain_sec = 0.0;  // energy gains of each direction alone (current solve)
  double beta_sec = 0.0;                  // K-optimal step along the previous change (current solve)
  // the dot products of a solve's prediction and statistics, on the device:
  // [0] u.u [1] r.r [2] s.s [3] s.r [4] s.Ks [5] d.r [6] d.Kd [7] d.Ks (s the
  // previous change, d the network's direction), [8] r0.r0 after the start is
  // formed, [9] [10] the K-norm errors of u_prev and of the start, [11] the sum
  // of squares of the training target
  DBuf<double> dots;
  // statistics of the last solve whose values [8]-[10] are still on the device,
  // folded in by the next read (the next prediction or a statistics query)
  bool pending_r0 = false, pending_energy = false;
  double pending_rr = 0.0, pending_gain_net = 0.0, pending_gain_sec = 0.0;
  DBuf<double> dprev;           // last displacement change per load case (ncases x nfree)
  std::vector<char> have_dprev;
  std::vector<DBuf<uint16_t>> xs;  // replay buffer: packed inputs
  std::vector<DBuf<float>> ts, ws; // targets and loss weights
  int count = 0, next = 0, solves = 0;
  int pending = -1;  // replay slot whose target the current solve provides
  bool predicted = false;  // the current solve started from x0 = u_prev + a d
  bool uold_valid = false; // uold holds the current solve's u_prev
  std::mt19937_64 rng;
};

// hooks of the compliance evaluation (called around each load case's solve)
void warm_before_solve(Context& c, int load_case);
void warm_after_solve(Context& c, int load_case);
void warm_after_evaluate(Context& c);
// folds the statistics still on the device into stats (one read)
void warm_flush_stats(Context& c);

}  // namespace slx
