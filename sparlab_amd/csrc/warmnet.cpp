// Host side of the warm-start network: parameter layout, forward and backward
// passes as kernel sequences on one stream, Adam. No counterpart exists in the
// reference; the network feeds the initial guess of the CG solve that
// ComplianceObjective warm-starts from the previous solution
// (src/topopt/Sensitivity.cpp:169-176).
#include <cmath>
#include <random>

#include "slx_net.hpp"

namespace slx {

namespace {
constexpr int kChunk = 2048;  // voxels per weight-gradient block
}

void WarmNet::configure(int X, int Y, int Z, int nin, int nout, unsigned long long seed, hipStream_t s) {
  if (X < 1 || Y < 1 || Z < 1) fail(SLX_CONFIG_ERROR, "warm-start network: empty node grid");
  if (nin < 1 || nin > kIn || nout < 1 || nout > 16) fail(SLX_CONFIG_ERROR, "warm-start network: 1..16 input and output channels");
  lvl_X[0] = X, lvl_Y[0] = Y, lvl_Z[0] = Z;
  for (int l = 1; l < 3; ++l) {
    lvl_X[l] = (lvl_X[l - 1] + 1) / 2;
    lvl_Y[l] = (lvl_Y[l - 1] + 1) / 2;
    lvl_Z[l] = (lvl_Z[l - 1] + 1) / 2;
  }
  for (int l = 0; l < 3; ++l) {
    const long long v = (long long)lvl_X[l] * lvl_Y[l] * lvl_Z[l];
    if (v > (1LL << 30) / 64) fail(SLX_CONFIG_ERROR, "warm-start network: node grid too large");
    lvl_V[l] = (int)v;
  }
  V1 = lvl_V[0];
  nin_ = nin;
  nout_ = nout;
  step_ = 0;
  const int spec[kLayers][3] = {{16, 16, 0}, {16, 32, 1}, {32, 64, 2}, {64, 32, 1}, {32, 16, 0}, {16, 16, 0}};
  layers_.clear();
  size_t off = 0;
  for (auto& sp : spec) {
    Layer L{sp[0], sp[1], sp[2], off, 0};
    off += (size_t)L.cout * 27 * L.cin;
    L.b_off = off;
    off += (size_t)L.cout;
    layers_.push_back(L);
  }
  nparams = off;
  // He-normal weights over the fan-in that carries signal; zero biases; the padded
  // input channels of layer 1 and the unused outputs of layer 6 start (and stay) at zero
  std::vector<float> h(nparams, 0.0f);
  std::mt19937_64 rng(seed);
  for (int li = 0; li < kLayers; ++li) {
    const Layer& L = layers_[(size_t)li];
    const int cin_used = li == 0 ? nin : L.cin;
    const int cout_used = li == kLayers - 1 ? nout : L.cout;
    std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (27.0 * cin_used)) * (li == kLayers - 1 ? 0.1 : 1.0));
    for (int co = 0; co < cout_used; ++co)
      for (int t = 0; t < 27; ++t)
        for (int ci = 0; ci < cin_used; ++ci) h[L.w_off + ((size_t)co * 27 + t) * L.cin + ci] = (float)nd(rng);
  }
  params.upload(h, s);
  grad.alloc(nparams);
  adam_m.alloc(nparams);
  adam_v.alloc(nparams);
  reset_optimizer();
  w16.alloc(nparams);
  wt16.alloc(nparams);
  const size_t v1 = (size_t)lvl_V[0], v2 = (size_t)lvl_V[1], v3 = (size_t)lvl_V[2];
  x16.alloc(v1 * 16);
  e1.alloc(v1 * 16);
  p1.alloc(v2 * 16);
  e2.alloc(v2 * 32);
  p2.alloc(v3 * 32);
  e3.alloc(v3 * 64);
  u3.alloc(v2 * 64);
  d2.alloc(v2 * 32);
  u2.alloc(v1 * 32);
  d1.alloc(v1 * 16);
  y.alloc(v1 * 16);
  gy.alloc(v1 * 16);
  g_d1.alloc(v1 * 16);
  g_z5.alloc(v1 * 16);
  g_u2.alloc(v1 * 32);
  g_d2.alloc(v2 * 32);
  g_z4.alloc(v2 * 32);
  g_u3.alloc(v2 * 64);
  g_e3.alloc(v3 * 64);
  g_z3.alloc(v3 * 64);
  g_p2.alloc(v3 * 32);
  g_e2.alloc(v2 * 32);
  g_z2.alloc(v2 * 32);
  g_p1.alloc(v2 * 16);
  g_e1.alloc(v1 * 16);
  g_z1.alloc(v1 * 16);
  gy16.alloc(v1 * 16);
  gz5_16.alloc(v1 * 16);
  gz4_16.alloc(v2 * 32);
  gz3_16.alloc(v3 * 64);
  gz2_16.alloc(v2 * 32);
  gz1_16.alloc(v1 * 16);
  size_t wp = 0;
  for (const Layer& L : layers_)
    wp = std::max(wp, (size_t)knet::wgrad_chunks(lvl_V[L.level], kChunk) * L.cout * 27 * L.cin);
  wpart.alloc(wp);
  dpart.alloc((size_t)std::max(knet::mse_blocks(V1, 16), knet::colsum_blocks(V1) * 64) + 8);  // >= V1 / 256 too
  loss.alloc(1);
  wsum.alloc(1);
  SLX_HIP(hipMemsetAsync(x16.p, 0, x16.n * sizeof(uint16_t), s));
  prep(s);
  SLX_HIP(hipStreamSynchronize(s));
}

void WarmNet::reset_optimizer() {
  step_ = 0;
  if (adam_m.p) SLX_HIP(hipMemset(adam_m.p, 0, nparams * sizeof(float)));
  if (adam_v.p) SLX_HIP(hipMemset(adam_v.p, 0, nparams * sizeof(float)));
}

void WarmNet::prep(hipStream_t s) {
  for (const Layer& L : layers_)
    knet::prep_weights(L.cout, L.cin, params.p + L.w_off, w16.p + L.w_off, wt16.p + L.w_off, s);
}

void WarmNet::set_parameters(const float* host, hipStream_t s) {
  if (!ready()) fail(SLX_CONFIG_ERROR, "warm-start network: not configured");
  params.upload(host, nparams, s);
  prep(s);
  SLX_HIP(hipStreamSynchronize(s));
}

void WarmNet::get_parameters(float* host, hipStream_t s) const {
  if (!ready()) fail(SLX_CONFIG_ERROR, "warm-start network: not configured");
  params.download(host, nparams, s);
  SLX_HIP(hipStreamSynchronize(s));
}

void WarmNet::forward(hipStream_t s) {
  const Layer* L = layers_.data();
  auto conv = [&](int li, const uint16_t* in, const uint16_t* add16, int relu, float* o32, uint16_t* o16) {
    const Layer& l = L[li];
    knet::ConvArgs a{lvl_X[l.level], lvl_Y[l.level], lvl_Z[l.level], lvl_V[l.level], l.cin, l.cout,
                     in, w16.p + l.w_off, params.p + l.b_off, nullptr, add16, relu, o32, o16};
    knet::conv(a, s);
  };
  conv(0, x16.p, nullptr, 1, nullptr, e1.p);
  knet::pool(lvl_X[0], lvl_Y[0], lvl_Z[0], 16, e1.p, p1.p, s);
  conv(1, p1.p, nullptr, 1, nullptr, e2.p);
  knet::pool(lvl_X[1], lvl_Y[1], lvl_Z[1], 32, e2.p, p2.p, s);
  conv(2, p2.p, nullptr, 1, nullptr, e3.p);
  knet::upsample(lvl_X[1], lvl_Y[1], lvl_Z[1], 64, e3.p, u3.p, s);
  conv(3, u3.p, e2.p, 1, nullptr, d2.p);
  knet::upsample(lvl_X[0], lvl_Y[0], lvl_Z[0], 32, d2.p, u2.p, s);
  conv(4, u2.p, e1.p, 1, nullptr, d1.p);
  conv(5, d1.p, nullptr, 0, y.p, nullptr);
}

void WarmNet::backward(const float* t, const float* w, hipStream_t s) {
  const Layer* L = layers_.data();
  // adjoint convolution: g_in = conv with flipped, transposed weights (+ add)
  auto conv_adj = [&](int li, const uint16_t* gz16, const float* add, float* out) {
    const Layer& l = L[li];
    knet::ConvArgs a{lvl_X[l.level], lvl_Y[l.level], lvl_Z[l.level], lvl_V[l.level], l.cout, l.cin,
                     gz16, wt16.p + l.w_off, nullptr, add, nullptr, 0, out, nullptr};
    knet::conv(a, s);
  };
  auto wgrad = [&](int li, const uint16_t* gz16, const float* gz, const uint16_t* in) {
    const Layer& l = L[li];
    const int V = lvl_V[l.level];
    knet::WgradArgs a{lvl_X[l.level], lvl_Y[l.level], lvl_Z[l.level], V, kChunk, gz16, in, wpart.p};
    knet::conv_wgrad(a, l.cout, l.cin, s);
    knet::sum_chunks(knet::wgrad_chunks(V, kChunk), (long long)l.cout * 27 * l.cin, wpart.p, grad.p + l.w_off, s);
    knet::colsum(V, l.cout, gz, dpart.p, grad.p + l.b_off, s);
  };
  const int X1 = lvl_X[0], Y1 = lvl_Y[0], Z1 = lvl_Z[0], X2 = lvl_X[1], Y2 = lvl_Y[1], Z2 = lvl_Z[1];
  const long long v1 = lvl_V[0], v2 = lvl_V[1], v3 = lvl_V[2];
  if (w) knet::sum_float(V1, w, dpart.p, wsum.p, s);
  knet::mse(V1, 16, nout_, y.p, t, w, wsum.p, gy.p, gy16.p, dpart.p, loss.p, s);
  // layer 6
  wgrad(5, gy16.p, gy.p, d1.p);
  conv_adj(5, gy16.p, nullptr, g_d1.p);
  // layer 5: d1 = relu(z5 + e1)
  knet::relu_adjoint(v1 * 16, g_d1.p, d1.p, g_z5.p, gz5_16.p, s);
  wgrad(4, gz5_16.p, g_z5.p, u2.p);
  conv_adj(4, gz5_16.p, nullptr, g_u2.p);
  knet::upsample_adjoint(X1, Y1, Z1, 32, g_u2.p, g_d2.p, s);
  // layer 4: d2 = relu(z4 + e2)
  knet::relu_adjoint(v2 * 32, g_d2.p, d2.p, g_z4.p, gz4_16.p, s);
  wgrad(3, gz4_16.p, g_z4.p, u3.p);
  conv_adj(3, gz4_16.p, nullptr, g_u3.p);
  knet::upsample_adjoint(X2, Y2, Z2, 64, g_u3.p, g_e3.p, s);
  // layer 3: e3 = relu(z3)
  knet::relu_adjoint(v3 * 64, g_e3.p, e3.p, g_z3.p, gz3_16.p, s);
  wgrad(2, gz3_16.p, g_z3.p, p2.p);
  conv_adj(2, gz3_16.p, nullptr, g_p2.p);
  knet::pool_adjoint(X2, Y2, Z2, 32, g_p2.p, g_z4.p, g_e2.p, s);  // + skip into d2
  // layer 2: e2 = relu(z2)
  knet::relu_adjoint(v2 * 32, g_e2.p, e2.p, g_z2.p, gz2_16.p, s);
  wgrad(1, gz2_16.p, g_z2.p, p1.p);
  conv_adj(1, gz2_16.p, nullptr, g_p1.p);
  knet::pool_adjoint(X1, Y1, Z1, 16, g_p1.p, g_z5.p, g_e1.p, s);  // + skip into d1
  // layer 1: e1 = relu(z1)
  knet::relu_adjoint(v1 * 16, g_e1.p, e1.p, g_z1.p, gz1_16.p, s);
  wgrad(0, gz1_16.p, g_z1.p, x16.p);
}

void WarmNet::adam_step(float lr, hipStream_t s) {
  ++step_;
  knet::adam((long long)nparams, params.p, grad.p, adam_m.p, adam_v.p, lr, 0.9f, 0.999f, 1.0e-8f, step_, s);
  prep(s);
}

}  // namespace slx
