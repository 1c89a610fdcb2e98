// Learned warm start of the compliance solves and the slx_warmstart_* /
// slx_net_* entry points. The reference starts each load case's CG from the
// previous solution (src/topopt/Sensitivity.cpp:169-176); here a network
// trained on the solves of the same run proposes a displacement change d and CG
// starts from the point of u_prev + span(d, previous change) closest to the
// solution in the K-norm (never further than u_prev). CG then runs to the
// unchanged tolerance, so the result is judged by the same test.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>

#include "slx_kernels.hpp"
#include "slx_net.hpp"
#include "slx_solver.hpp"

namespace slx {
namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// the node grid of the structured model and the channels the features use
void ensure_model_net(Context& c, WarmStart& w) {
  if (w.model_grid) return;
  if (c.dim != 2 && c.dim != 3) fail(SLX_CONFIG_ERROR, "learned warm start: structured Hex8 or Q4 models only");
  const int Z = c.dim == 3 ? c.nz + 1 : 1;
  w.net.configure(c.nx + 1, c.ny + 1, Z, 2 + 3 * c.dim, c.dim, w.opt.seed, c.stream);
  const size_t V = (size_t)w.net.voxels();
  w.xs.clear();
  w.ts.clear();
  w.ws.clear();
  w.xs.resize((size_t)w.opt.buffer);
  w.ts.resize((size_t)w.opt.buffer);
  w.ws.resize((size_t)w.opt.buffer + 1);  // the last one receives the weights of a prediction-only solve
  for (int i = 0; i < w.opt.buffer; ++i) {
    w.xs[(size_t)i].alloc(V * WarmNet::kIn);
    w.ts[(size_t)i].alloc(V * (size_t)c.dim);
  }
  for (auto& b : w.ws) b.alloc(V);
  const size_t nf = c.free_dofs.size();
  w.scales.alloc(4);
  w.dots.alloc(16);
  w.pending_r0 = w.pending_energy = false;
  w.dprev.alloc(nf * (size_t)c.ncases);
  w.have_dprev.assign((size_t)c.ncases, 0);
  w.uold.alloc(nf);
  w.r.alloc(nf);
  w.dir.alloc(nf);
  w.kdir.alloc(nf);
  w.ksec.alloc(nf);
  w.x0.alloc(nf);
  w.phys_prev.alloc((size_t)c.nelem);
  w.model_grid = true;
}

double read_loss(Context& c, const WarmNet& n) {
  double* hb = c.host_buffer(16);
  SLX_HIP(hipMemcpyAsync(hb + 13, n.loss_device(), sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SLX_HIP(hipStreamSynchronize(c.stream));
  return hb[13];
}

void train(Context& c, WarmStart& w) {
  const auto t0 = Clock::now();
  const size_t V = (size_t)w.net.voxels();
  std::uniform_int_distribution<int> pick(0, w.count - 1);
  for (int k = 0; k < w.opt.steps_per_solve; ++k) {
    // the newest sample first, then random replay
    const int slot = k == 0 ? (w.next + w.opt.buffer - 1) % w.opt.buffer : pick(w.rng);
    SLX_HIP(hipMemcpyAsync(w.net.input(), w.xs[(size_t)slot].p, V * WarmNet::kIn * sizeof(uint16_t),
                           hipMemcpyDeviceToDevice, c.stream));
    w.net.forward(c.stream);
    w.net.backward(w.ts[(size_t)slot].p, w.ws[(size_t)slot].p, c.stream);
    if (w.stats.train_steps == 0) w.stats.first_loss = read_loss(c, w.net);
    w.net.adam_step((float)w.opt.learning_rate, c.stream);
    ++w.stats.train_steps;
  }
  w.stats.last_loss = read_loss(c, w.net);
  w.stats.train_seconds += seconds_since(t0);
}

// the solve's dot products [0, count) on the host, in one read
const double* read_dots(Context& c, WarmStart& w, int count) {
  double* hb = c.host_buffer(16);
  SLX_HIP(hipMemcpyAsync(hb, w.dots.p, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SLX_HIP(hipStreamSynchronize(c.stream));
  return hb;
}

// folds the statistics of the previous solve (dots [8]-[10]) into stats
void fold_pending(WarmStart& w, const double* h) {
  if (w.pending_r0) {
    w.stats.residual_ratio_sum += w.pending_rr > 0.0 ? std::sqrt(h[8] / w.pending_rr) : 1.0;
    w.pending_r0 = false;
  }
  if (w.pending_energy) {
    const double e_plain = h[9], e_start = h[10];
    auto ratio = [&](double e) { return e_plain > 0.0 ? std::sqrt(std::max(e, 0.0) / e_plain) : 1.0; };
    w.stats.energy_ratio_sum += ratio(e_start);
    w.stats.last_energy_ratio = ratio(e_start);
    w.stats.net_ratio_sum += ratio(e_plain - w.pending_gain_net);
    w.stats.secant_ratio_sum += ratio(e_plain - w.pending_gain_sec);
    w.pending_energy = false;
  }
}

}  // namespace

void warm_before_solve(Context& c, int l) {
  WarmStart* wp = c.ws.get();
  if (!wp || !wp->opt.enable || !wp->have_prev || !c.have_u[(size_t)l]) return;
  WarmStart& w = *wp;
  ensure_model_net(c, w);
  const auto t0 = Clock::now();
  hipStream_t s = c.stream;
  const size_t nf = c.free_dofs.size();
  const long long n = (long long)nf;
  double* u = c.d_u.p + (size_t)l * nf;
  const double* b = c.d_loads_free.p + (size_t)l * nf;
  const bool training = w.solves < w.opt.train_solves;
  const bool predicting = w.count >= w.opt.min_samples;
  w.predicted = false;
  w.beta_sec = 0.0;
  double* D = w.dots.p;
  // the plain warm start's residual and the feature scales (rms of u_prev, r and
  // the previous change), formed on the device
  op_residual(c, u, b, w.r.p);
  const double* dprev = w.have_dprev[(size_t)l] ? w.dprev.p + (size_t)l * nf : nullptr;
  k::dot(n, u, u, c.partials.p, D + 0, s);
  k::dot(n, w.r.p, w.r.p, c.partials.p, D + 1, s);
  if (dprev)
    k::dot(n, dprev, dprev, c.partials.p, D + 2, s);
  else
    k::fill(1, 0.0, D + 2, s);
  knet::rms(3, D, (double)n, w.scales.p, s);  // the fourth (the target's scale) is set after the solve
  k::copy(n, u, w.uold.p, s);
  w.uold_valid = true;
  // the previous change s and its K-optimal step
  if (dprev) {
    op_apply(c, dprev, w.ksec.p);
    k::dot(n, dprev, w.r.p, c.partials.p, D + 3, s);
    k::dot(n, dprev, w.ksec.p, c.partials.p, D + 4, s);
  }
  const size_t V = (size_t)w.net.voxels();
  if (training || predicting) {
    const int slot = training ? w.next : w.opt.buffer;  // weights of a prediction-only solve go to the spare
    knet::features(c.dim, c.nx, c.ny, c.nz, c.d_phys.p, w.phys_prev.p, c.d_fac.p, c.d_reduced.p, u, w.r.p, dprev,
                   w.scales.p, WarmNet::kIn, w.net.input(), w.ws[(size_t)slot].p, s);
  }
  if (training) {
    // keep the input; the target follows from this solve
    w.pending = w.next;
    SLX_HIP(hipMemcpyAsync(w.xs[(size_t)w.pending].p, w.net.input(), V * WarmNet::kIn * sizeof(uint16_t),
                           hipMemcpyDeviceToDevice, s));
  }
  if (predicting) {
    w.net.forward(s);
    knet::prediction(c.dim, (long long)V, WarmNet::kIn, c.d_reduced.p, w.net.output(), w.dir.p, s);
    op_apply(c, w.dir.p, w.kdir.p);
    k::dot(n, w.dir.p, w.r.p, c.partials.p, D + 5, s);
    k::dot(n, w.dir.p, w.kdir.p, c.partials.p, D + 6, s);
    if (dprev) k::dot(n, w.dir.p, w.ksec.p, c.partials.p, D + 7, s);
  }
  // one read: this solve's dots and the previous solve's statistics
  double h[11];
  {
    const double* hb = read_dots(c, w, 11);
    std::copy(hb, hb + 11, h);
  }
  fold_pending(w, h);
  const double rr = h[1];
  const double sr = dprev ? h[3] : 0.0, sks = dprev ? h[4] : 0.0;
  const bool sec_ok = dprev && sks > 0.0 && std::isfinite(sks) && std::isfinite(sr);
  if (sec_ok) w.beta_sec = sr / sks;
  // x0 = u_prev + a d + b s, (a, b) from the Galerkin condition on span(d, s):
  // the point of that plane closest to the solution in the K-norm
  const double dr = predicting ? h[5] : 0.0, dkd = predicting ? h[6] : 0.0;
  const double dks = predicting && sec_ok ? h[7] : 0.0;
  const bool net_ok = predicting && dkd > 0.0 && std::isfinite(dkd) && std::isfinite(dr);
  double a = 0.0, bs = 0.0;
  const double det = dkd * sks - dks * dks;
  if (net_ok && sec_ok && det > 1.0e-12 * dkd * sks) {
    a = (dr * sks - sr * dks) / det;
    bs = (sr * dkd - dr * dks) / det;
  } else if (net_ok) {
    a = dr / dkd;
  } else if (sec_ok) {
    bs = sr / sks;
  }
  if (net_ok || sec_ok) {
    if (a != 0.0) {
      k::axpy(n, a, w.dir.p, u, s);
      k::axpy(n, -a, w.kdir.p, w.r.p, s);
    }
    if (bs != 0.0) {
      k::axpy(n, bs, dprev, u, s);
      k::axpy(n, -bs, w.ksec.p, w.r.p, s);
    }
    k::copy(n, u, w.x0.p, s);
    // the start's residual norm, read with the next prediction
    k::dot(n, w.r.p, w.r.p, c.partials.p, D + 8, s);
    w.pending_r0 = true;
    w.pending_rr = rr;
    ++w.stats.predictions;
    w.stats.step_sum += a;
    // energy gains of each direction alone: E(u_prev + t v) = E(u_prev) - (v.r)^2 / v.K v
    w.gain_net = net_ok ? dr * dr / dkd : 0.0;
    w.gain_sec = sec_ok ? sr * sr / sks : 0.0;
    w.predicted = true;
  } else if (predicting) {
    ++w.stats.rejected;
  }
  w.stats.predict_seconds += seconds_since(t0);
}

void warm_after_solve(Context& c, int l) {
  WarmStart* wp = c.ws.get();
  if (!wp || !wp->opt.enable) return;
  WarmStart& w = *wp;
  hipStream_t s = c.stream;
  const size_t nf = c.free_dofs.size();
  const long long n = (long long)nf;
  const double* un = c.d_u.p + (size_t)l * nf;
  double* D = w.dots.p;
  if (w.predicted) {
    // energy-norm distance of both starting points from the converged solution,
    // read with the next prediction
    auto energy = [&](const double* from, double* out) {
      k::copy(n, un, w.dir.p, s);
      k::axpy(n, -1.0, from, w.dir.p, s);
      op_apply(c, w.dir.p, w.kdir.p);
      k::dot(n, w.dir.p, w.kdir.p, c.partials.p, out, s);
    };
    energy(w.uold.p, D + 9);
    energy(w.x0.p, D + 10);
    w.pending_energy = true;
    w.pending_gain_net = w.gain_net;
    w.pending_gain_sec = w.gain_sec;
    w.predicted = false;
  }
  if (!w.uold_valid) return;
  w.uold_valid = false;
  double* dl = w.dprev.p + (size_t)l * nf;
  // this solve's displacement change (w.dir) ...
  k::copy(n, un, w.dir.p, s);
  k::axpy(n, -1.0, w.uold.p, w.dir.p, s);
  if (w.pending >= 0) {
    // ... minus what the previous change explains is the target, scale-free (the
    // Galerkin step fixes the magnitude); its rms goes to scales[3] on the device
    k::copy(n, w.dir.p, w.kdir.p, s);
    if (w.beta_sec != 0.0 && w.have_dprev[(size_t)l]) k::axpy(n, -w.beta_sec, dl, w.kdir.p, s);
    k::dot(n, w.kdir.p, w.kdir.p, c.partials.p, D + 11, s);
    knet::rms(1, D + 11, (double)n, w.scales.p + 3, s);
    const size_t V = (size_t)w.net.voxels();
    knet::target(c.dim, (long long)V, c.d_reduced.p, w.kdir.p, w.scales.p + 3, w.ts[(size_t)w.pending].p, s);
  }
  // ... is the next solve's previous change
  k::copy(n, w.dir.p, dl, s);
  w.have_dprev[(size_t)l] = 1;
  if (w.pending < 0) return;
  w.next = (w.pending + 1) % w.opt.buffer;
  w.count = std::min(w.count + 1, w.opt.buffer);
  w.pending = -1;
  ++w.stats.samples;
  ++w.solves;
  train(c, w);
}

void warm_flush_stats(Context& c) {
  WarmStart* wp = c.ws.get();
  if (!wp || !(wp->pending_r0 || wp->pending_energy) || !wp->dots.p) return;
  double h[11];
  const double* hb = read_dots(c, *wp, 11);
  std::copy(hb, hb + 11, h);
  fold_pending(*wp, h);
}

void warm_after_evaluate(Context& c) {
  WarmStart* wp = c.ws.get();
  if (!wp || !wp->opt.enable) return;
  ensure_model_net(c, *wp);
  k::copy(c.nelem, c.d_phys.p, wp->phys_prev.p, c.stream);
  wp->have_prev = true;
}

}  // namespace slx

using namespace slx;

namespace {

template <class F>
slx_status net_guarded(slx_context* ctx, F&& f) {
  Context* c = reinterpret_cast<Context*>(ctx);
  if (!c) return SLX_CONFIG_ERROR;
  try {
    SLX_HIP(hipSetDevice(c->device));
    if (!c->ws) c->ws = std::make_shared<WarmStart>();
    f(*c, *c->ws);
    c->last_error.clear();
    return SLX_OK;
  } catch (const Error& e) {
    c->last_error = e.msg;
    return (slx_status)e.code;
  } catch (const std::exception& e) {
    c->last_error = e.what();
    return SLX_DEVICE_ERROR;
  }
}

void require_net(const WarmStart& w) {
  if (!w.net.ready()) fail(SLX_CONFIG_ERROR, "warm-start network: call slx_net_configure first");
}

void load_pair(Context& c, WarmStart& w, const float* x, const float* t, const float* wt, DBuf<float>& tx,
               DBuf<float>& tt, DBuf<float>& tw) {
  const size_t V = (size_t)w.net.voxels();
  tx.upload(x, V * (size_t)w.net.nin(), c.stream);
  knet::pack_input((int)V, w.net.nin(), WarmNet::kIn, tx.p, w.net.input(), c.stream);
  if (t) tt.upload(t, V * (size_t)w.net.nout(), c.stream);
  if (wt) tw.upload(wt, V, c.stream);
}

}  // namespace

extern "C" {

void slx_warmstart_default_options(slx_warmstart_options* o) {
  if (!o) return;
  o->enable = 1;
  o->train_solves = 40;
  o->min_samples = 4;
  o->steps_per_solve = 16;
  o->buffer = 8;
  o->learning_rate = 1.0e-3;
  o->seed = 20261203ULL;
}

slx_status slx_warmstart_configure(slx_context* ctx, const slx_warmstart_options* opt) {
  return net_guarded(ctx, [&](Context& c, WarmStart& w) {
    if (!opt) fail(SLX_CONFIG_ERROR, "learned warm start: options are required");
    if (opt->enable) {
      if (!c.finalized) fail(SLX_MODEL_ERROR, "learned warm start: finalise the structured model first");
      if (opt->buffer < 1 || opt->min_samples < 1 || opt->steps_per_solve < 0 || opt->train_solves < 0 ||
          !(opt->learning_rate > 0.0))
        fail(SLX_CONFIG_ERROR, "learned warm start: buffer, min_samples >= 1, steps and solves >= 0, rate > 0");
    }
    c.ws = std::make_shared<WarmStart>();
    c.ws->opt = *opt;
    c.ws->rng.seed(opt->seed);
    if (opt->enable) ensure_model_net(c, *c.ws);
  });
}

slx_status slx_warmstart_stats_get(const slx_context* ctx, slx_warmstart_stats* out) {
  if (!ctx || !out) return SLX_CONFIG_ERROR;
  Context* c = reinterpret_cast<Context*>(const_cast<slx_context*>(ctx));
  if (!c->ws) {
    *out = slx_warmstart_stats{};
    return SLX_OK;
  }
  // the last solve's statistics may still be on the device
  return net_guarded(const_cast<slx_context*>(ctx), [&](Context& cc, WarmStart& w) {
    warm_flush_stats(cc);
    *out = w.stats;
  });
}

slx_status slx_net_configure(slx_context* ctx, int X, int Y, int Z, int nin, int nout, unsigned long long seed) {
  return net_guarded(ctx, [&](Context& c, WarmStart& w) {
    w.net.configure(X, Y, Z, nin, nout, seed, c.stream);
    w.model_grid = false;
  });
}

slx_status slx_net_parameter_count(const slx_context* ctx, long long* count) {
  const Context* c = reinterpret_cast<const Context*>(ctx);
  if (!c || !count) return SLX_CONFIG_ERROR;
  *count = c->ws ? (long long)c->ws->net.parameter_count() : 0;
  return SLX_OK;
}

slx_status slx_net_parameters_get(slx_context* ctx, float* params) {
  return net_guarded(ctx, [&](Context& c, WarmStart& w) {
    require_net(w);
    w.net.get_parameters(params, c.stream);
  });
}

slx_status slx_net_parameters_set(slx_context* ctx, const float* params) {
  return net_guarded(ctx, [&](Context& c, WarmStart& w) {
    require_net(w);
    w.net.set_parameters(params, c.stream);
  });
}

slx_status slx_net_forward(slx_context* ctx, const float* x, float* y) {
  return net_guarded(ctx, [&](Context& c, WarmStart& w) {
    require_net(w);
    DBuf<float> tx, tt, tw;
    load_pair(c, w, x, nullptr, nullptr, tx, tt, tw);
    w.net.forward(c.stream);
    const size_t V = (size_t)w.net.voxels();
    std::vector<float> full(V * 16);
    SLX_HIP(hipMemcpyAsync(full.data(), w.net.output(), full.size() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    SLX_HIP(hipStreamSynchronize(c.stream));
    const int no = w.net.nout();
    for (size_t v = 0; v < V; ++v)
      for (int k = 0; k < no; ++k) y[v * (size_t)no + (size_t)k] = full[v * 16 + (size_t)k];
  });
}

slx_status slx_net_gradients(slx_context* ctx, const float* x, const float* t, const float* wt, double* loss,
                             float* grad) {
  return net_guarded(ctx, [&](Context& c, WarmStart& w) {
    require_net(w);
    DBuf<float> tx, tt, tw;
    load_pair(c, w, x, t, wt, tx, tt, tw);
    w.net.forward(c.stream);
    w.net.backward(tt.p, wt ? tw.p : nullptr, c.stream);
    SLX_HIP(hipMemcpyAsync(grad, w.net.gradients(), w.net.parameter_count() * sizeof(float), hipMemcpyDeviceToHost,
                           c.stream));
    SLX_HIP(hipMemcpyAsync(loss, w.net.loss_device(), sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SLX_HIP(hipStreamSynchronize(c.stream));
  });
}

slx_status slx_net_train(slx_context* ctx, const float* x, const float* t, const float* wt, int steps,
                         double learning_rate, double* loss_first, double* loss_last) {
  return net_guarded(ctx, [&](Context& c, WarmStart& w) {
    require_net(w);
    if (steps < 1 || !(learning_rate > 0.0)) fail(SLX_CONFIG_ERROR, "warm-start network: steps >= 1, rate > 0");
    DBuf<float> tx, tt, tw;
    load_pair(c, w, x, t, wt, tx, tt, tw);
    const float* wd = wt ? tw.p : nullptr;
    for (int k = 0; k < steps; ++k) {
      w.net.forward(c.stream);
      w.net.backward(tt.p, wd, c.stream);
      if (k == 0 && loss_first) *loss_first = read_loss(c, w.net);
      w.net.adam_step((float)learning_rate, c.stream);
    }
    w.net.forward(c.stream);
    w.net.backward(tt.p, wd, c.stream);
    if (loss_last) *loss_last = read_loss(c, w.net);
  });
}

}  // extern "C"
