// Solver-level functions shared by the library's translation units.
#pragma once

#include <functional>

#include "slx_internal.hpp"

namespace slx {

// SA-AMG setup of the context's current K_ff (multigrid hierarchy on the device).
// The coarsest level's dense inverse is queued on the device. defer_coarse:
// its pivot check is left to finish_coarse() (called by every V-cycle entry),
// so the setup returns without reading the pivots back
void amg_setup(Context& c, bool defer_coarse = false);
void finish_coarse(Context& c);
// level 0's diagonal and spectral bound from the current element factors, the
// coarse levels kept (matrix-free Hex8; the coarse_reuse option)
void amg_refresh_level0(Context& c);
// average milliseconds of one Chebyshev step at a level above the coarsest,
// over reps back-to-back launches between two events (measurement only)
double time_smoother(Context& c, int level, int reps);
// The matrix-free operator's tile shape for the V-cycle's precision, once per
// structure: level 0's Chebyshev step is timed in each shape (kmf::tile_shape)
// and the fastest kept. A node's sums run in one order whatever the tile, so the
// choice changes time only. SPARLAB_MF_TILE=t fixes shape t; SPARLAB_NO_AUTOTUNE=1
// keeps shape 0.
void tune_tile(Context& c);
// FP32 copies of the hierarchy for the mixed-precision V-cycle.
void prepare_mixed(Context& c);
// z = M^-1 r: one V-cycle from a zero guess (device pointers, length n of level 0).
void vcycle(Context& c, const double* r, double* z);
// Preconditioned CG on the device (src/fem/Multigrid.cpp:1152-1237); b, x device pointers.
slx_pcg_result pcg(Context& c, const double* b, double* x, double tol, int max_iterations);
// The same for nv load cases at once (case-major b and x, nv x n): one operator
// product and one V-cycle pass per iteration for every case; each case's
// result equals pcg() on that case alone (FP64 or FP32 V-cycle).
void pcg_block(Context& c, int nv, const double* b, double* x, double tol, int max_iterations,
               std::vector<slx_pcg_result>& out);

// Independent PCG solves on the diagonal blocks [seg_ptr[s], seg_ptr[s+1]) of the
// factorized matrix, all segments per kernel launch; b, x device pointers.
void seg_pcg(Context& c, const std::vector<int>& seg_ptr, const double* b, double* x, double tol, int max_iterations,
             std::vector<slx_pcg_result>& results);

// K_ff products on the device: matrix-free for structured Hex8 models
// (after slx_assemble), the assembled CSR otherwise
void op_apply(Context& c, const double* x, double* y);
void op_residual(Context& c, const double* x, const double* b, double* r);
int op_dot_blocks(const Context& c);
void op_apply_dot(Context& c, const double* x, double* y, double* partials);
// Lanczos step at level 0: w = sq (K (sq v)) - (*beta) vprev
// one Lanczos step's product w = s (K (s v)) - (*beta) vprev with the block
// partials of w.v; returns the number of partials
int op_lanczos(Context& c, const double* sq, const double* v, const double* vprev, const double* beta, double* w,
               double* partials);
// whether level-0 products run matrix-free (SPARLAB_NO_MATRIX_FREE=1 turns it off)
bool matrix_free_enabled();

// launches recorded into an executable graph when SPARLAB_GRAPHS=1 (default: eager launches)
bool graphs_enabled();
// several load cases of an evaluation through pcg_block (SPARLAB_NO_BLOCK_CASES=1: one by one)
bool block_cases_enabled();
hipGraphExec_t capture_graph(Context& c, const std::function<void()>& body);

// drop the captured CG graphs (after a setup or a reallocation)
void invalidate_graphs(Context& c);
// workspace sized for the current K_ff
void ensure_workspace(Context& c);
double device_dot(Context& c, long long n, const double* a, const double* b);
bool device_all_finite(Context& c, long long n, const double* v);

// RAII event pair around a kernel launch when timing is on
struct Timed {
  Context& c;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(Context& ctx, const char* n);
  ~Timed();
};

}  // namespace slx
