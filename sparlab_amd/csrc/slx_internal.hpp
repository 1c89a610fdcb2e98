// Internal types of the sparlab_hip library: host CSR matrices, device buffers,
// the multigrid level and the context. Host code and HIP kernels share them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "sparlab_hip.h"

namespace slx {

struct Error : std::exception {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};
[[noreturn]] void fail(int code, const std::string& msg);
// device address of the no-allocation scratch array on `device` (noalloc.hip),
// taken once per device with that device current
double* noalloc_scratch(int device);
void hip_check(hipError_t e, const char* what);
#define SLX_HIP(expr) ::slx::hip_check((expr), #expr)

// Host CSR matrix, rows sorted by column.
struct HCsr {
  int rows = 0, cols = 0;
  std::vector<int> ptr, col;
  std::vector<double> val;
  long long nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

// Large host <-> device copies through a pinned staging buffer (two halves, the
// host copy of one overlapping the DMA of the other); both complete before
// returning. Pageable copies of hundreds of MB run several times slower.
constexpr size_t kStageMinBytes = size_t(16) << 20;
void stage_copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s);
void stage_copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s);

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    n = count;
    if (count) SLX_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Host data may be pageable and short-lived: the copy completes before return.
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count * sizeof(T) >= kStageMinBytes) {
      stage_copy_h2d(p, h, count * sizeof(T), s);
    } else if (count) {
      SLX_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
      SLX_HIP(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  // small copies are queued (the caller synchronises); large ones complete here
  void download(T* h, size_t count, hipStream_t s) const {
    if (count * sizeof(T) >= kStageMinBytes) stage_copy_d2h(h, p, count * sizeof(T), s);
    else if (count) SLX_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// Largest coarsest level the device's dense coarse solver takes: its FP64 inverse
// (n^2 doubles, 2 GiB at this size) is formed by blocked Gauss-Jordan at every setup.
constexpr int kDenseCoarseMax = 16384;

// Device CSR matrix (int32 indices, FP64 values, optional FP32 and bfloat16 copies of the values).
struct DCsr {
  int rows = 0, cols = 0;
  long long nnz = 0;
  int max_row = 0;  // longest row
  int lanes = 0;    // lanes per row of the CSR kernels, set by the autotuner (0: by row length)
  DBuf<int> ptr, col;
  DBuf<double> val;
  DBuf<float> val32;
  DBuf<uint16_t> val16;  // bfloat16 copy (coarse operators of precision 2); FP32 products read it when present
  // shared column lists (rows of one node have one pattern): the V-cycle's
  // products read entry k of row i's column at ncol[cbase[i] + k]
  DBuf<int> cbase, ncol;
  void upload_pattern(const HCsr& h, hipStream_t s);
  // values keep 8 doubles of padding past the last row (the row products read
  // up to 8 values from the start of a run)
  void upload_values(const double* h, hipStream_t s) {
    if (val.n < (size_t)nnz + 8) val.alloc((size_t)nnz + 8);
    if (sizeof(double) * (size_t)nnz >= kStageMinBytes)
      stage_copy_h2d(val.p, h, sizeof(double) * (size_t)nnz, s);
    else if (nnz)
      SLX_HIP(hipMemcpyAsync(val.p, h, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice, s));
    SLX_HIP(hipStreamSynchronize(s));
  }
  void upload(const HCsr& h, hipStream_t s) {
    upload_pattern(h, s);
    upload_values(h.val.data(), s);
  }
};

// Contribution lists of one block product C = A B (kernels_amg.hip): one item
// per aggregate block of each C node-row, each with its (A run, B block) list.
struct ProductPlan {
  long long nitems = 0;
  int dmax = 0, wmax = 0;  // most rows per C node, widest block (lanes per item)
  DBuf<int> item_node, item_agg, item_pos;
  DBuf<long long> item_begin;
  DBuf<uint2> list;  // two words per entry (kernels_amg.hip)
};

// Row plan of C = A B (one wavefront per C node): per C node its runs of A's
// row (one per B node met), per run the B node's aggregate blocks with their
// positions in B's and C's rows (kernels_amg.hip k_row_product).
struct RowPlan {
  int nnodes = 0;
  int dmax = 0, jmax = 0, wmax = 0, lenc_max = 0, emax = 0, lena_max = 0, wpb = 0;
  bool pre = false, sa = false;
  long long runs = 0, entries = 0;
  DBuf<int> order, run_ptr;
  DBuf<uint4> run;
  DBuf<unsigned> entry;
};

// One level of the SA-AMG hierarchy (structure of src/fem/Multigrid.cpp:336-354).
// Structure (node blocks, aggregates, tentative prolongator, product patterns)
// is built on the host once per sparsity structure; every numeric array lives
// on the device. Level 0's operator is the context's K_ff.
struct Level {
  int n = 0;
  int num_nodes = 0;
  std::vector<int> node_ptr;
  std::vector<double> nullspace;  // n x m, row-major
  std::vector<int> aggregate;
  double lambda_max = 0.0;
  bool has_next = false;
  DCsr A;                  // operator (levels > 0)
  DCsr Ptent, P, R, AP, RAW;
  DCsr Rt;                 // P-hat^T: with the matrix-free operator, the level-0 transfers
                           // apply P = (I - omega D^-1 K) P-hat and R = P-hat^T (I - omega K D^-1)
  double omega = 0.0;      // prolongator damping / lambda_max
  DBuf<int> r_src;         // R.val[k] = P.val[r_src[k]]
  bool r_nodes = false;    // R's rows come in coarse-node groups reading consecutive P values (r_cnode)
  DBuf<int> r_cnode;       // the next level's node pointer, for that gather
  DBuf<int> sym_a, sym_b;  // next level's A from RAW (symmetrisation)
  // block-product plans (see kernels_amg.hip); node_ok when every operator of
  // this level has node-uniform rows in complete aggregate blocks
  bool node_ok = false;
  DBuf<int> f_node_ptr, c_node_ptr, row_agg;
  ProductPlan plan_p;          // P = (I - w D^-1 A) P-hat
  RowPlan plan_ap, plan_raw;   // A P, R (A P)
  DBuf<double> dinv, sq, lz_v, lz_vprev, lz_w, lz_start, lz_scal, lz_part;
  // P-hat's aggregate blocks for the level-0 transfers (amg.cpp build_tent_blocks)
  bool tent_ok = false;
  int tent_naggs = 0;
  DBuf<int> tent_c0len, tent_agg_ptr, tent_agg_rows, tent_agg_c0len;
  DBuf<double> tent_ell, tent_ella;  // by fine row; in the aggregates' member order
  DBuf<float> tent_ell32, tent_ella32;
  // symmetric Gauss-Seidel smoother: the rows grouped by dependency level for the
  // forward and the backward sweep (amg.cpp gs_schedule), once per structure
  DBuf<int> gs_frows, gs_brows;
  std::vector<int> gs_fptr, gs_bptr;
  // V-cycle vectors: x the pre-smoothing iterate, x2 the post-smoothed result
  DBuf<double> x, x2, b, res, d, d2;
  DBuf<float> dinv32, x32, x232, b32, res32, d32, d232;
  DBuf<double> vx, vx2, vb, vres, vd, vd2;  // a block of load cases side by side (case-major)
  DBuf<float> vx32, vx232, vb32, vres32, vd32, vd232;
};

struct TimerSlot {
  double total_ms = 0.0;
  long long launches = 0;
};

struct WarmStart;  // learned warm start (slx_net.hpp)
struct SegState;   // segmented CG of a block-diagonal batch (seg.cpp)

// Problem / solver context: one GPU stream, one structured SIMP problem.
struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;

  // ---- model ---------------------------------------------------------------
  int dim = 0, npe = 0, ndpn = 0;
  int nx = 0, ny = 0, nz = 0;
  double lx = 0, ly = 0, lz = 0, E = 1, nu = 0.3, thickness = 1;
  long long nnodes = 0, nelem = 0, ndofs = 0;
  std::vector<unsigned char> fixed;
  std::vector<int> free_dofs, reduced;  // free list (ascending), DOF -> reduced index or -1
  std::vector<double> ke0;              // edofs x edofs
  std::vector<double> coords;           // nnodes x dim
  std::vector<double> volumes;          // element volumes
  double total_volume = 0.0;            // their sum in element order (the domain volume)
  bool unit_bounds = true;              // every design bound lies in [0, 1]
  bool design_dirty = true;             // host bounds not yet on the device
  DBuf<double> d_htvol;                 // H^T vol = dV/dx, fixed by the filter and the volumes
  bool htvol_valid = false;
  bool finalized = false;
  int ncases = 0;
  std::vector<double> loads;    // ncases x ndofs
  std::vector<double> weights;  // normalised
  std::vector<double> case_compliance;  // f_l . u_l of the last evaluation
  // device model state
  DBuf<int> d_free, d_reduced;
  DBuf<int> d_nodemap;  // Hex8: packed per-node free map of the matrix-free kernels
  // per free DOF its grid node and component, packed for the assembly (x | y << 10 |
  // z << 20 | c << 30 in 3-D, x | y << 15 | c << 30 in 2-D); empty when the grid does not fit
  DBuf<unsigned> d_dofxyz;
  DBuf<double> d_ke0, d_scale, d_loads_free, d_weights;  // loads restricted to the free DOFs (ncases x nfree)
  DBuf<double> d_u;                           // last solutions (ncases x nfree)
  std::vector<char> have_u;
  DCsr K;  // K_ff on the device
  // what K holds: the structured pattern only (after slx_model_finalize), an
  // assembly of caller factors (slx_assemble), the last objective evaluation,
  // the modal analysis's stiffness, or a caller-assembled matrix
  // (slx_factorize_csr, which leaves the structured model unusable until it is
  // finalised again)
  enum class KSource { none, assembled, evaluation, modal, caller };
  KSource k_source = KSource::none;
  unsigned long long k_pattern_hash = 0;  // of a caller matrix's pattern (ptr, col)
  long long k_pattern_version = 0;        // changes whenever K's pattern is replaced
  HCsr k_host;                            // K's pattern (ptr, col) on the host
  long long k_host_version = -1;          // k_pattern_version it belongs to
  bool k_on_host = false;
  // K_ff's values lag the element factors: an evaluation that reused the coarse
  // operators refreshed only level 0 and skipped the assembly (coarse_reuse)
  bool k_stale = false;
  int coarse_lag = 0;      // evaluations since the last full numeric setup
  int lag_base_its = 0;    // PCG iterations of the evaluation right after it
  int lag_last_its = 0;    // and of the latest evaluation
  // matrix-free K_ff of a structured Hex8 model: the element factors of the last
  // assembly (FP64 and FP32) and Ke0 in FP32; used for level-0 products when mf
  bool mf = false;
  DBuf<double> mf_scale;
  DBuf<float> mf_scale32, mf_ke32;

  // ---- multigrid -------------------------------------------------------------
  slx_amg_options amg_opt{};
  std::vector<Level> levels;
  int nullspace_dim = 0;
  bool amg_ready = false;
  bool reused = false;
  long long fp_rows = -1, fp_nnz = -1;
  long long fp_layout_version = -1;  // layout_version at the last setup
  long long fp_pattern_version = -1;  // k_pattern_version at the last setup
  std::vector<double> layout_coords;  // coords of the DOF layout (nnodes x dim)
  int layout_dim = 0, layout_ndpn = 0, layout_nnodes = 0;
  std::vector<int> layout_unknowns;
  long long layout_version = 0;  // changes whenever layout_unknowns does
  DBuf<int> setup_flags;         // device error flags of a setup (amg.cpp)
  DBuf<double> coarse_inv;  // dense inverse of the coarsest operator
  DBuf<float> coarse_inv32;
  DBuf<double> coarse_piv, coarse_work;  // its pivots, the inversion's panels
  int coarse_n = 0;
  // the inverse is queued by a setup; finish_coarse() checks its pivots before
  // the first V-cycle
  bool coarse_pending = false;
  double coarse_pivot_ratio = 0.0;
  double setup_seconds = 0.0;
  int tuned_precision = -1;  // precision the CSR kernels were last autotuned for
  int mf_tile[2] = {0, 0};   // tile shapes of the FP64 and FP32 matrix-free kernels (kmf::tile_shape)
  double phase_seconds[7] = {0, 0, 0, 0, 0, 0, 0};  // check, spectral, aggregation, smoothing, Galerkin, coarse, tentative

  // ---- PCG workspace ------------------------------------------------------------
  DBuf<double> r, z, p, q, xw, bw, xpcg, zold;
  // the CG iteration as three captured graphs: q = K p (timed alone), the
  // alpha / x / r update, and the preconditioner with the direction update
  hipGraphExec_t g_update = nullptr, g_precond = nullptr;  // captured CG iteration parts
  bool capturing = false;
  DBuf<double> partials;  // block partial sums
  DBuf<double> eval_dots;  // an evaluation's f.u per load case and its volume, read together
  int oc_last_bisections = 0;  // bisection steps of the last OC update (the next one's first batch)
  DBuf<double> scal;      // device scalars
  std::vector<double> h_scal;
  // several load cases through one operator and hierarchy pass (pcg_block):
  // case-major vectors, 16 scalars and one active flag per case
  DBuf<double> vx, vr, vz, vp, vq, vpart, vscal;
  DBuf<int> vactive;
  int block_nv = 0;          // cases the block workspace and graphs were made for
  int block_precision = -1;  // and the V-cycle precision
  hipGraphExec_t g_vupdate = nullptr, g_vprecond = nullptr;
  // pinned host memory for small device -> host reads (scalars, per-segment sums)
  double* pinned = nullptr;
  size_t pinned_n = 0;
  double* host_buffer(size_t n);

  // ---- topology optimisation -------------------------------------------------
  double volume_fraction = 0.5;
  std::vector<double> lower, upper;
  std::vector<unsigned char> passive;
  bool filter_on = false;
  double filter_radius = 0.0;
  HCsr H, Ht;  // filter and its transpose
  DCsr dH, dHt;
  DBuf<double> d_x, d_phys, d_fac, d_dfac, d_dc, d_energy, d_vol, d_tmp, d_tmp2, d_lower, d_upper, d_dcx, d_dvx;
  DBuf<double> d_dproj;                   // d physical / d filtered under the projection
  std::vector<DBuf<double>> adjoint_slots;  // last adjoint solution per warm-start slot (free DOFs)
  long long adjoint_solves = 0, adjoint_iterations = 0;
  double proj_beta = 0.0, proj_eta = 0.5;  // projection of the last evaluation (beta 0: off)
  DBuf<double> oc_state;  // the device bisection state
  DBuf<unsigned char> d_passive;

  // ---- learned warm start -------------------------------------------------------
  std::shared_ptr<WarmStart> ws;
  // ---- batched (segmented) CG ------------------------------------------------------
  std::shared_ptr<SegState> seg;
  long long setup_id = 0;  // counts completed multigrid setups

  // ---- measurement ------------------------------------------------------------
  bool timing = false;
  std::map<std::string, TimerSlot> timers;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  void flush_timers();
};

}  // namespace slx
