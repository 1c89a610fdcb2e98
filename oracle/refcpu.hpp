// refcpu: CPU restatement of SparLab's SIMP linear-solve path.
//
// TEST INFRASTRUCTURE ONLY. This library is the oracle the HIP product is
// checked against, and the CPU baseline timed beside it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product library under sparlab_amd/ never links or calls it.
//
// Every function restates one piece of the reference algorithm and cites the
// reference file:line it follows (paths relative to the SparLab source tree).
// The reference itself cannot be compiled here (it needs Eigen 3, which is not
// installed), so the parts Eigen computes are restated from the published
// algorithm Eigen implements:
//   * column-pivoted Householder QR (Businger-Golub, LAPACK xGEQP3 norm
//     downdating, LAWN 176), used by the tentative prolongator;
//   * LDL^T with diagonal pivoting, used for the coarsest level;
//   * the largest eigenvalue of a symmetric tridiagonal matrix (Sturm-sequence
//     bisection instead of Eigen's implicit QR; same value to rounding).
// Those parts are pinned by invariants, not bits ("parity unpinned" at the
// Eigen boundary for the Q/R signs, the Lanczos eigenvalue and dot orders).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace refcpu {

using Index = int;  // Types.hpp:24-44: Index = StorageIndex = int

struct Error : std::exception {
  int code;  // 1 config, 2 solver, 3 convergence, 4 model
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

struct Csr {
  Index rows = 0, cols = 0;
  std::vector<Index> ptr, col;
  std::vector<double> val;
  Index nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

// ---------------------------------------------------------------------------
// Mesh (StructuredMesh.cpp:157-247; numbering StructuredMesh.hpp:6-27)
// ---------------------------------------------------------------------------
struct Mesh {
  int dim = 0;
  int npe = 0;
  Index nn = 0, ne = 0;
  Index nx = 0, ny = 0, nz = 0;
  double lx = 0, ly = 0, lz = 0;
  std::vector<double> coords;  // nn x dim, node-major
  std::vector<Index> conn;     // ne x npe, VTK corner order
  const Index* element_nodes(Index e) const { return conn.data() + (size_t)e * npe; }
};
Mesh structured_quad(Index nx, Index ny, double lx, double ly);
Mesh structured_hex(Index nx, Index ny, Index nz, double lx, double ly, double lz);
void element_centroid(const Mesh& m, Index e, double* c);   // Mesh.cpp:370-375
double element_measure(const Mesh& m, Index e);              // Mesh.cpp:377-396 (Q4 shoelace, Hex8 2x2x2 det)
double mean_element_size(const Mesh& m);                     // Mesh.cpp:398-406

// ---------------------------------------------------------------------------
// Material and element stiffness
// ---------------------------------------------------------------------------
void d_plane_stress(double E, double nu, double d[9]);       // IsotropicMaterial.cpp:34-43
void d_three_dimensional(double E, double nu, double d[36]); // IsotropicMaterial.cpp:56-66
// Hex8 stiffness (Hex8.cpp:46-101,172-192), 2x2x2 Gauss (Quadrature.cpp:74-95).
// xe: 3 x 8 corner coordinates (xe[i*8+a]); ke: 24 x 24 row-major.
void hex8_stiffness(const double* xe, const double* d, double* ke);
// Q4 stiffness (Quad4.cpp:44-81,117-137), 2x2 Gauss (Quadrature.cpp:54-72).
void quad4_stiffness(const double* xe, const double* d, double t, double* ke);

// ---------------------------------------------------------------------------
// DOFs, pattern, assembly (DofManager.cpp:58-71, Assembler.cpp:85-225,410-438)
// ---------------------------------------------------------------------------
struct Dofs {
  int ndpn = 0;
  Index ndofs = 0;
  std::vector<char> fixed;
  std::vector<Index> free, reduced;
  void rebuild();
};
struct Pattern {
  Index n = 0;
  std::vector<Index> outer, inner, block_offset;
};
Pattern build_pattern(const Mesh& m, int ndpn);
// values = sum_e s_e Ke0 scattered in ascending element order (Assembler.cpp:182-214)
void scatter(const Pattern& p, const Mesh& m, int ndpn, const double* ke, const double* scale,
             std::vector<double>& values);
struct Reduction {
  Index rows = 0;
  std::vector<Index> ptr, col, src;
};
Reduction free_free_map(const Pattern& p, const Dofs& dofs);  // Assembler.cpp:410-438
Csr reduce(const Reduction& r, const Pattern& p, const std::vector<double>& full);  // :216-225

// ---------------------------------------------------------------------------
// SA-AMG and PCG (Multigrid.cpp)
// ---------------------------------------------------------------------------
struct AmgOptions {  // Multigrid.hpp:75-103
  double strength_threshold = 0.02;
  int max_levels = 12;
  Index coarse_size = 1500;
  int smoother = 0;  // 0 Chebyshev, 1 symmetric Gauss-Seidel
  int smoother_degree = 3;
  double chebyshev_ratio = 30.0;
  double prolongator_damping = 4.0 / 3.0;
  int lanczos_steps = 12;
  double coarse_pivot_tolerance = 1.0e-13;
  bool reuse_aggregates = true;
};
struct Layout {  // Multigrid.hpp:105-117
  int dim = 0;
  int ndpn = 0;
  const double* coords = nullptr;  // nn x dim node-major
  Index nn = 0;
  const Index* unknowns = nullptr;  // ascending global DOFs
  Index n = 0;
};
struct LevelStats {
  Index unknowns = 0, nonzeros = 0, aggregates = 0;
  double lambda_max = 0;
};
struct AmgStats {  // Multigrid.hpp:126-150
  std::vector<LevelStats> levels;
  double operator_complexity = 0, grid_complexity = 0, coarse_pivot_ratio = 0;
  bool reused_aggregates = false;
  int near_null_space_dimension = 0;
  long long storage_nonzeros = 0;  // Multigrid.cpp:1089-1100
};

double chunked_dot(const double* a, const double* b, Index n);  // Multigrid.cpp:74-95
void spmv(const Csr& a, const double* x, double* y);            // :98-107
void residual(const Csr& a, const double* x, const double* b, double* r);  // :110-119
Csr multiply(const Csr& a, const Csr& b);                        // :122-194
void multiply_numeric(const Csr& a, const Csr& b, Csr& c);       // :199-225
Csr transpose(const Csr& a);                                     // :227-247
Csr symmetrize(const Csr& a);                                    // :250-284
double lanczos_lambda_max(const Csr& a, const std::vector<double>& inv_diag, int steps);  // :288-327
double tridiagonal_max_eigenvalue(const std::vector<double>& diag, const std::vector<double>& sub);

// Column-pivoted Householder QR of a rows x m block (row-major). Returns rank,
// q (rows x rank, row-major) and r (rank x m, row-major, columns un-pivoted).
Index cpqr_block(const std::vector<double>& block, Index rows, Index m, std::vector<double>& q,
                 std::vector<double>& r);

struct Level {
  Csr a;  // level operator (level 0: a copy of the caller's matrix)
  std::vector<double> inv_diag;
  double lambda_max = 0;
  Index num_nodes = 0;
  std::vector<Index> node_ptr;
  std::vector<double> nullspace;  // n x m row-major
  std::vector<Index> aggregate;
  Csr p_tent, p, r, ap, coarse_raw;
  mutable std::vector<double> x, b, res, d;
};

class Amg {
 public:
  explicit Amg(AmgOptions o = AmgOptions());
  void setup(const Csr& a, const Layout& layout);  // Multigrid.cpp:929-1111
  void apply(const double* r, double* z) const;     // :1113-1121
  const AmgStats& stats() const { return stats_; }
  const std::vector<Level>& levels() const { return levels_; }
  const AmgOptions& options() const { return opt_; }
  int nullspace_dim() const { return m_; }

 private:
  void build_finest_nodes(const Layout& layout, Index n);  // :383-455
  void aggregate_level(Level& lvl);                        // :457-565
  void tentative_prolongator(Level& lvl, Level& next);     // :567-669
  void numeric_level(size_t l);                            // :718-747
  void smooth_prolongator(Level& lvl, bool numeric_only);  // :671-716
  void coarse_factorize(Level& lvl);                       // :749-781
  void smooth(const Level& lvl, bool pre, bool zero_guess) const;  // :783-855
  void cycle(size_t l) const;                                       // :857-885

  AmgOptions opt_;
  AmgStats stats_;
  std::vector<Level> levels_;
  int m_ = 0;
  bool ready_ = false;
  Index fp_rows_ = -1, fp_nnz_ = -1;
  std::vector<Index> fp_unknowns_;
  // coarsest dense LDL^T with diagonal pivoting
  std::vector<double> ldl_;  // n x n, unit lower L below the diagonal, D on it
  std::vector<Index> perm_;
};

struct PcgResult {
  int iterations = 0;
  double relative_residual = 0;
  bool converged = false;
};
PcgResult pcg_solve(const Csr& a, const double* b, double* x, const Amg& m, double tol,
                    int max_iterations);  // Multigrid.cpp:1152-1237

// ---------------------------------------------------------------------------
// Topology optimisation
// ---------------------------------------------------------------------------
struct Filter {  // DensityFilter.cpp:19-195
  Csr h;         // row-normalised weights, rows = elements, columns ascending
  std::vector<double> row_sums;
  void build(const Mesh& m, double radius);
  void to_physical(const double* x, double* y) const;  // y = H x
  void pull_back(const double* v, double* y) const;    // y = H^T v
};

struct Simp {  // SimpInterpolation.hpp:71-80
  double penalty = 3.0;
  double emin_ratio = 1.0e-9;
};
double simp_factor(double rho, const Simp& s);      // SimpInterpolation.cpp:59-62
double simp_derivative(double rho, const Simp& s);  // :64-70

struct OcOptions {  // OptimalityCriteria.hpp:35-43
  double move_limit = 0.2;
  double damping = 0.5;
  double lambda_lower = 1.0e-12;
  double lambda_upper = 1.0e12;
  int max_bisections = 200;
  double volume_tolerance = 1.0e-10;
};
struct OcStep {
  std::vector<double> x;
  double lambda = 0, achieved_volume = 0, max_change = 0;
  int bisections = 0;
  bool volume_converged = false;
};

// A structured SIMP problem: mesh, material, supports, load cases, design
// domain and the solver state that persists over a run (Sensitivity.cpp:110-237).
// Heaviside projection (src/topopt/Projection.cpp:58-94, include/sparlab/topopt/Projection.hpp:68-97)
double heaviside_project(double r, double beta, double eta);
double heaviside_derivative(double r, double beta, double eta);
struct Projection {  // ProjectionOptions without the robust formulation
  bool enabled = false;
  double eta = 0.5, beta_start = 1.0, beta_max = 32.0, beta_factor = 2.0;
  int beta_interval = 50;
  bool advance_on_convergence = true;
  bool robust = false;             // objective on eta + delta, volume on eta - delta
  double robust_delta = 0.1;
  int robust_volume_interval = 1;  // iterations between dilated-target updates
  void validate() const;  // Projection.cpp:10-56
  int num_stages() const;
  double beta_of_stage(int k) const;
};

struct Problem {
  Mesh mesh;
  int ndpn = 0;
  double E = 1, nu = 0.3, thickness = 1;
  Dofs dofs;
  std::vector<double> ke0;  // edofs x edofs row-major
  Pattern pattern;
  Reduction ff;
  std::vector<std::vector<double>> loads;  // full-length load vectors
  std::vector<double> weights;             // normalised load-case weights
  std::vector<double> volumes;             // element volumes
  std::vector<double> lower, upper;        // design bounds (passive elements pinned)
  std::vector<char> passive;
  Filter filter;
  bool filter_on = false;
  double volume_fraction = 0.5;
  Simp simp;
  AmgOptions amg_options;
  double tolerance = 1.0e-8;
  int max_iterations = 1000;
  bool warm_start = true;
  double beta = 0.0, eta = 0.5;  // projection of the physical density (beta 0: off)
  Projection projection;         // run_oc's beta continuation
  // state
  Amg amg;
  Csr kff;
  std::vector<std::vector<double>> previous;
  long total_iterations = 0;
  // wall seconds of the last evaluate(): K_ff assembly, AMG setup, PCG, and the whole call
  double seconds_assemble = 0, seconds_setup = 0, seconds_pcg = 0, seconds_total = 0;

  void finalize();  // DOFs, Ke0, pattern, reduction, volumes
  Csr assemble_kff(const double* scale) const;
  double domain_volume() const;
};

struct Evaluation {
  double compliance = 0, volume = 0, volume_fraction = 0;
  int linear_iterations = 0;
  std::vector<double> physical, factors, dc_dx, dv_dx, energy;
  std::vector<double> load_case_compliance;
  std::vector<std::vector<double>> u;  // full displacement vectors
};
Evaluation evaluate(Problem& p, const std::vector<double>& x, bool need_gradients);
// volume_target > 0 replaces the domain's target; vol_eta >= 0 measures the volume
// projected at that threshold; step_toward: step to the nearest reachable volume
// when the move-limited box cannot reach the target (OptimalityCriteria.cpp:40-110)
OcStep oc_update(const Problem& p, const std::vector<double>& x, const std::vector<double>& dc,
                 const std::vector<double>& dv, const OcOptions& o, double volume_target = 0.0,
                 double vol_eta = -1.0, bool step_toward = false);
struct RunOptions {
  int max_iterations = 200;
  double change_tolerance = 1.0e-3;
  double objective_tolerance = 0.0;
  int objective_window = 5;
  int continuation_steps = 1;
  double penalty_start = 1.0;
  int continuation_iterations = 25;
  OcOptions oc;
};
struct IterationRecord {  // one history.csv row (ResultWriter.cpp:1428-1470)
  int iteration;
  double penalty, compliance, volume, volume_fraction, max_change, lambda, gray;
  int bisections;
  bool volume_converged;
  int linear_iterations;
  double beta;
  double eroded_vf = 0, dilated_vf = 0, dilated_target_fraction = 0;  // robust runs
};
struct RunResult {
  std::vector<IterationRecord> history;
  std::vector<double> x;
  double compliance = 0, volume = 0, volume_fraction = 0, gray = 0;
  std::vector<double> load_case_compliance;
  int iterations = 0;
  bool converged = false;
  int stop_reason = 0;  // 0 iteration_cap, 1 design_change, 2 objective_stall
  double final_beta = 0.0;
  // robust runs (RobustRecord): the blueprint (eta) is the result above
  double compliance_eroded = 0, compliance_dilated = 0, vf_eroded = 0, vf_dilated = 0, dilated_target_fraction = 0;
};
RunResult run_oc(Problem& p, const RunOptions& o, const std::vector<double>& x0);  // TopologyOptimizer.cpp:556-724

}  // namespace refcpu
