/* sparlab_hip.h - C ABI of the MI355X (gfx950) SIMP linear-solve path.
 *
 * One opaque context owns one GPU stream and the device-resident state of one
 * structured SIMP problem: the cached K_ff pattern and values, the SA-AMG
 * hierarchy, the load cases and their last solutions, the density filter and
 * the design vectors. Every entry point takes plain pointers and sizes; host
 * arrays are caller-owned and read or written before the call returns. Each
 * call returns a status code; the message of the last failure on a context is
 * available from slx_last_error(). Status codes map one to one onto the
 * reference's exception classes (include/sparlab/core/Exceptions.hpp:15-64):
 * SLX_CONFIG_ERROR -> ConfigError, SLX_SOLVER_ERROR -> SolverError,
 * SLX_CONVERGENCE_ERROR -> ConvergenceError, SLX_MODEL_ERROR -> ModelError.
 *
 * The entry points replace these seams of the reference (paths relative to the
 * SparLab source tree):
 *   slx_model_*          FemModel + DofManager + Assembler construction
 *                        (src/fem/FemModel.cpp, src/fem/DofManager.cpp:58-71,
 *                         src/fem/Assembler.cpp:41-58,85-172)
 *   slx_assemble         Assembler::assemble_stiffness + reduce_free_free
 *                        (src/fem/Assembler.cpp:182-258,410-438)
 *   slx_amg_setup        LinearSolver::factorize of AmgCgSolver
 *                        (src/fem/LinearSolver.cpp:229-239 -> src/fem/Multigrid.cpp:929-1111)
 *   slx_factorize_csr    the same seam for a caller-assembled CSR matrix and DofLayout
 *                        (include/sparlab/fem/LinearSolver.hpp:71,87)
 *   slx_solve            LinearSolver::solve_from (src/fem/LinearSolver.cpp:246-281 ->
 *                        src/fem/Multigrid.cpp:1152-1237)
 *   slx_solve_case       StaticAnalysis::solve_load_vector with the warm start of
 *                        ComplianceObjective (src/fem/StaticAnalysis.cpp:114-143,
 *                        src/topopt/Sensitivity.cpp:169-176)
 *   slx_filter_*         DensityFilter build / to_physical / pull_back
 *                        (src/topopt/DensityFilter.cpp:115-195)
 *   slx_evaluate         ComplianceObjective::evaluate (src/topopt/Sensitivity.cpp:110-237)
 *   slx_oc_update        optimality_criteria_update (src/topopt/OptimalityCriteria.cpp:40-149)
 *   slx_run_oc           TopologyOptimizer::run_oc (src/topopt/TopologyOptimizer.cpp:556-724)
 *   slx_solve_adjoint    ComplianceObjective::solve_adjoint with warm slots
 *                        (src/topopt/Sensitivity.cpp:239-257)
 *   slx_modal_solve      solve_modal's shift-invert subspace iteration with warm-started
 *                        AMG-CG solves (src/fem/ModalAnalysis.cpp:120-418)
 *
 * Additions with no counterpart in the reference:
 *   slx_warmstart_*      a learned initial guess for the compliance solves, which
 *                        the reference warm-starts from the previous solution
 *                        (src/topopt/Sensitivity.cpp:169-176); CG still runs to
 *                        the same tolerance
 *   slx_net_*            direct access to that network (parameters, forward pass,
 *                        gradients, training) for tests against a CPU restatement
 */
#ifndef SPARLAB_HIP_H
#define SPARLAB_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slx_context slx_context;

typedef enum {
  SLX_OK = 0,
  SLX_CONFIG_ERROR = 1,
  SLX_SOLVER_ERROR = 2,
  SLX_CONVERGENCE_ERROR = 3,
  SLX_MODEL_ERROR = 4,
  SLX_DEVICE_ERROR = 5
} slx_status;

/* Multigrid options (include/sparlab/fem/Multigrid.hpp:75-103), plus the
 * precision of the hierarchy: 0 = FP64 everywhere (the drop-in path),
 * 1 = FP32 hierarchy and V-cycle under an FP64 outer (flexible) CG,
 * 2 = as 1 with the coarse-level operators stored in bfloat16 (FP64 row sums). */
typedef struct {
  double strength_threshold;     /* 0.02 */
  int max_levels;                /* 12 */
  int coarse_size;               /* 1500 */
  int smoother;                  /* 0 Chebyshev, 1 symmetric Gauss-Seidel (sweeps by dependency level) */
  int smoother_degree;           /* 3 */
  double chebyshev_ratio;        /* 30 */
  double prolongator_damping;    /* 4/3 */
  int lanczos_steps;             /* 12 */
  double coarse_pivot_tolerance; /* 1e-13 */
  int reuse_aggregates;          /* 1 */
  int precision;                 /* 0 */
  /* Reuse of the coarse operators across objective evaluations (no counterpart in
   * the reference, which recomputes every level per evaluation; default 0 = that
   * behaviour). k > 0, structured Hex8 with the matrix-free operator only: after a
   * full numeric setup, up to k evaluations refresh level 0's diagonal and spectral
   * bound from the new element factors and keep A_1 and below, skipping the K_ff
   * assembly and the Galerkin products; a full setup follows as soon as an
   * evaluation's PCG iterations exceed those after the last full setup by 20 %
   * (and at least 2). PCG still runs to the same tolerance. */
  int coarse_reuse;              /* 0 */
} slx_amg_options;

typedef struct {
  int iterations;
  double relative_residual; /* true ||b - A x|| / ||b|| at exit */
  int converged;
} slx_pcg_result;

typedef struct {
  int levels;
  int reused_aggregates;
  double operator_complexity;
  double grid_complexity;
  double coarse_pivot_ratio;
  int near_null_space_dimension;
  long long unknowns[16];
  long long nonzeros[16];
  long long aggregates[16];
  double lambda_max[16];
  double setup_seconds;
  /* setup phases (include/sparlab/fem/Multigrid.hpp:140-149) */
  double check_seconds;
  double spectral_seconds;
  double aggregation_seconds;
  double smoothing_seconds;
  double galerkin_seconds;
  double coarse_seconds;
  int node_product_levels; /* levels whose products ran on the node-row kernels */
  /* AmgStats::storage_nonzeros (src/fem/Multigrid.cpp:1089-1100): the coarse
   * operators plus P-hat, P, R, A P and R A P of every level, plus n_c^2 when the
   * coarsest level has at most 4000 unknowns; what LinearSolver::storage_nonzeros()
   * reports (src/fem/LinearSolver.cpp:258) */
  long long storage_nonzeros;
  double tentative_seconds; /* AmgStats::tentative_seconds: the tentative prolongators */
} slx_amg_stats;

/* Evaluation options: SIMP law (include/sparlab/topopt/SimpInterpolation.hpp:71-80)
 * and the linear-solver settings ComplianceObjective uses. */
typedef struct {
  double penalty;         /* 3 */
  double emin_ratio;      /* 1e-9 */
  double tolerance;       /* relative residual of PCG */
  int max_iterations;     /* 0 -> 1000 */
  int warm_start;         /* 1 */
  /* Heaviside projection of the filtered density (ComplianceObjective::set_projection,
   * src/topopt/Sensitivity.cpp:42-56; src/topopt/Projection.cpp:58-94): physical =
   * (tanh(beta eta) + tanh(beta (rho - eta))) / (tanh(beta eta) + tanh(beta (1 - eta))).
   * beta = 0: off (default). */
  double beta;
  double eta;             /* 0.5 */
} slx_eval_options;

typedef struct {
  double compliance;
  double volume;
  double volume_fraction;
  int linear_iterations;
  double assemble_seconds;
  double solve_seconds;
} slx_evaluation;

/* include/sparlab/topopt/OptimalityCriteria.hpp:35-52 */
typedef struct {
  double move_limit;       /* 0.2 */
  double damping;          /* 0.5 */
  double lambda_lower;     /* 1e-12 */
  double lambda_upper;     /* 1e12 */
  int max_bisections;      /* 200 */
  double volume_tolerance; /* 1e-10 */
} slx_oc_options;

typedef struct {
  double lambda;
  double achieved_volume;
  double max_change;
  int bisections;
  int volume_converged;
} slx_oc_step;

/* Beta continuation of the projection (ProjectionOptions,
 * include/sparlab/topopt/Projection.hpp:68-97; the schedule of
 * src/topopt/TopologyOptimizer.cpp:17-46). Stage k runs at
 * min(beta_start beta_factor^k, beta_max); a stage ends after beta_interval
 * iterations or, with advance_on_convergence, when the design change reaches
 * change_tolerance. Only the last stage may stop the loop. */
typedef struct {
  int enabled;                 /* 0 */
  double eta;                  /* 0.5 */
  double beta_start;           /* 1 */
  double beta_max;             /* 32 */
  double beta_factor;          /* 2 */
  int beta_interval;           /* 50 */
  int advance_on_convergence;  /* 1 */
  /* robust formulation (TopologyOptimizer.cpp:60-80,566-640): the objective on the
   * eroded design (eta + delta), the volume constraint on the dilated design
   * (eta - delta) against V* V_dilated / V_blueprint, rescaled every
   * robust_volume_interval iterations; the blueprint (eta) is the result */
  int robust;                  /* 0 */
  double robust_delta;         /* 0.1 */
  int robust_volume_interval;  /* 1 */
} slx_projection_options;

/* include/sparlab/topopt/TopologyOptimizer.hpp:81-121 (OC subset) */
typedef struct {
  int max_iterations;
  double change_tolerance;
  double objective_tolerance;
  int objective_window;
  int continuation_steps;
  double penalty_start;
  int continuation_iterations;
  slx_oc_options oc;
  slx_eval_options eval;       /* its beta is ignored: the projection schedule sets it */
  slx_projection_options projection;
} slx_run_options;

/* TopologyOptimizationResult (include/sparlab/topopt/TopologyOptimizer.hpp:193-258,
 * the OC fields): the figures of the final evaluation at the returned design
 * (src/topopt/TopologyOptimizer.cpp:263-283) and how the loop ended
 * (src/topopt/TopologyOptimizer.cpp:672-701). */
typedef enum {
  SLX_STOP_ITERATION_CAP = 0,   /* "iteration_cap" */
  SLX_STOP_DESIGN_CHANGE = 1,   /* "design_change" */
  SLX_STOP_OBJECTIVE_STALL = 2  /* "objective_stall" */
} slx_stop_reason;

typedef struct {
  int iterations;                /* performed */
  int converged;
  int stop_reason;               /* slx_stop_reason */
  double compliance;             /* J, weighted over the load cases */
  double volume;                 /* m^3 of the physical density */
  double volume_fraction;
  double grey_level;             /* 4/n sum rho (1 - rho) of the physical density */
  double final_design_change;
  double final_objective_change; /* relative compliance spread over the window (inf before it fills) */
  long long linear_solves;       /* one per load case and evaluation, the final one included */
  long long linear_iterations;
  double total_seconds;
  double final_beta;             /* beta of the last stage (0 without projection) */
  /* robust runs (RobustRecord, TopologyOptimizer.hpp): the designs projected at
   * eta +- delta next to the blueprint the fields above describe */
  double compliance_eroded;
  double compliance_dilated;
  double volume_fraction_eroded;
  double volume_fraction_dilated;
  double dilated_target_fraction; /* of the last rescaling */
} slx_run_result;

/* Columns of one optimisation history row, in the order of the reference's
 * history.csv (src/io/ResultWriter.cpp:1428-1470, OC): the 12 columns of a run
 * without projection, then beta (0 without projection), then the eroded and
 * dilated volume fractions and the dilated target fraction of a robust run (0
 * otherwise; the volume columns then hold the blueprint's). */
#define SLX_HISTORY_COLUMNS 16

/* Learned warm start: a 3-level 3-D convolutional network on the node grid maps
 * per-node (density, density change, previous displacement, residual of the
 * previous displacement, previous displacement change s) to the direction d of
 * the part of the coming displacement change that s does not explain. It trains
 * on the first train_solves solves of the evaluation sequence (loss weighted by
 * the nodes' stiffness factors) and then proposes d for every solve; CG starts
 * from x0 = u_prev + a d + b s with (a, b) from the Galerkin condition on
 * span(d, s) -- the point of that plane closest to the solution in the K-norm,
 * so x0 is never further from it than u_prev. With train_solves = 0 and
 * min_samples beyond reach only s is used. CG then runs to the unchanged
 * tolerance. */
typedef struct {
  int enable;             /* 0: plain warm start */
  int train_solves;       /* solves whose results train the network (40) */
  int min_samples;        /* training samples before the first prediction (4) */
  int steps_per_solve;    /* Adam steps after each training solve (16) */
  int buffer;             /* replay samples kept on the device (8) */
  double learning_rate;   /* 1e-3 */
  unsigned long long seed;
} slx_warmstart_options;

typedef struct {
  int samples;                /* training samples collected */
  int train_steps;            /* Adam steps taken */
  double first_loss;          /* loss of the first training step */
  double last_loss;           /* loss of the last training step */
  int predictions;            /* solves started from x0 */
  int rejected;               /* proposals with no usable direction (CG started from u_prev) */
  double step_sum;            /* sum of the network coefficients a */
  double energy_ratio_sum;    /* sum of ||u - x0||_K / ||u - u_prev||_K, u the converged solution */
  double last_energy_ratio;
  double net_ratio_sum;       /* the same ratio for the network's direction alone */
  double secant_ratio_sum;    /* and for the previous change alone */
  double residual_ratio_sum;  /* sum of ||b - K x0|| / ||b - K u_prev|| */
  double train_seconds;
  double predict_seconds;
} slx_warmstart_stats;

/* ---- lifecycle ------------------------------------------------------------ */
slx_status slx_create(int device, slx_context** out);
void slx_destroy(slx_context* ctx);
const char* slx_last_error(const slx_context* ctx);
void slx_amg_default_options(slx_amg_options* o);
void slx_eval_default_options(slx_eval_options* o);
void slx_oc_default_options(slx_oc_options* o);

/* ---- model: structured Hex8 (dim 3) or Q4 plane-stress (dim 2) grid -------
 * Node / element numbering of src/mesh/StructuredMesh.hpp:6-27. */
slx_status slx_model_structured(slx_context* ctx, int dim, int nx, int ny, int nz, double lx, double ly,
                                double lz, double E, double nu, double thickness);
/* fixed[ndofs]: 1 for a prescribed (zero) DOF, DOF = node * dim + component. */
slx_status slx_model_set_supports(slx_context* ctx, const unsigned char* fixed);
/* loads[ncases * ndofs] full-length load vectors; weights are normalised to sum 1. */
slx_status slx_model_set_loads(slx_context* ctx, int ncases, const double* loads, const double* weights);
/* Builds Ke0, the DOF map and the K_ff pattern on the device. */
slx_status slx_model_finalize(slx_context* ctx);
/* out[0..5] = nnodes, nelem, ndofs, nfree, kff_nnz, ncases */
slx_status slx_model_sizes(const slx_context* ctx, long long* out);
slx_status slx_ke0_get(const slx_context* ctx, double* ke);
slx_status slx_free_dofs_get(const slx_context* ctx, int* free_dofs);

/* ---- assembly: K_ff = reduce(sum_e scale_e Ke0) on the device -------------- */
slx_status slx_assemble(slx_context* ctx, const double* scale /* nelem, host, or NULL for ones */);
slx_status slx_kff_get(const slx_context* ctx, int* ptr, int* col, double* val);
/* y = K_ff x with the operator the solver uses (host arrays of nfree); matrix_free
 * (optional) reports whether it ran matrix-free (structured Hex8 after slx_assemble,
 * unless SPARLAB_NO_MATRIX_FREE=1 was set at assembly) or on the assembled CSR. */
slx_status slx_kff_apply(slx_context* ctx, const double* x, double* y, int* matrix_free);
/* slx_kff_apply without any device allocation, and x.K_ff x into *xkx (optional).
 * No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, hosts that forbid
 * hipMalloc). x and y are staged through one file-scope device array per GPU
 * (35.7 MB, held until the process ends): at most 2 097 152 unknowns; larger
 * systems return SLX_CONFIG_ERROR. Every call on a GPU shares that array, so calls
 * that overlap on one GPU (from several threads or contexts) race. */
slx_status slx_kff_apply_noalloc(slx_context* ctx, const double* x, double* y, double* xkx, int* matrix_free);

/* ---- linear solver ---------------------------------------------------------- */
slx_status slx_amg_setup(slx_context* ctx, const slx_amg_options* opt);
slx_status slx_factorize_csr(slx_context* ctx, int n, const int* ptr, const int* col, const double* val, int dim,
                             int ndpn, int nnodes, const double* coords, const int* unknowns,
                             const slx_amg_options* opt);
slx_status slx_solve(slx_context* ctx, const double* b, double* x, double tol, int max_iterations,
                     slx_pcg_result* res);
slx_status slx_solve_case(slx_context* ctx, int load_case, int warm_start, double tol, int max_iterations,
                          slx_pcg_result* res);
/* Batched solve (no counterpart in the reference, which solves designs and load
 * cases one at a time): the factorized matrix holds independent systems on its
 * diagonal blocks -- designs of different sizes, or copies of one design's K_ff
 * for its load cases -- with rows [seg_ptr[s], seg_ptr[s+1]) in segment s
 * (seg_ptr[0] = 0, seg_ptr[nseg] = n). Each segment runs PCG with its own step
 * lengths, residual, convergence test and restarts (the semantics of slx_solve);
 * every kernel launch covers all segments. A matrix that couples two segments
 * is rejected (SLX_CONFIG_ERROR). results[nseg]. */
slx_status slx_solve_segmented(slx_context* ctx, int nseg, const int* seg_ptr, const double* b, double* x, double tol,
                               int max_iterations, slx_pcg_result* results);
slx_status slx_amg_apply(slx_context* ctx, const double* r, double* z);
/* Measurement (no counterpart in the reference): the average time of one
 * Chebyshev step of a level above the coarsest (level 0: the matrix-free
 * operator's step when the level-0 products are matrix-free), over reps
 * back-to-back launches between two device events. It overwrites that level's
 * work vectors (the next V-cycle sets them again). */
slx_status slx_amg_time_smoother(slx_context* ctx, int level, int reps, double* ms_per_step);
slx_status slx_amg_stats_get(const slx_context* ctx, slx_amg_stats* out);
/* Test access to the hierarchy. which: 0 A_l, 1 P, 2 P-hat, 3 R, 4 A P, 5 R (A P):
 * info = rows, cols, nnz (slx_amg_level_get: 0..3 only); 6, 7: the device plans of
 * the products A P and R (A P): info = output nodes, runs, entries (all 0 when the
 * level uses the generic product); 8: the level's operator: nodes, length of the
 * shared column lists (0: none), lanes per row of its CSR products; 9 (level 0): the
 * nodes per direction of the matrix-free operator's tile the setup timed fastest for
 * the V-cycle's precision (all 0 when level 0 is assembled). */
slx_status slx_amg_level_info(const slx_context* ctx, int level, int which, long long* info);
slx_status slx_amg_level_get(const slx_context* ctx, int level, int which, int* ptr, int* col, double* val);
slx_status slx_amg_aggregates_get(const slx_context* ctx, int level, int* aggregates);
/* u[nfree] of the last solve of a load case */
slx_status slx_solution_get(const slx_context* ctx, int load_case, double* u);

/* ---- topology optimisation ------------------------------------------------- */
slx_status slx_design_set(slx_context* ctx, double volume_fraction, const double* lower, const double* upper,
                          const unsigned char* passive);
slx_status slx_filter_build(slx_context* ctx, double radius);
slx_status slx_filter_apply(slx_context* ctx, int adjoint, const double* in, double* out);
slx_status slx_evaluate(slx_context* ctx, const double* x, const slx_eval_options* opt, slx_evaluation* out,
                        double* dc_dx, double* dv_dx, double* energy);
slx_status slx_oc_update(slx_context* ctx, const double* x, const double* dc_dx, const double* dv_dx,
                         const slx_oc_options* opt, double* x_new, slx_oc_step* step);
/* x[nelem] in: the initial design, out: the final one. history[max_iterations *
 * SLX_HISTORY_COLUMNS] (may be NULL), one row per performed iteration:
 * iteration, penalty, compliance [J], volume [m^3], volume fraction, max design
 * change, Lagrange multiplier, grey level, OC bisections, volume converged (0/1),
 * seconds, linear iterations, beta, eroded / dilated volume fraction, dilated
 * target fraction. */
slx_status slx_run_oc(slx_context* ctx, const slx_run_options* opt, double* x, double* history,
                      slx_run_result* result);
/* ComplianceObjective::solve_adjoint (src/topopt/Sensitivity.cpp:239-257): solves
 * K_ff lambda = rhs on the operator of the last evaluation (rhs and lambda full
 * length ndofs, lambda zero on the fixed DOFs) with the context's AMG-PCG to tol.
 * slot >= 0 warm-starts from that slot's previous solution and stores the new
 * one; slot < 0 starts from zero and stores nothing. */
slx_status slx_solve_adjoint(slx_context* ctx, const double* rhs, int slot, double tol, int max_iterations,
                             double* lambda, slx_pcg_result* result);
/* compliance[ncases]: f_l . u_l of each load case at the last evaluation
 * (ObjectiveEvaluation::load_case_compliance, include/sparlab/topopt/Sensitivity.hpp:78) */
slx_status slx_load_case_compliance_get(const slx_context* ctx, double* compliance);

/* ---- free vibration ----------------------------------------------------------
 * ModalAnalysisOptions and ModalResult (include/sparlab/fem/ModalAnalysis.hpp:44-90).
 * The shifted systems are solved with the context's AMG-PCG (the options of the
 * last slx_amg_setup, or the defaults), each warm started from the converging
 * subspace when warm_start is set. A system of at most 400 free DOFs, or one
 * the subspace covers, takes the dense generalised eigensolve instead
 * (iterations 0), as the reference does. Only shift_factor = 0 (the
 * reference's default) is supported. */
typedef struct {
  int num_modes;              /* lowest modes requested (6) */
  int lumped;                 /* 0: consistent mass (default), 1: row-sum lumped */
  int max_iterations;         /* subspace iterations (200) */
  double tolerance;           /* relative eigenvalue change (1e-10) */
  double residual_tolerance;  /* eigenpair residual |K phi - lambda M phi| / |lambda M phi| (1e-6) */
  double rigid_body_ratio;    /* eigenvalues below this times max K_ii / max M_ii are flagged (1e-10) */
  double shift_factor;        /* must be 0 */
  double density;             /* kg/m^3, > 0 */
  double linear_tolerance;    /* relative residual of each AMG-PCG solve (1e-8) */
  int linear_max_iterations;  /* per solve (1000) */
  int warm_start;             /* start each solve from the last Ritz pair (1; linear.warm_start) */
} slx_modal_options;

typedef struct {
  int num_modes;              /* returned modes: min(num_modes, free DOFs) */
  int subspace_size;          /* q = min(n, max(2 m, m + 8)) */
  int iterations;
  int converged;
  int rigid_body_modes;       /* modes flagged as rigid-body / mechanism (warnings in the reference) */
  int residual_warnings;      /* modes whose final residual exceeds residual_tolerance */
  long long linear_iterations;
  double final_change;
  double total_mass;          /* kg: sum over elements of scale * density * volume */
  double seconds;
} slx_modal_result;

void slx_modal_default_options(slx_modal_options* o);
/* Lowest eigenpairs of (K_ff, M_ff) for the context's model. stiffness_scale,
 * mass_scale [nelem]: per-element factors (NULL: 1). eigenvalues[num_modes]
 * (1/s^2, ascending), residuals[num_modes], shapes[ndofs * num_modes] (may be
 * NULL; column-major, M-orthonormal, zero on the fixed DOFs). Leaves the
 * context's K_ff assembled and factorized for stiffness_scale. */
slx_status slx_modal_solve(slx_context* ctx, const slx_modal_options* opt, const double* stiffness_scale,
                           const double* mass_scale, double* eigenvalues, double* residuals, double* shapes,
                           slx_modal_result* result);

/* ---- learned warm start ------------------------------------------------------ */
void slx_warmstart_default_options(slx_warmstart_options* o);
/* Structured models only; resets the network and its training state. */
slx_status slx_warmstart_configure(slx_context* ctx, const slx_warmstart_options* opt);
slx_status slx_warmstart_stats_get(const slx_context* ctx, slx_warmstart_stats* out);

/* The network on an arbitrary X x Y x Z grid (tests). Parameters are FP32,
 * per layer: weights [cout][27][cin] then biases [cout], layers in order
 * (cin, cout) = (16,16) (16,32) (32,64) (64,32) (32,16) (16,16); tap =
 * (dz+1)*9 + (dy+1)*3 + (dx+1). x is [V][nin], y and t are [V][nout], voxel
 * v = (z*Y + y)*X + x. */
slx_status slx_net_configure(slx_context* ctx, int X, int Y, int Z, int nin, int nout, unsigned long long seed);
slx_status slx_net_parameter_count(const slx_context* ctx, long long* count);
slx_status slx_net_parameters_get(slx_context* ctx, float* params);
slx_status slx_net_parameters_set(slx_context* ctx, const float* params);
slx_status slx_net_forward(slx_context* ctx, const float* x, float* y);
/* loss = sum_v w_v sum_k (y - t)^2 / (nout sum_v w_v) (w = 1 when NULL) and its
 * gradient with respect to every parameter */
slx_status slx_net_gradients(slx_context* ctx, const float* x, const float* t, const float* w, double* loss,
                             float* grad);
/* steps Adam steps on one (x, t, w) triple; losses before the first and after the last step */
slx_status slx_net_train(slx_context* ctx, const float* x, const float* t, const float* w, int steps,
                         double learning_rate, double* loss_first, double* loss_last);

/* ---- measurement ------------------------------------------------------------
 * With timing on, every launch of the named kernels is bracketed by HIP events
 * on the context's stream and the elapsed times are summed. Names: "spmv"
 * (finest-level operator products of PCG and the V-cycle), "assemble". */
slx_status slx_timing_enable(slx_context* ctx, int on);
slx_status slx_timing_get(slx_context* ctx, const char* kernel, double* total_ms, long long* launches);
slx_status slx_synchronize(slx_context* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SPARLAB_HIP_H */
