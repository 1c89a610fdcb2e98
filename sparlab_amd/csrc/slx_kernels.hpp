// Host-callable launchers of the HIP kernels (kernels.hip). Every launcher
// enqueues on the given stream and returns immediately.
#pragma once

#include "slx_internal.hpp"

namespace slx {
namespace k {

// The tentative prolongator's aggregate blocks (built in amg.cpp): per fine row
// c0 * 8 + len (its len <= 6 consecutive columns from c0), values column-major
// (ell[k * n + r]); per aggregate its fine rows (agg_ptr, agg_rows) and c0 * 8 + len.
struct TentBlocks {
  int n = 0, naggs = 0;
  const int* c0len = nullptr;
  const int* agg_ptr = nullptr;
  const int* agg_rows = nullptr;
  const int* agg_c0len = nullptr;
};
// y = P-hat xc (ell: values by fine row) and xc = P-hat^T y (ella: the same
// values in the aggregates' member order, ella[k * n + i] for agg_rows[i]);
// with dinv set, tent_restrict also starts the coarse level's smoother from a
// zero guess: d = dinv xc / theta, x2 = d
template <typename T>
void tent_prolong(const TentBlocks& t, const T* ell, const T* xc, T* y, hipStream_t s);
template <typename T>
void tent_restrict(const TentBlocks& t, const T* ella, const T* y, T* xc, hipStream_t s, const T* dinv = nullptr,
                   double theta = 1.0, T* d = nullptr, T* x2 = nullptr);


// operator products: y = A x, y += A x, r = b - A x (FP64 and FP32 storage).
// nv > 1: nv load cases side by side (case-major vectors, x at stride A.cols,
// b, r and y at stride A.rows), each case's sums in the single-case order.
void spmv(const DCsr& A, const double* x, double* y, hipStream_t s, int nv = 1);
void spmv_add(const DCsr& A, const double* x, double* y, hipStream_t s, int nv = 1);
void residual(const DCsr& A, const double* x, const double* b, double* r, hipStream_t s, int nv = 1);
void spmv32(const DCsr& A, const float* x, float* y, hipStream_t s, int nv = 1);
void spmv_add32(const DCsr& A, const float* x, float* y, hipStream_t s, int nv = 1);
void residual32(const DCsr& A, const float* x, const float* b, float* r, hipStream_t s, int nv = 1);
// The smoother start fused into a product: spmv_start forms y = A x and then
// the zero-guess start of y's level, d = dinv y / theta and x2 = d (the result
// of a restriction, cheb_first_zero in one pass); residual_first forms
// r = b - A x and then d = dinv r / theta, x2 = x + d (cheb_first after the
// post-smoothing residual; x2 must not alias x). Same bits as the two passes.
void spmv_start(const DCsr& A, const double* x, double* y, const double* dinv, double theta, double* d, double* x2,
                hipStream_t s, int nv = 1);
void residual_first(const DCsr& A, const double* x, const double* b, double* r, const double* dinv, double theta,
                    double* d, double* x2, hipStream_t s, int nv = 1);
void spmv_start32(const DCsr& A, const float* x, float* y, const float* dinv, double theta, float* d, float* x2,
                  hipStream_t s, int nv = 1);
void residual_first32(const DCsr& A, const float* x, const float* b, float* r, const float* dinv, double theta,
                      float* d, float* x2, hipStream_t s, int nv = 1);
// y = A x and per-block partial sums of x.y (spmv_dot_blocks(A) partials)
int spmv_dot_blocks(const DCsr& A);
void spmv_dot(const DCsr& A, const double* x, double* y, double* partials, hipStream_t s, int nv = 1);
// one thread per row, column-order sums (bitwise equal to the CPU restatement)
void spmv_serial(const DCsr& A, const double* x, double* y, hipStream_t s);

// Chebyshev smoother pieces
void cheb_step(const DCsr& A, const double* dinv, const double* r_in, double* r_out, const double* d_in, double* d_out,
               double* x, double c1, double c2, hipStream_t s, int nv = 1);
void cheb_step32(const DCsr& A, const float* dinv, const float* r_in, float* r_out, const float* d_in, float* d_out,
                 float* x, double c1, double c2, hipStream_t s, int nv = 1);
void cheb_first_zero(int n, const double* dinv, const double* b, double theta, double* x, double* d, hipStream_t s,
                     int nv = 1);
void cheb_first(int n, const double* dinv, const double* r, double theta, double* x, double* d, hipStream_t s,
                int nv = 1);
void cheb_first_zero32(int n, const float* dinv, const float* b, double theta, float* x, float* d, hipStream_t s,
                       int nv = 1);
void cheb_first32(int n, const float* dinv, const float* r, double theta, float* x, float* d, hipStream_t s,
                  int nv = 1);
void dense_gemv(int n, const double* M, const double* b, double* y, hipStream_t s, int nv = 1);
// a[n*n] (row-major) = the inverse of the symmetric positive definite CSR matrix
// (ptr, col, val), by blocked in-place Gauss-Jordan on the device; piv[n] = the
// scalar pivots (the D of its unpivoted LDL^T). work: dense_inverse_work(n) doubles.
int dense_inverse_work(int n);
void dense_inverse(int n, const int* ptr, const int* col, const double* val, double* a, double* piv, double* work,
                   hipStream_t s);
void dense_gemv32(int n, const float* M, const float* b, float* y, hipStream_t s, int nv = 1);

// deterministic reductions: partial buffer of at least 2048 doubles
int reduce_blocks(long long n);
// nv > 1: case v's dot (vectors at stride n) into out[v * out_stride]; partials
// then hold nv * reduce_blocks(n) doubles
void dot(long long n, const double* a, const double* b, double* partials, double* out, hipStream_t s, int nv = 1,
         int out_stride = 0);
// out[0] = sum_i rho_i (1 - rho_i) (the grey level times n / 4)
void grey_sum(long long n, const double* rho, double* partials, double* out, hipStream_t s);
void reduce(int m, const double* partials, double* out, int take_sqrt, hipStream_t s, int nv = 1,
            int out_stride = 0);
// x += alpha p, r -= alpha q, out = r.r with alpha = scal[i_rz] / scal[i_pq].
// nv > 1: case v at vector stride n, scalars at scal + 16 v and out + 16 v; a case
// with active[v] == 0 (active may be null) keeps x and r and only recomputes r.r
void pcg_update(long long n, double* x, double* r, const double* p, const double* q, const double* scal, int i_rz,
                int i_pq, double* partials, double* out, hipStream_t s, int nv = 1, const int* active = nullptr);
// p = z + beta p, beta = scal[i_new] / scal[i_old] (or scal[i_new] when i_old < 0)
void pcg_direction(long long n, double* p, const double* z, const double* scal, int i_new, int i_old, hipStream_t s,
                   int nv = 1);
// z = zc (null: z is already the cycle's result), r.z (and r.z_old when
// flexible), beta into scal[i_beta], scal[i_rzn] = scal[i_rz] = r.z;
// partials: 2 per block (reduce_blocks(n))
void precond_tail(long long n, const float* zc, const double* r, double* z, double* scal, int i_rz, int i_rzn,
                  int i_beta, bool flexible, double* partials, hipStream_t s, int nv = 1);
void cheb_first_zero_from64(int n, const float* dinv, const double* r, double theta, float* b, float* x, float* d,
                            hipStream_t s, int nv = 1);
void copy_scalar(double* scal, int from, int to, hipStream_t s);
void flexible_beta(double* scal, int i_new, int i_cross, int i_old, int i_out, hipStream_t s);
void to_float(long long n, const double* in, float* out, hipStream_t s);
void to_double(long long n, const float* in, double* out, hipStream_t s);
void to_bf16(long long n, const double* in, uint16_t* out, hipStream_t s);
void axpy(long long n, double a, const double* x, double* y, hipStream_t s);

// structured-grid assembly of the K_ff values into K's pattern; Hex8 with the
// packed node map (kmf::node_map) goes node block by node block (the same values;
// SPARLAB_ROW_ASSEMBLY=1 keeps the per-entry kernel)
void assemble_structured(int dim, const DCsr& K, const int* free_dofs, const double* scale, const double* ke0, int nx,
                         int ny, int nz, hipStream_t s,
                         const unsigned* xyzc = nullptr, const int* nodemap = nullptr);
// element loop of the sensitivity
void sensitivity(int dim, int nelem, int nx, int ny, const double* ke0, const int* reduced, const double* u,
                 long long nfree, int ncases, const double* w, const double* fac, const double* dfac, double* energy,
                 double* dc, hipStream_t s);
void simp(int n, const double* filtered, double penalty, double emin, double* phys, double* fac, double* dfac,
          hipStream_t s);
// out = Heaviside projection of clamp(in) (in == out allowed), dproj its derivative (may be null)
void project(int n, const double* in, double beta, double eta, double* out, double* dproj, hipStream_t s);
void mul(int n, const double* a, const double* b, double* c, hipStream_t s);
void oc_candidate(int n, const double* x, const double* dc, const double* dv, const double* lower, const double* upper,
                  const unsigned char* passive, double lambda, double move, double damping, double* xn,
                  hipStream_t s);
// one OC bisection trial on the device (state: lo, hi, volume, lambda, trials, done, converged);
// explicit_lambda: evaluate lam and only record the volume
void oc_trial(int n, const double* x, const double* dc, const double* dv, const double* lower, const double* upper,
              const unsigned char* passive, double move, double damping, const double* w, double* state,
              int explicit_lambda, double lam, double target, double vol_tol, int max_bisections, double* xn,
              double* partials, hipStream_t s);
int oc_trial_blocks(int n);
int filter_volume_blocks(const DCsr& H);
void filter_volume(const DCsr& H, const double* x, const double* vol, double* partials, double* out, hipStream_t s);
void max_abs_diff(int n, const double* a, const double* b, double* partials, double* out, hipStream_t s);
void clamp(int n, double* x, const double* lo, const double* hi, hipStream_t s);
// modal analysis (kernels_modal.hip). y = M x on the free DOFs of a structured
// grid (rho_v = density * element volume; scale: per-element factors or null),
// or its diagonal; lumped: row-sum lumping
void mass_apply(int dim, int nx, int ny, int nz, double rho_v, bool lumped, bool diagonal, const double* scale,
                const int* reduced, const double* x, double* y, hipStream_t s);
void csr_diagonal(const DCsr& A, double* d, hipStream_t s);
// doubles of partials gram() and eig_residuals() need for n rows and q columns
int modal_partials(int n, int q);
// out[i * q + j] = A_i . B_j for column-major n x q blocks A, B
void gram(int n, int q, const double* A, const double* B, double* partials, double* out, hipStream_t s);
// Y = X R: X n x q, R q x m column-major, Y n x m (q, m <= 64)
void block_combine(int n, int q, int m, const double* X, const double* R, double* Y, hipStream_t s);
// out[2 i] = |KX_i - mu_i MX_i|^2, out[2 i + 1] = |mu_i MX_i|^2 (mu on the device)
void eig_residuals(int n, int m, const double* KX, const double* MX, const double* mu, double* partials, double* out,
                   hipStream_t s);
void scale_copy(int n, double a, const double* x, double* y, hipStream_t s);

// Strength of connection at level 0 (kernels_amg.hip): per node and per run of
// its first row's pattern (one neighbour node), the sum of squares over the
// node's rows (run_ptr: the runs of each node, in row order)
void node_block_sums(int nn, const int* node_ptr, const int* node_of, const int* ptr, const int* col,
                     const double* val, const int* run_ptr, double* sums, hipStream_t s);
// *bad = min over asymmetric entries of (i << 32) | j (|a_ij - a_ji| > tol max|a|);
// set *bad to all ones before; scratch: 1025 doubles
void symmetry_check(const DCsr& A, double tol, double* scratch, unsigned long long* bad, hipStream_t s);

// SA-AMG numeric setup (kernels_amg.hip)
struct BlockProductArgs {
  long long nitems;           // C blocks (one per aggregate block of each C node-row)
  int dmax;                   // most rows in a C node (at most 6)
  int wmax;                   // widest aggregate block: lanes per item
  const int* item_node;       // C node of the item
  const int* item_agg;        // its aggregate
  const int* item_pos;        // position of the block within each row of the node
  const long long* item_begin;  // contribution list of item i: [item_begin[i], item_begin[i+1])
  // entry: x = index in B's values of the block in the run's first B row,
  // y = (offset of the run in A's row << 16) | (B row length << 3) | B rows
  const uint2* list;
  const int* out_node_ptr;    // rows of C node I: [out_node_ptr[I], out_node_ptr[I+1])
  const int* aptr;
  const double* aval;
  const double* bval;
  const int* cptr;
  double* cval;
  const int* agg_off;         // aggregate -> first coarse column (width = next - this)
  const double* dinv;         // prolongator smoothing only
  double omega;
  const int* tptr;
  const double* tval;
  const int* row_agg;         // aggregate of each fine row
};
void block_product(const BlockProductArgs& a, bool smooth, hipStream_t s);
// C = A B, one wavefront per C node (kernels_amg.hip k_row_product)
struct RowProductArgs {
  int nnodes;                 // C nodes
  int dmax, jmax, wmax;       // most rows per C node, most B rows per run, widest aggregate block
  int lenc_max, emax;         // longest C row, most entries of one C node
  int lena_max;               // longest A row
  int wpb;                    // wavefronts (C nodes) per block
  bool pre;                   // stage each node's entries in LDS
  bool sa;                    // stage each node's A rows in LDS (at most 3 rows per node)
  const int* order;           // C nodes in processing order
  const int* node_ptr;        // rows of C node I: [node_ptr[I], node_ptr[I+1])
  const int* aptr;
  const double* aval;
  const int* run_ptr;         // runs of C node I: [run_ptr[I], run_ptr[I+1])
  // run: x = index in B's values of the run's first B row,
  // y = (offset of the run in A's row << 16) | (B row length << 3) | B rows,
  // entries [z, w)
  const uint4* run;
  // entry: one B block of the run's B node: (position of its aggregate block
  // in C's row << 16) | (position in B's row << 3) | width
  const unsigned* entry;
  const double* bval;
  const int* cptr;
  double* cval;
};
// LDS bytes of one wavefront of k_row_product
inline size_t row_product_lds(int dmax, int lenc_max, int emax, int lena) {
  return sizeof(double) * ((size_t)dmax * ((size_t)lenc_max + (size_t)lena) + (size_t)(emax + 1) / 2);
}
void row_product(const RowProductArgs& a, hipStream_t s);
void spgemm_numeric(const DCsr& A, const DCsr& B, DCsr& C, int* error_flag, hipStream_t s);
void smooth_prolongator(const DCsr& A, const DCsr& Ptent, DCsr& P, const double* dinv, double omega, int* error_flag,
                        hipStream_t s);
void gather(long long n, const int* src, const double* in, double* out, hipStream_t s);
// R = P^T by coarse node (cnode: the coarse nodes' first rows, nnodes + 1; every
// row of a node has the pattern of its first and reads src + q)
void gather_nodes(int nnodes, const int* cnode, const int* rptr, const int* src, const double* in, double* out,
                  hipStream_t s);
void symmetrize_values(long long n, const int* ia, const int* ib, const double* raw, double* out, hipStream_t s);
void inv_diag(const DCsr& A, double* dinv, int* bad_row, hipStream_t s);
void vsqrt(int n, const double* in, double* out, hipStream_t s);
// w = s (A (s v)) - beta vprev with the partials of w.v: lanczos_blocks(n) of them
int lanczos_blocks(int n);
void lanczos_apply(const DCsr& A, const double* sq, const double* v, const double* vprev, const double* beta,
                   double* w, double* partials, hipStream_t s);
// blocks (and partials) of the Lanczos vector steps below
int axmy_blocks(int n);
// Lanczos vector steps with the scalar reductions folded in (every block sums
// the m partials of the previous kernel itself; block 0 stores the scalar):
// alpha = sum part_in, w -= alpha v, part_out = partials of w.w;
void axmy_dot_sum(int n, int m, const double* part_in, double* alpha, const double* v, double* w, double* part_out,
                  hipStream_t s);
// beta = sqrt(sum part_in), vprev = v, v = w / beta (axmy_blocks(n) blocks)
void lanczos_shift_sum(int n, int m, const double* part_in, double* beta, const double* w, double* v, double* vprev,
                       hipStream_t s);
void lanczos_shift(int n, const double* b, const double* w, double* v, double* vprev, hipStream_t s);
void sqrt_scalar(double* p, hipStream_t s);
void fill(long long n, double v, double* out, hipStream_t s);
void fill32(long long n, float v, float* out, hipStream_t s);
// one symmetric Gauss-Seidel sweep of A (forward or backward by the schedule
// given): rows[level_ptr[L] .. level_ptr[L + 1]) is dependency level L, run in
// order, one launch each (kernels_gs.hip); nv cases side by side (stride rows)
void gs_sweep(const DCsr& A, const int* rows, const std::vector<int>& level_ptr, const double* dinv, const double* b,
              double* x, hipStream_t s, int nv = 1);
void gs_sweep32(const DCsr& A, const int* rows, const std::vector<int>& level_ptr, const float* dinv, const float* b,
                float* x, hipStream_t s, int nv = 1);
void copy(long long n, const double* in, double* out, hipStream_t s);

}  // namespace k
}  // namespace slx
