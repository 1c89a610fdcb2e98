/* A plain C caller of the C ABI (include/sparlab_hip.h), built by
 * tests/test_gpu_capi_caller.py with gcc and linked against
 * libsparlab_hip.so. It runs the LinearSolver seam the way a C or C++ host
 * would bind it: the sparlab_bench Hex8 box (apps/sparlab_bench.cpp:87-127)
 * assembled through the structured entry points in one context, its K_ff
 * handed to slx_factorize_csr in a second context with the DOF layout, solved
 * by slx_solve, and the AMG statistics read back. It also walks the error
 * paths a caller sees. Every result is printed as "key value" lines.
 *
 * usage: capi_caller NX  (ny = NX / 2, nz = NX / 4)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sparlab_hip.h"

static int check(slx_context* c, slx_status s, const char* what) {
  if (s != SLX_OK) {
    printf("error %s status %d message %s\n", what, (int)s, slx_last_error(c));
    return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  const int nx = argc > 1 ? atoi(argv[1]) : 16;
  const int ny = nx / 2, nz = ny / 2;
  const int nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const long long nnodes = (long long)nnx * nny * nnz, ndofs = 3 * nnodes;
  slx_context *a = NULL, *b = NULL;
  if (slx_create(0, &a) != SLX_OK || slx_create(0, &b) != SLX_OK) {
    printf("error create\n");
    return 1;
  }
  /* the structured model: x = 0 clamped, -1000 N along y at the node nearest (2, 0, 0) */
  if (check(a, slx_model_structured(a, 3, nx, ny, nz, 2.0, 1.0, 0.5, 70.0e9, 0.3, 1.0), "structured")) return 1;
  unsigned char* fixed = calloc((size_t)ndofs, 1);
  double* load = calloc((size_t)ndofs, sizeof(double));
  for (long long n = 0; n < nnodes; ++n)
    if (n % nnx == 0) fixed[3 * n] = fixed[3 * n + 1] = fixed[3 * n + 2] = 1;
  load[3 * (long long)nx + 1] = -1000.0;
  const double w = 1.0;
  if (check(a, slx_model_set_supports(a, fixed), "supports")) return 1;
  if (check(a, slx_model_set_loads(a, 1, load, &w), "loads")) return 1;
  if (check(a, slx_model_finalize(a), "finalize")) return 1;
  long long sizes[6];
  if (check(a, slx_model_sizes(a, sizes), "sizes")) return 1;
  const int nf = (int)sizes[3];
  const long long nnzk = sizes[4];
  int* ptr = malloc(sizeof(int) * ((size_t)nf + 1));
  int* col = malloc(sizeof(int) * (size_t)nnzk);
  double* val = malloc(sizeof(double) * (size_t)nnzk);
  int* freed = malloc(sizeof(int) * (size_t)nf);
  if (check(a, slx_assemble(a, NULL), "assemble")) return 1;
  if (check(a, slx_kff_get(a, ptr, col, val), "kff_get")) return 1;
  if (check(a, slx_free_dofs_get(a, freed), "free_dofs")) return 1;
  printf("free_dofs %d\nkff_nnz %lld\n", nf, nnzk);

  /* error path: a solve before any factorisation */
  double* rhs = malloc(sizeof(double) * (size_t)nf);
  double* x = calloc((size_t)nf, sizeof(double));
  for (int i = 0; i < nf; ++i) rhs[i] = load[freed[i]];
  slx_pcg_result res;
  slx_status s = slx_solve(b, rhs, x, 1e-10, 0, &res);
  printf("solve_before_factorize_status %d\n", (int)s);
  printf("solve_before_factorize_message_nonempty %d\n", (int)(strlen(slx_last_error(b)) > 0));

  /* the seam: factorize the caller's K_ff with the DOF layout (node coordinates, free DOFs) */
  double* coords = malloc(sizeof(double) * 3 * (size_t)nnodes);
  for (long long n = 0; n < nnodes; ++n) {
    const long long i = n % nnx, j = (n / nnx) % nny, k = n / ((long long)nnx * nny);
    coords[3 * n + 0] = (double)i * (2.0 / nx);
    coords[3 * n + 1] = (double)j * (1.0 / ny);
    coords[3 * n + 2] = (double)k * (0.5 / nz);
  }
  slx_amg_options opt;
  slx_amg_default_options(&opt);
  if (check(b, slx_factorize_csr(b, nf, ptr, col, val, 3, 3, (int)nnodes, coords, freed, &opt), "factorize")) return 1;
  if (check(b, slx_solve(b, rhs, x, 1e-10, 0, &res), "solve")) return 1;
  double c = 0.0;
  for (int i = 0; i < nf; ++i) c += rhs[i] * x[i];
  printf("compliance %.17g\niterations %d\nrelative_residual %.6e\nconverged %d\n", c, res.iterations,
         res.relative_residual, res.converged);
  slx_amg_stats st;
  if (check(b, slx_amg_stats_get(b, &st), "stats")) return 1;
  printf("levels %d\noperator_complexity %.17g\nstorage_nonzeros %lld\nnonzeros0 %lld\n", st.levels,
         st.operator_complexity, st.storage_nonzeros, st.nonzeros[0]);
  printf("setup_phases_nonnegative %d\n",
         (int)(st.setup_seconds >= 0 && st.aggregation_seconds >= 0 && st.tentative_seconds >= 0 &&
               st.galerkin_seconds >= 0));

  /* a context that held the structured model refuses it once a caller matrix replaces K_ff */
  if (check(a, slx_factorize_csr(a, nf, ptr, col, val, 3, 3, (int)nnodes, coords, freed, &opt), "factorize_a"))
    return 1;
  s = slx_solve_case(a, 0, 1, 1e-10, 0, &res);
  printf("solve_case_after_caller_matrix_status %d\n", (int)s);
  slx_eval_options eo;
  slx_eval_default_options(&eo);
  const long long ne = sizes[1];
  double* design = malloc(sizeof(double) * (size_t)ne);
  for (long long e = 0; e < ne; ++e) design[e] = 0.5;
  slx_evaluation ev;
  s = slx_evaluate(a, design, &eo, &ev, NULL, NULL, NULL);
  printf("evaluate_after_caller_matrix_status %d\n", (int)s);
  /* finalising again restores the structured model */
  if (check(a, slx_model_finalize(a), "refinalize")) return 1;
  eo.tolerance = 1e-10;
  s = slx_evaluate(a, design, &eo, &ev, NULL, NULL, NULL);
  printf("evaluate_after_refinalize_status %d\nevaluate_compliance %.17g\n", (int)s, ev.compliance);

  slx_destroy(a);
  slx_destroy(b);
  free(fixed); free(load); free(ptr); free(col); free(val); free(freed); free(rhs); free(x); free(coords);
  free(design);
  printf("done 1\n");
  return 0;
}
