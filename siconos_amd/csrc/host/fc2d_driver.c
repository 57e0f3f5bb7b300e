/* fc2d_driver.c — fc2d_driver, fc2d_compute_error and the
 * SICONOS_FRICTION_2D_NSGS_HIP solver: the reference's 2-D friction-contact
 * NSGS (fc2d_nsgs.c:297-538) on the GPU behind its driver and registry.
 *
 * The host turns W into scalar rows in the order the reference sums them
 * (fc2d_rows), then fc2d_hip.h solves on the device:
 *   NM_SPARSE_BLOCK (2x2 blocks): blocks in stored order, column 0 then 1 of
 *     each (mvp2x2, op3x3.h:316-334); the diagonal block is skipped by the
 *     sweep (SBM_row_prod_no_diag_2x2, SparseBlockMatrix.c:1492-1547);
 *   NM_SPARSE: the rows of the transposed CSC, i.e. ascending columns
 *     (NM_row_prod_no_diag2, NumericsMatrix.c:710-741, and the gemv
 *     CSparseMatrix_aaxpby, CSparseMatrix.c:109-140); the contact's own
 *     reaction enters as zero;
 *   NM_DENSE: every non-zero in ascending column order, summed from zero
 *     and then added to q (cblas_ddot, NumericsMatrix.c:682-700); the vendor
 *     BLAS order of the reference is not reproduced. */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fc2d_hip.h"
#include "fc2d_numerics.h"

typedef struct {
  int nc;
  int64_t* rp;   /* 2 nc + 1 */
  int64_t* col;
  double* val;
  double* diag;  /* 4 nc, column-major */
  int row_mode;
  int own_zero;
} fc2d_rows_t;

static void rows_free(fc2d_rows_t* R) {
  free(R->rp);
  free(R->col);
  free(R->val);
  free(R->diag);
}

/* scalar rows of an NM_SPARSE matrix: cs_transpose of its CSC */
static int rows_sparse(NumericsMatrix* M, fc2d_rows_t* R) {
  const CSparseMatrix* A = NM_csc(M);
  const int64_t n2 = 2 * (int64_t)R->nc;
  if (!A || A->m != n2 || A->n != n2) return NUMERICS_ERR_INVALID_ARGUMENT;
  const int64_t nnz = A->p[n2];
  R->rp = (int64_t*)calloc((size_t)n2 + 1, sizeof(int64_t));
  R->col = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  R->val = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  int64_t* w = (int64_t*)calloc((size_t)n2 + 1, sizeof(int64_t));
  if (!R->rp || !R->col || !R->val || !w) {
    free(w);
    return NUMERICS_ERR_MEMORY;
  }
  for (int64_t p = 0; p < nnz; p++) w[A->i[p]]++;
  for (int64_t i = 0; i < n2; i++) R->rp[i + 1] = R->rp[i] + w[i];
  for (int64_t i = 0; i < n2; i++) w[i] = R->rp[i];
  for (int64_t j = 0; j < n2; j++)
    for (int64_t p = A->p[j]; p < A->p[j + 1]; p++) {
      const int64_t q = w[A->i[p]]++;
      R->col[q] = j;
      R->val[q] = A->x[p];
    }
  free(w);
  /* NSM_extract_block (NumericsSparseMatrix.c:782-820): rows of a column
   * below the block are skipped, the first one past it ends the column */
  for (int c = 0; c < R->nc; c++) {
    double* B = R->diag + 4 * (size_t)c;
    const int64_t r0 = 2 * (int64_t)c;
    for (int64_t j = r0; j < r0 + 2; j++)
      for (int64_t p = A->p[j]; p < A->p[j + 1]; p++) {
        const int64_t row = A->i[p];
        if (row >= r0) {
          if (row >= r0 + 2) break;
          B[(j - r0) * 2 + row - r0] = A->x[p];
        }
      }
  }
  R->row_mode = FC2D_ROWS_ACCUMULATE;
  R->own_zero = 1;
  return 0;
}

static int rows_sbm(const SparseBlockStructuredMatrix* S, fc2d_rows_t* R) {
  const int nc = R->nc;
  if (!S || S->filled1 != (size_t)nc + 1 || !S->index1_data || !S->index2_data)
    return NUMERICS_ERR_INVALID_ARGUMENT;
  const size_t nb = S->index1_data[nc];
  /* every block puts two entries in each of its two rows */
  R->rp = (int64_t*)calloc(2 * (size_t)nc + 1, sizeof(int64_t));
  R->col = (int64_t*)malloc(sizeof(int64_t) * 4 * (nb ? nb : 1));
  R->val = (double*)malloc(sizeof(double) * 4 * (nb ? nb : 1));
  if (!R->rp || !R->col || !R->val) return NUMERICS_ERR_MEMORY;
  int64_t k = 0;
  for (int c = 0; c < nc; c++) {
    int found = 0;
    for (int i = 0; i < 2; i++) {
      for (size_t b = S->index1_data[c]; b < S->index1_data[c + 1]; b++) {
        const size_t d = S->index2_data[b];
        if (d >= (size_t)nc) return NUMERICS_ERR_INVALID_ARGUMENT;
        const double* blk = S->block[b];
        /* row i of the block: column 0 (entry i), then column 1 (entry 2 + i) */
        R->col[k] = 2 * (int64_t)d;
        R->val[k++] = blk[i];
        R->col[k] = 2 * (int64_t)d + 1;
        R->val[k++] = blk[2 + i];
        if (i == 0 && d == (size_t)c && !found) {
          memcpy(R->diag + 4 * (size_t)c, blk, 4 * sizeof(double));
          found = 1;
        }
      }
      R->rp[2 * (size_t)c + i + 1] = k;
    }
    if (!found) return NUMERICS_ERR_NOT_IMPLEMENTED; /* no diagonal block */
  }
  R->row_mode = FC2D_ROWS_ACCUMULATE;
  R->own_zero = 0;
  return 0;
}

static int rows_dense(const double* M, fc2d_rows_t* R) {
  const int64_t n2 = 2 * (int64_t)R->nc;
  int64_t nnz = 0;
  for (int64_t i = 0; i < n2; i++)
    for (int64_t j = 0; j < n2; j++)
      if (M[i + j * n2] != 0.0 || (j >> 1) == (i >> 1)) nnz++;
  R->rp = (int64_t*)calloc((size_t)n2 + 1, sizeof(int64_t));
  R->col = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  R->val = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  if (!R->rp || !R->col || !R->val) return NUMERICS_ERR_MEMORY;
  int64_t k = 0;
  for (int64_t i = 0; i < n2; i++) {
    for (int64_t j = 0; j < n2; j++)
      if (M[i + j * n2] != 0.0 || (j >> 1) == (i >> 1)) {
        R->col[k] = j;
        R->val[k++] = M[i + j * n2];
      }
    R->rp[i + 1] = k;
  }
  /* NM_extract_diag_block2, dense branch (NumericsMatrix.c:1802-1812) */
  for (int c = 0; c < R->nc; c++) {
    const double* p = M + (n2 + 1) * (2 * (int64_t)c);
    double* B = R->diag + 4 * (size_t)c;
    B[0] = p[0];
    B[1] = p[1];
    B[2] = p[n2];
    B[3] = p[n2 + 1];
  }
  R->row_mode = FC2D_ROWS_DOT;
  R->own_zero = 1;
  return 0;
}

static int fc2d_rows(FrictionContactProblem* problem, fc2d_rows_t* R) {
  memset(R, 0, sizeof *R);
  if (!problem || !problem->M || !problem->q || !problem->mu) return NUMERICS_ERR_NULL_POINTER;
  if (problem->dimension != 2 || problem->numberOfContacts <= 0)
    return NUMERICS_ERR_INVALID_ARGUMENT;
  R->nc = problem->numberOfContacts;
  NumericsMatrix* M = problem->M;
  if (M->size0 != 2 * R->nc || M->size1 != 2 * R->nc) return NUMERICS_ERR_INVALID_ARGUMENT;
  R->diag = (double*)calloc(4 * (size_t)R->nc, sizeof(double));
  if (!R->diag) return NUMERICS_ERR_MEMORY;
  int rc;
  if (M->storageType == NM_SPARSE)
    rc = rows_sparse(M, R);
  else if (M->storageType == NM_SPARSE_BLOCK)
    rc = rows_sbm(M->matrix1, R);
  else if (M->storageType == NM_DENSE && M->matrix0)
    rc = rows_dense(M->matrix0, R);
  else
    rc = NUMERICS_ERR_INVALID_ARGUMENT;
  if (rc) rows_free(R);
  return rc;
}

int fc2d_problem_to_device(FrictionContactProblem* problem, int device, fc2d_hip_problem** h) {
  fc2d_rows_t R;
  int rc = fc2d_rows(problem, &R);
  if (rc) return rc;
  rc = fc2d_hip_problem_create(R.nc, R.rp, R.col, R.val, R.row_mode, R.own_zero, R.diag,
                               problem->q, problem->mu, device, h);
  rows_free(&R);
  return rc;
}

void fc2d_nsgs_hip_set_default(SolverOptions* options) {
  if (!options) return;
  options->iparam[SICONOS_IPARAM_NSGS_SHUFFLE] = 0;
  options->iparam[SICONOS_NSGS_ERROR_EVALUATION_TYPE] =
      SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL;
}

void fc2d_nsgs_hip(FrictionContactProblem* problem, double* reaction, double* velocity,
                   int* info, SolverOptions* options) {
  if (!info) return;
  if (!problem || !reaction || !velocity || !options) {
    *info = NUMERICS_ERR_NULL_POINTER;
    return;
  }
  if (options->iparam[SICONOS_NSGS_FREEZING_CONTACT] > 0) {
    fprintf(stderr, "fc2d_nsgs_hip: contact freezing is not supported\n");
    *info = NUMERICS_ERR_NOT_IMPLEMENTED;
    return;
  }
  fc2d_hip_problem* h = NULL;
  int rc = fc2d_problem_to_device(problem, options->iparam[SICONOS_FRICTION_3D_NSGS_HIP_DEVICE], &h);
  if (!rc) rc = fc2d_hip_set_reaction(h, reaction);
  fc2d_hip_params p;
  fc2d_hip_params_default(&p);
  p.max_iter = options->iparam[SICONOS_IPARAM_MAX_ITER];
  p.tol = options->dparam[SICONOS_DPARAM_TOL];
  p.error_evaluation = options->iparam[SICONOS_NSGS_ERROR_EVALUATION_TYPE];
  fc2d_hip_stats s;
  memset(&s, 0, sizeof s);
  if (!rc) rc = fc2d_hip_solve(h, &p, &s);
  if (!rc) rc = fc2d_hip_get_reaction(h, reaction);
  if (!rc && s.velocity_set) rc = fc2d_hip_get_velocity(h, velocity);
  fc2d_hip_problem_destroy(h);
  if (rc) {
    fprintf(stderr, "fc2d_nsgs_hip: %s (%s)\n", numerics_error_string(rc), fc3d_hip_last_error());
    *info = rc;
    return;
  }
  options->iparam[SICONOS_IPARAM_ITER_DONE] = s.iterations;
  options->dparam[SICONOS_DPARAM_RESIDU] = s.residual;
  options->iparam[SICONOS_FRICTION_3D_NSGS_HIP_NUM_COLOURS] = s.num_colours;
  options->dparam[SICONOS_FRICTION_3D_NSGS_HIP_TIME_SOLVE] = s.solve_ms;
  *info = s.info;
}

static int fc2d_init_wrap(void* problem, SolverOptions* options) {
  (void)problem;
  fc2d_nsgs_hip_set_default(options); /* as fc2d_nsgs_init_wrap, fc2d_nsgs.c:802-805 */
  return NUMERICS_OK;
}

static int fc2d_solve_wrap(void* problem, double* reaction, double* velocity,
                           SolverOptions* options) {
  int info = NUMERICS_OK;
  fc2d_nsgs_hip((FrictionContactProblem*)problem, reaction, velocity, &info, options);
  return info;
}

FC3D_REGISTER_SOLVER(SICONOS_FRICTION_2D_NSGS_HIP, "FC2D_NSGS_HIP",
                     "Coloured non-smooth Gauss-Seidel for 2D friction contact on MI355X",
                     fc2d_init_wrap, fc2d_solve_wrap, NULL, NULL, fc2d_nsgs_hip_set_default,
                     1000, 1e-4, 0)

#define CHECK(cond, code, msg)                                                        \
  do {                                                                                \
    if (!(cond)) {                                                                    \
      fprintf(stderr, "[ERROR] %s:%d (%s) - %s\n", __FILE__, __LINE__, __func__, msg); \
      return code;                                                                    \
    }                                                                                 \
  } while (0)

int fc2d_driver(FrictionContactProblem* problem, double* reaction, double* velocity,
                SolverOptions* options) {
  CHECK(problem, NUMERICS_ERR_NULL_POINTER, "problem is NULL");
  CHECK(reaction, NUMERICS_ERR_NULL_POINTER, "reaction is NULL");
  CHECK(velocity, NUMERICS_ERR_NULL_POINTER, "velocity is NULL");
  CHECK(options, NUMERICS_ERR_NULL_POINTER, "options is NULL");
  CHECK(options->iparam && options->dparam, NUMERICS_ERR_INVALID_OPTION, "iparam/dparam NULL");
  CHECK(problem->M, NUMERICS_ERR_NULL_POINTER, "problem->M is NULL");
  CHECK(problem->q, NUMERICS_ERR_NULL_POINTER, "problem->q is NULL");
  CHECK(problem->mu, NUMERICS_ERR_NULL_POINTER, "problem->mu is NULL");
  CHECK(problem->numberOfContacts > 0, NUMERICS_ERR_INVALID_ARGUMENT,
        "Number of contacts must be positive");
  CHECK(problem->dimension == 2, NUMERICS_ERR_INVALID_ARGUMENT, "dimension must be 2");
  options->iparam[SICONOS_IPARAM_ITER_DONE] = 0;
  options->dparam[SICONOS_DPARAM_RESIDU] = 0.0;
  CHECK(options->solverId >= 400 && options->solverId < 500, NUMERICS_ERR_INVALID_SOLVER,
        "not an fc2d solver id");
  const SolverEntry* s = solver_registry_lookup(options->solverId);
  CHECK(s != NULL, NUMERICS_ERR_INVALID_SOLVER, "Solver ID not found in registry");
  CHECK(!s->is_local_solver, NUMERICS_ERR_INVALID_SOLVER, "local solver as main solver");
  CHECK(s->solve != NULL, NUMERICS_ERR_INVALID_SOLVER, "Solver has no solve function");
  if (s->init) {
    int st = s->init(problem, options);
    if (st != NUMERICS_OK) return st;
  }
  int st = s->solve(problem, reaction, velocity, options);
  if (s->free) s->free(problem, options);
  return st;
}

int fc2d_compute_error(FrictionContactProblem* problem, double* z, double* w, double tolerance,
                       double norm, double* error) {
  if (!problem || !z || !w || !error) return NUMERICS_ERR_NULL_POINTER;
  fc2d_hip_problem* h = NULL;
  int rc = fc2d_problem_to_device(problem, 0, &h);
  if (!rc) rc = fc2d_hip_compute_error(h, z, w, norm, error);
  fc2d_hip_problem_destroy(h);
  if (rc) return rc;
  return *error > tolerance ? 1 : 0;
}
