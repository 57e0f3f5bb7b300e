/* fc3d_oracle.h — CPU oracle for the fc3d NSGS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py, as the reference results are compared against.  The product
 * library never links, loads or calls it.  Each function cites the reference
 * file:line it follows
 * (paths relative to the Siconos tree).
 *
 * Pinning (what checks each part against something other than itself) is
 * listed in the header comment of fc3d_oracle.c.
 */
#ifndef FC3D_ORACLE_H
#define FC3D_ORACLE_H

#include "fc3d_numerics.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_API __attribute__((visibility("default")))

/* Planted mistakes, for tests that show a pin catches them (0 = none). */
enum oracle_mutation {
  ORACLE_MUT_NONE = 0,
  ORACLE_MUT_AC_JACOBIAN_SIGN = 1,  /* wrong sign of A[1,0] in the AC Jacobian   */
  ORACLE_MUT_MVP_TRANSPOSED = 2,    /* 3x3 block applied transposed in the gemv  */
  ORACLE_MUT_ERROR_DROP_MU = 3,     /* drops mu*|u_t| from the residual          */
  ORACLE_MUT_CONE_SIGN = 4,         /* boundary case of the cone projection off  */
  ORACLE_MUT_COLOUR_OFF_BY_ONE = 5, /* colouring ignores the first neighbour     */
  ORACLE_MUT_GEPP_PIVOT = 6,        /* 3x3 solve without row exchange            */
  ORACLE_MUT_PROJ_DROP_MU = 7,      /* OC_PROJ step without the mu*|u_t| term    */
  ORACLE_MUT_PROJ_RHO = 8,          /* OC_PROJ step with rho = 1/W11, not 1/W00  */
  ORACLE_MUT_EIG_PHASE = 9,         /* eig_3x3 with the angle's sign flipped     */
  ORACLE_MUT_FREEZE_NO_STALE = 10   /* a frozen contact adds no light-error term */
};
ORACLE_API void oracle_set_mutation(int m);
/* Work counters: [0] Newton calls, [1] Newton iterations, [2] PLI calls,
 * [3] PLI iterations, [4] line-search trials, [5] non-trivial local solves. */
ORACLE_API void oracle_counters(long* out, int reset);

/* op3x3.h:342-363, 449-465, 868-981 */
ORACLE_API void oracle_mvp3x3(const double* a, const double* v, double* r);
ORACLE_API void oracle_mm3x3(const double* a, const double* b, double* c);
ORACLE_API int oracle_solve_3x3_gepp(const double* a, double* b);
/* op3x3.h:817-858 */
ORACLE_API int oracle_orthoBaseFromVector(double* n, double* t1, double* t2);

/* projectionOnCone.c:32-73 */
ORACLE_API unsigned oracle_projectionOnCone(double* r, double mu);
ORACLE_API unsigned oracle_projectionOnDualCone(double* u, double mu);

/* fc3d_AlartCurnier_functions.c:506-646 (STD), 649-753 (JeanMoreau) */
ORACLE_API void oracle_computeAlartCurnierSTD(const double* R, const double* u, double mu,
                                              const double* rho, double* F, double* A,
                                              double* B);
ORACLE_API void oracle_computeAlartCurnierJeanMoreau(const double* R, const double* u, double mu,
                                                     const double* rho, double* F, double* A,
                                                     double* B);
/* Independent formulation: F = r - P(r - rho*u) with P the normal/disk
 * projections of the Alart-Curnier map, differentiated by forward-mode dual
 * numbers (A = dF/du, B = dF/dr).  Not a restatement of any reference
 * function; it cross-checks computeAlartCurnierSTD the way
 * test/fc3d_AlartCurnierFunctions_test.c:122-151 checks it against the
 * generated implementation. */
ORACLE_API void oracle_alart_curnier_dual(const double* R, const double* u, double mu,
                                          const double* rho, double* F, double* A, double* B);

/* fc3d_AlartCurnier_functions.c:755-812 */
ORACLE_API void oracle_rho_split_spectral_norm(const double* W, double* rho);
ORACLE_API void oracle_rho_split_spectral_norm_cond(const double* W, double* rho);
ORACLE_API void oracle_eig_3x3(const double* a, double* eig);
ORACLE_API void oracle_rho_spectral_norm(const double* W, double* rho);

/* fc3d_compute_error.c:42-56 */
ORACLE_API void oracle_unitary_compute_and_add_error(const double* r, const double* u, double mu,
                                                     double* error);

/* SparseBlockMatrix.c:360-419, 1431-1490 */
ORACLE_API void oracle_SBM_gemv_3x3(const SparseBlockStructuredMatrix* A, const double* x,
                                    double* y);
ORACLE_API void oracle_SBM_row_prod_no_diag_3x3(const SparseBlockStructuredMatrix* A, size_t row,
                                                const double* x, double* y);

/* fc3d_compute_error.c:70-122 */
ORACLE_API int oracle_fc3d_compute_error(FrictionContactProblem* problem, const double* z,
                                         double* w, double tolerance, double norm,
                                         double* error);

/* Options with the reference defaults: fc3d_nsgs.c:919-950 for the NSGS
 * level and fc3d_onecontact_nonsmooth_Newton_solvers.c:1017-1056 /
 * fc3d_projection.c:917-1002 for the local level. */
ORACLE_API SolverOptions* oracle_nsgs_options_new(void);
ORACLE_API SolverOptions* oracle_local_options_new(int local_solver_id);
ORACLE_API void oracle_set_local_solver(SolverOptions* nsgs_options, int local_solver_id);
ORACLE_API void oracle_options_free(SolverOptions* options);

/* One local problem (W column-major, q, mu) solved in place from r, with the
 * per-contact state the reference keeps in local_opts->dWork
 * (onecontact...c:62-187, fc3d_projection.c:329-341).  Returns the local
 * solver's info.  `state` is 4 doubles: [PLI rho, AC rho x3]. */
ORACLE_API int oracle_local_solve(const double* W, const double* q, double mu, double* r,
                                  SolverOptions* local_options, double* state);
/* Initial per-contact state for a diagonal block W (ρ from the rho strategy,
 * PLI ρ = 1). */
ORACLE_API int oracle_local_state_init(const double* W, SolverOptions* local_options,
                                       double* state);

/* fc3d_nsgs.c:560-917.  `order` (may be NULL) gives the contact processed at
 * each position of a sweep; NULL is the reference's natural order. */
ORACLE_API void oracle_fc3d_nsgs(FrictionContactProblem* problem, double* reaction,
                                 double* velocity, int* info, SolverOptions* options,
                                 const int* order);
/* fc3d_driver.c:51-145 restricted to the NSGS solver. */
ORACLE_API int oracle_fc3d_driver(FrictionContactProblem* problem, double* reaction,
                                  double* velocity, SolverOptions* options, const int* order);

/* Reference colouring: sequential greedy in index order (the definition the
 * GPU colouring must reproduce bit for bit). */
ORACLE_API int oracle_greedy_colouring(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                                       int* colour);

/* Problem helpers for tests (SBM built from CSR; freed with
 * oracle_problem_free). */
ORACLE_API FrictionContactProblem* oracle_problem_from_csr(int nc, const int64_t* row_ptr,
                                                          const int64_t* col_idx,
                                                          const double* blocks,
                                                          const double* q, const double* mu);
ORACLE_API void oracle_problem_free(FrictionContactProblem* p);

/* Batched CPU baseline: solves `nprob` independent problems given as
 * concatenated CSR pieces, one problem per thread over `nthreads` threads.
 * Returns the number of problems whose final residual is <= tol. */
ORACLE_API int oracle_solve_many(int nprob, const int* nc, const int64_t* row_ptr_off,
                                 const int64_t* block_off, const int64_t* row_ptr,
                                 const int64_t* col_idx, const double* blocks, const double* q,
                                 const double* mu, double* reaction, double tol, int max_iter,
                                 int nthreads, int* iters, double* residuals);

/* ---- fc2d (fc2d_oracle.c) ---------------------------------------------- */
/* Reads a dimension-2 .dat problem (dense, SBM of 2x2 blocks or CSparse);
 * NULL on failure.  The handle is released with oracle_fc2d_free. */
ORACLE_API void* oracle_fc2d_read(const char* path);
ORACLE_API void oracle_fc2d_free(void* h);
/* out[0] contacts, out[1] storage type, out[2] sparse origin */
ORACLE_API int oracle_fc2d_info(void* h, int* out);
/* dense column-major copy of W (2nc x 2nc), q and mu */
ORACLE_API int oracle_fc2d_dense(void* h, double* W, double* q, double* mu);
/* fc2d_nsgs.c:97-113: q_loc of contact c at reaction x */
ORACLE_API int oracle_fc2d_local_q(void* h, int c, double* x, double* y);
/* fc2d_nsgs.c:247-284 */
ORACLE_API void oracle_fc2d_local_solve(const double* W, double D, const double* q, double mu,
                                        double* P);
/* fc2d_compute_error.c:104-148 with norm = |q| */
ORACLE_API int oracle_fc2d_compute_error(void* h, const double* z, double* w, double* error);
/* fc2d_nsgs.c:297-538; mode is a SICONOS_NSGS_ERROR_EVALUATION_* value;
 * returns info, out = {iterations, residual} */
ORACLE_API int oracle_fc2d_nsgs(void* h, double* reaction, double* velocity, double tol,
                                int itermax, int mode, const int* order, double* out);
/* 1: slide directions swapped in the local solve (a planted mistake) */
ORACLE_API void oracle_fc2d_set_mutation(int m);

#ifdef __cplusplus
}
#endif
#endif
