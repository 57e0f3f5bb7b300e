/* fc2d_hip.h — device layer of the fc2d (2-D friction contact) NSGS on HIP.
 *
 * Plain C ABI: pointers and sizes only.  It is what fc2d_driver's
 * SICONOS_FRICTION_2D_NSGS_HIP solver (fc2d_numerics.h) calls; a cgo, JNI or
 * ctypes binding can call it directly.
 *
 * The reference's fc2d_nsgs (fc2d_nsgs.c:297-538) sweeps contacts in natural
 * order on the CPU; its CUDA counterpart (parallel/gpu/
 * fc2d_nsgs_graph_permut_cuda.cu) sweeps a colouring of the contact graph.
 * Here one workgroup runs the whole solve: colour passes separated by
 * workgroup barriers, the light error, the LIGHT_WITH_FULL_FINAL logic and
 * fc2d_compute_error all on the device.  Sweeps follow the greedy colouring
 * in index order, then index order inside a colour (fc2d_hip_problem_order);
 * in that order the result equals the reference fc2d_nsgs bit for bit for
 * SBM and CSparse input.  For dense input the reference's row product goes
 * through a vendor cblas_ddot whose summation order is not fixed; here the
 * terms are summed in ascending column order.
 */
#ifndef FC2D_HIP_H
#define FC2D_HIP_H

#include <stdint.h>

#include "fc3d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fc2d_hip_problem fc2d_hip_problem;

/* How a row's terms enter the local problem of the sweep (fc2d_nsgs.c:97-113
 * through NM_row_prod_no_diag2, NumericsMatrix.c:674-749):
 *   FC2D_ROWS_ACCUMULATE: q_loc = q, then q_loc += w x term by term, in the
 *     given order (NM_SPARSE_BLOCK through mvp2x2, NM_SPARSE through the
 *     rows of the transposed CSC);
 *   FC2D_ROWS_DOT: q_loc = q + (sum of the terms from zero) (NM_DENSE
 *     through cblas_ddot). */
enum { FC2D_ROWS_ACCUMULATE = 0, FC2D_ROWS_DOT = 1 };

/* Creates a resident problem from its scalar rows.  Row i (0 <= i < 2 nc)
 * holds entries row_ptr[i] .. row_ptr[i+1]-1: column cols[e] (0 <= cols[e]
 * < 2 nc) and value vals[e], in the order the reference sums them.  Entries
 * whose column belongs to the row's own contact enter the sweep's local
 * problem multiplied by zero when own_zero = 1 (the dense and CSparse forms
 * zero the contact's own reaction), and are skipped when own_zero = 0 (the
 * SBM form skips the diagonal block).  The residual u = q + W r uses every
 * entry, in the given order.  diag holds the 2x2 diagonal block of each
 * contact, column-major (NM_extract_diag_block2, NumericsMatrix.c:1799-1826).
 * A contact conflicts with every contact its rows reference, or that
 * references it.  0 or a NUMERICS_ERR_* code. */
FC3D_API int fc2d_hip_problem_create(int nc, const int64_t* row_ptr, const int64_t* cols,
                                     const double* vals, int row_mode, int own_zero,
                                     const double* diag, const double* q, const double* mu,
                                     int device, fc2d_hip_problem** out);
FC3D_API void fc2d_hip_problem_destroy(fc2d_hip_problem* h);
FC3D_API int fc2d_hip_problem_num_colours(const fc2d_hip_problem* h);
/* the sweep order: order[k] is the k-th contact swept */
FC3D_API int fc2d_hip_problem_order(const fc2d_hip_problem* h, int* order);
/* reaction / velocity in the caller's contact order (2 nc doubles) */
FC3D_API int fc2d_hip_set_reaction(fc2d_hip_problem* h, const double* reaction);
FC3D_API int fc2d_hip_get_reaction(fc2d_hip_problem* h, double* reaction);
FC3D_API int fc2d_hip_get_velocity(fc2d_hip_problem* h, double* velocity);

/* How the sweeps are run (both give the same iterates bit for bit):
 *   FC2D_SCHEDULE_ONE_WORKGROUP: one launch of one 1024-thread workgroup
 *     runs the whole solve, colours separated by workgroup barriers;
 *   FC2D_SCHEDULE_PER_COLOUR: one thread per contact across the GPU, one
 *     launch per colour, sweeps replayed from a captured graph;
 *   FC2D_SCHEDULE_AUTO: per colour when some colour holds more than 2048
 *     contacts, one workgroup otherwise. */
enum { FC2D_SCHEDULE_AUTO = 0, FC2D_SCHEDULE_ONE_WORKGROUP = 1, FC2D_SCHEDULE_PER_COLOUR = 2 };

typedef struct fc2d_hip_params {
  int max_iter;         /* iparam[SICONOS_IPARAM_MAX_ITER]                    */
  double tol;           /* dparam[SICONOS_DPARAM_TOL]                         */
  int error_evaluation; /* iparam[13]: LIGHT (1) or LIGHT_WITH_FULL_FINAL (2); *
                         * fc2d_nsgs evaluates no other mode (it then sweeps  *
                         * max_iter times and reports info 1)                */
  int schedule;         /* FC2D_SCHEDULE_*                                    */
} fc2d_hip_params;

typedef struct fc2d_hip_stats {
  int info;            /* fc2d_nsgs's: 0 converged, 1 not (or a singular     *
                        * diagonal block, fc2d_nsgs.c:318-334)               */
  int iterations;
  double residual;     /* LIGHT: the last light error; LIGHT_WITH_FULL_FINAL:
                          fc2d_compute_error of the result                   */
  int num_colours;
  int velocity_set;    /* 1 if u = q + W r was computed (full error)         */
  double solve_ms;     /* device time of the solve                           */
  int schedule;        /* the FC2D_SCHEDULE_* used                            */
} fc2d_hip_stats;

/* fc2d_nsgs_set_default (fc2d_nsgs.c:785-789) and the registry defaults:
 * max_iter 1000, tol 1e-4, LIGHT_WITH_FULL_FINAL; schedule automatic */
FC3D_API void fc2d_hip_params_default(fc2d_hip_params* p);
/* One solve from the resident reaction (fc2d_nsgs.c:297-538, without contact
 * freezing or shuffling).  0 or a NUMERICS_ERR_* code; the reference's info
 * is in stats->info. */
FC3D_API int fc2d_hip_solve(fc2d_hip_problem* h, const fc2d_hip_params* p, fc2d_hip_stats* s);
/* fc2d_compute_error (fc2d_compute_error.c:104-148) of `reaction` (caller
 * order, or NULL for the resident reaction); velocity (may be NULL) receives
 * u = q + W r. */
FC3D_API int fc2d_hip_compute_error(fc2d_hip_problem* h, const double* reaction,
                                    double* velocity, double norm, double* error);

#ifdef __cplusplus
}
#endif

#endif
