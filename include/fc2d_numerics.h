/* fc2d_numerics.h — drop-in C ABI of the 2-D friction-contact NSGS.
 *
 * The types (FrictionContactProblem with dimension 2, NumericsMatrix in
 * NM_DENSE, NM_SPARSE_BLOCK with 2x2 blocks, or NM_SPARSE; SolverOptions;
 * the registry) are those of fc3d_numerics.h.  Each declaration cites the
 * reference interface it replaces.
 */
#ifndef FC2D_NUMERICS_H
#define FC2D_NUMERICS_H

#include "fc3d_numerics.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FrictionContact_options.h:60-73 lists the reference fc2d ids (400-411; 409
 * and 410 are its CUDA sweeps).  This library registers one more: */
enum { SICONOS_FRICTION_2D_NSGS_HIP = 440 };

/* fc2d_driver.c:38-150: validation, registry lookup, init / solve / free.
 * Only fc2d solvers (ids 400-499) are accepted; this library registers
 * SICONOS_FRICTION_2D_NSGS_HIP. */
FC3D_API int fc2d_driver(FrictionContactProblem* problem, double* reaction, double* velocity,
                         SolverOptions* options);

/* fc2d_compute_error.c:104-148 — computed on the GPU; returns 1 if
 * error > tolerance, 0 otherwise, negative on failure. */
FC3D_API int fc2d_compute_error(FrictionContactProblem* problem, double* z, double* w,
                                double tolerance, double norm, double* error);

/* The FC2D_NSGS_HIP solver (fc2d_nsgs.c:297-538 in colour order, on the GPU):
 * the caller's reaction is the start; options keep fc2d_nsgs's meaning
 * (iparam[0] max iterations, dparam[0] tolerance, iparam[13] error
 * evaluation); contact freezing (iparam[15] > 0) is refused with
 * NUMERICS_ERR_NOT_IMPLEMENTED (its criteria depend on the natural order),
 * and shuffling is not used, as in fc2d_nsgs.  Outputs iparam[1]
 * iterations, iparam[25] colours, dparam[1] residual, dparam[24] device time
 * in ms. */
FC3D_API void fc2d_nsgs_hip(FrictionContactProblem* problem, double* reaction, double* velocity,
                            int* info, SolverOptions* options);
/* fc2d_nsgs.c:785-789 */
FC3D_API void fc2d_nsgs_hip_set_default(SolverOptions* options);

/* Uploads a dimension-2 problem as a resident fc2d_hip_problem (fc2d_hip.h):
 * its rows in the reference's summation order for the matrix's storage, the
 * diagonal blocks, q and mu.  No reference counterpart (the reference keeps
 * the problem on the host); callers that solve repeatedly on the same W keep
 * the handle.  0 or a NUMERICS_ERR_* code. */
struct fc2d_hip_problem;
FC3D_API int fc2d_problem_to_device(FrictionContactProblem* problem, int device,
                                    struct fc2d_hip_problem** out);

#ifdef __cplusplus
}
#endif

#endif
