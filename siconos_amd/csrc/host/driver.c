/* driver.c — fc3d_driver and fc3d_compute_error (behaviour of
 * numerics/src/FrictionContact/fc3d_driver.c:51-190 and
 * fc3d_compute_error.c:70-122; the residual itself is computed on the GPU). */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>

#include "fc3d_hip.h"
#include "fc3d_numerics.h"

int fc3d_resident_acquire(FrictionContactProblem* problem, int device, int mode,
                          fc3d_hip_problem** out, int* token);
void fc3d_resident_release(fc3d_hip_problem* h, int token);

int fc3d_checkTrivialCase(FrictionContactProblem* problem, double* velocity, double* reaction,
                          SolverOptions* options) {
  (void)options;
  int nc = problem->numberOfContacts;
  const double* q = problem->q;
  for (int i = 0; i < nc; i++)
    if (q[3 * i] < -DBL_EPSILON) return NUMERICS_ERR_INFEASIBLE;
  for (int i = 0; i < 3 * nc; i++) {
    velocity[i] = q[i];
    reaction[i] = 0.;
  }
  return NUMERICS_OK;
}

#define CHECK(cond, code, msg)                                                        \
  do {                                                                                \
    if (!(cond)) {                                                                    \
      fprintf(stderr, "[ERROR] %s:%d (%s) - %s\n", __FILE__, __LINE__, __func__, msg); \
      return code;                                                                    \
    }                                                                                 \
  } while (0)

int fc3d_driver(FrictionContactProblem* problem, double* reaction, double* velocity,
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
  CHECK(problem->dimension == 3, NUMERICS_ERR_INVALID_ARGUMENT, "dimension must be 3");
  options->iparam[SICONOS_IPARAM_ITER_DONE] = 0;
  options->dparam[SICONOS_DPARAM_RESIDU] = 0.0;
  if (fc3d_checkTrivialCase(problem, velocity, reaction, options) == NUMERICS_OK) return NUMERICS_OK;
  const SolverEntry* s = solver_registry_lookup(options->solverId);
  CHECK(s != NULL, NUMERICS_ERR_INVALID_SOLVER, "Solver ID not found in registry");
  CHECK(s->solve != NULL, NUMERICS_ERR_INVALID_SOLVER, "Solver has no solve function");
  if (s->init) {
    int st = s->init(problem, options);
    if (st != NUMERICS_OK) {
      fprintf(stderr, "[ERROR] fc3d_driver: solver initialization failed: %s\n",
              numerics_error_string(st));
      return st;
    }
  }
  int info = s->solve(problem, reaction, velocity, options);
  if (s->free) s->free(problem, options);
  if (info != NUMERICS_OK)
    fprintf(stderr, "[WARN] fc3d_driver: solver returned status %d (%s)\n", info,
            numerics_error_string(info));
  return info;
}

SolverOptions* fc3d_solver_options_create(solver_id_t id) {
  const SolverEntry* s = solver_registry_lookup(id);
  if (!s) {
    fprintf(stderr, "[ERROR] fc3d_solver_options_create: solver ID %d not registered\n", id);
    return NULL;
  }
  if (s->is_local_solver) {
    fprintf(stderr, "[ERROR] fc3d_solver_options_create: '%s' is a local solver\n", s->name);
    return NULL;
  }
  return solver_options_create(id);
}

int fc3d_compute_error(FrictionContactProblem* problem, double* z, double* w, double tolerance,
                       SolverOptions* options, double norm, double* error) {
  if (!problem || !z || !w || !error || !problem->M || !problem->q || !problem->mu)
    return NUMERICS_ERR_NULL_POINTER;
  /* the device problem a solve on the same matrix left resident is reused;
   * the residency option is read from FC3D_NSGS_HIP options only */
  int mode = 0, device = 0;
  if (options && options->solverId == SICONOS_FRICTION_3D_NSGS_HIP && options->dparam &&
      options->iparam) {
    mode = (int)options->dparam[SICONOS_FRICTION_3D_NSGS_HIP_RESIDENCY];
    device = options->iparam[SICONOS_FRICTION_3D_NSGS_HIP_DEVICE];
  }
  fc3d_hip_problem* h = NULL;
  int token = -1;
  int rc = fc3d_resident_acquire(problem, device, mode, &h, &token);
  if (!rc) rc = fc3d_hip_set_reaction(h, z);
  if (!rc) rc = fc3d_hip_compute_error(h, norm, error);
  if (!rc) rc = fc3d_hip_get_velocity(h, w);
  if (h) fc3d_resident_release(h, token);
  if (rc) return rc;
  return *error > tolerance ? 1 : 0;
}
