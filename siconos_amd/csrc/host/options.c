/* options.c — SolverOptions life cycle, behaviour of
 * numerics/src/core/SolverOptions.c:44-94 (initialise), 176-268 (delete, copy,
 * update_internal), 270-379 (name lookup, registry-based create). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "fc3d_numerics.h"

static SolverOptions* options_initialize(int solver_id, int iter_max, double tol) {
  SolverOptions* o = (SolverOptions*)calloc(1, sizeof(SolverOptions));
  if (!o) return NULL;
  o->solverId = solver_id;
  o->iSize = OPTIONS_PARAM_SIZE;
  o->dSize = OPTIONS_PARAM_SIZE;
  o->iparam = (int*)calloc(o->iSize, sizeof(int));
  o->dparam = (double*)calloc(o->dSize, sizeof(double));
  o->iparam[SICONOS_IPARAM_MAX_ITER] = iter_max;
  o->dparam[SICONOS_DPARAM_TOL] = tol;
  o->dparam[SICONOS_DPARAM_RESIDU] = INFINITY;
  o->filterOn = true;
  o->isSet = true;
  return o;
}

void solver_options_delete(SolverOptions* op) {
  if (!op) return;
  free(op->solverParameters);
  free(op->solverData);
  free(op->callback);
  if (op->internalSolvers) {
    for (size_t i = 0; i < op->numberOfInternalSolvers; i++)
      solver_options_delete(op->internalSolvers[i]);
    free(op->internalSolvers);
  }
  free(op->iWork);
  free(op->dWork);
  free(op->iparam);
  free(op->dparam);
  free(op);
}

SolverOptions* solver_options_create(int solver_id) {
  const SolverEntry* s = solver_registry_lookup(solver_id);
  if (!s) {
    fprintf(stderr, "solver_options_create: solver '%i' not found in registry\n", solver_id);
    return NULL;
  }
  SolverOptions* o = options_initialize(solver_id, s->default_max_iter, s->default_tol);
  if (o && s->set_default) s->set_default(o);
  return o;
}

SolverOptions* solver_options_create_by_name(const char* name) {
  const SolverEntry* s = solver_registry_lookup_by_name(name);
  if (!s) {
    fprintf(stderr, "solver_options_create_by_name: solver '%s' not found in registry\n",
            name ? name : "(null)");
    return NULL;
  }
  return solver_options_create(s->id);
}

SolverOptions* solver_options_copy(SolverOptions* src) {
  if (!src) return NULL;
  SolverOptions* o = solver_options_create(src->solverId);
  if (!o) return NULL;
  for (size_t i = 0; i < OPTIONS_PARAM_SIZE; i++) {
    o->iparam[i] = src->iparam[i];
    o->dparam[i] = src->dparam[i];
  }
  if (src->dWork) {
    o->dWork = (double*)calloc(src->dWorkSize, sizeof(double));
    o->dWorkSize = src->dWorkSize;
    for (size_t i = 0; i < o->dWorkSize; i++) o->dWork[i] = src->dWork[i];
  }
  if (src->iWork) {
    o->iWork = (int*)calloc(src->iWorkSize, sizeof(int));
    o->iWorkSize = src->iWorkSize;
    for (size_t i = 0; i < o->iWorkSize; i++) o->iWork[i] = src->iWork[i];
  }
  for (size_t i = 0; i < o->numberOfInternalSolvers && i < src->numberOfInternalSolvers; i++) {
    solver_options_delete(o->internalSolvers[i]);
    o->internalSolvers[i] = solver_options_copy(src->internalSolvers[i]);
  }
  /* pointer links shared with the source, as SolverOptions.c:249-255 keeps
   * them (the copy does not own them) */
  if (src->callback) o->callback = src->callback;
  if (src->solverData) o->solverData = src->solverData;
  if (src->solverParameters) o->solverParameters = src->solverParameters;
  return o;
}

void solver_options_update_internal(SolverOptions* parent, size_t n, int solver_id) {
  if (!parent || n >= parent->numberOfInternalSolvers) return;
  solver_options_delete(parent->internalSolvers[n]);
  parent->internalSolvers[n] = solver_options_create(solver_id);
}

const char* solver_options_id_to_name(int id) {
  const SolverEntry* s = solver_registry_lookup(id);
  return s ? s->name : "NONAME";
}

int solver_options_name_to_id(const char* name) {
  const SolverEntry* s = solver_registry_lookup_by_name(name);
  return s ? s->id : 0;
}

SolverOptions* solver_options_get_internal_solver(SolverOptions* options, size_t n) {
  if (!options || n + 1 > options->numberOfInternalSolvers) return NULL;
  return options->internalSolvers[n];
}
