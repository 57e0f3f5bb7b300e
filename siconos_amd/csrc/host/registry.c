/* registry.c — constructor-populated solver table and error strings.
 * Behaviour of numerics/src/utils/solver_registry.c:28-95 and
 * numerics/src/utils/numerics_errors.h:70-116. */
#include <stdio.h>
#include <string.h>

#include "fc3d_numerics.h"

static const SolverEntry* g_registry[SOLVER_REGISTRY_MAX];
static size_t g_count = 0;

int solver_registry_register(const SolverEntry* entry) {
  if (!entry) return NUMERICS_ERR_NULL_POINTER;
  if (g_count >= SOLVER_REGISTRY_MAX) return -1;
  for (size_t i = 0; i < g_count; i++) {
    if (g_registry[i]->id == entry->id) {
      fprintf(stderr, "solver_registry: duplicate solver ID %d\n", entry->id);
      return -1;
    }
  }
  g_registry[g_count++] = entry;
  return 0;
}

const SolverEntry* solver_registry_lookup(solver_id_t id) {
  for (size_t i = 0; i < g_count; i++)
    if (g_registry[i]->id == id) return g_registry[i];
  return NULL;
}

const SolverEntry* solver_registry_lookup_by_name(const char* name) {
  if (!name) return NULL;
  for (size_t i = 0; i < g_count; i++)
    if (g_registry[i]->name && strcmp(g_registry[i]->name, name) == 0) return g_registry[i];
  return NULL;
}

size_t solver_registry_count(void) { return g_count; }

int solver_registry_exists(solver_id_t id) { return solver_registry_lookup(id) != NULL; }

const char* numerics_error_string(int code) {
  switch (code) {
    case NUMERICS_OK: return "Success";
    case NUMERICS_ERR_NULL_POINTER: return "Null pointer";
    case NUMERICS_ERR_INVALID_ARGUMENT: return "Invalid argument";
    case NUMERICS_ERR_INVALID_OPTION: return "Invalid option";
    case NUMERICS_ERR_INVALID_SOLVER: return "Invalid solver";
    case NUMERICS_ERR_NOT_IMPLEMENTED: return "Not implemented";
    case NUMERICS_ERR_MEMORY: return "Memory allocation failed";
    case NUMERICS_ERR_FILE_IO: return "File I/O error";
    case NUMERICS_ERR_MAX_ITER: return "Maximum iterations reached";
    case NUMERICS_ERR_DIVERGENCE: return "Divergence detected";
    case NUMERICS_ERR_STAGNATION: return "Stagnation detected";
    case NUMERICS_ERR_LINEAR_SOLVER: return "Linear solver failed";
    case NUMERICS_ERR_LOCAL_SOLVER: return "Local solver failed";
    case NUMERICS_ERR_PROJECTION: return "Projection failed";
    case NUMERICS_ERR_INFEASIBLE: return "Problem infeasible";
    case NUMERICS_ERR_UNBOUNDED: return "Problem unbounded";
    case NUMERICS_ERR_ILL_CONDITIONED: return "Ill-conditioned problem";
    case NUMERICS_ERR_SINGULAR_MATRIX: return "Singular matrix";
    case NUMERICS_ERR_DEVICE: return "GPU runtime or device failure";
    default: return "Unknown error";
  }
}
