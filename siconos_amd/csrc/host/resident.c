/* resident.c — device problems kept resident across fc3d_driver and
 * fc3d_compute_error calls on the same matrix.
 *
 * The reference rebuilds nothing on the device because it has no device; this
 * path's per-call set-up (upload, colouring, sliced layout, autotuning, graph
 * capture) costs about a second on 1M contacts, so the solver keeps up to
 * RESIDENT_MAX device problems keyed on the caller's NumericsMatrix pointer
 * (NumericsMatrix.h:73-92) and reuses one when the block structure is the same:
 * the colouring, layout, tuned schedule and captured graphs stay valid, and only
 * values are refreshed.
 *
 * Values.  q and mu are plain arrays with no version, so they are re-uploaded on
 * every call (32 bytes a contact).  W's blocks are re-uploaded on every call too,
 * unless the caller sets dparam[SICONOS_FRICTION_3D_NSGS_HIP_RESIDENCY] = 1: then
 * an unchanged, non-zero NumericsDataVersion of the storage in use
 * (NumericsMatrix.c:112-135) is taken to mean unchanged values, as the
 * reference's own caches of derived data trust the caller to invalidate them
 * (NM_LU_factorized, NumericsMatrix.c:2386-2398).  A value of -1 turns
 * residency off (a fresh device problem per call, destroyed after it).
 *
 * Input storage.  NM_SPARSE_BLOCK is used as is.  NM_DENSE and NM_SPARSE are
 * cut into 3x3 blocks, block columns ascending in each block row, diagonal
 * block always present.  For NM_SPARSE this keeps the reference's arithmetic:
 * its row product sums each scalar row in ascending column order over the
 * stored entries (NumericsMatrix.c:512-573, through NM_csc_trans), which is the
 * order mvp3x3 over column-sorted blocks adds the same products in; the
 * explicit zeros of a block add +0 or -0.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "fc3d_hip.h"
#include "fc3d_numerics.h"

int fc3d_sbm_to_csr(const SparseBlockStructuredMatrix* A, int64_t** rp, int64_t** ci,
                    double** vals);

#define RESIDENT_MAX 4

typedef struct {
  const NumericsMatrix* M;  /* key: the caller's matrix */
  int storage;              /* storageType at upload */
  const void* store;        /* matrix0, matrix1 or matrix2 at upload */
  version_t version;        /* of that storage at upload */
  int nc, device, busy;
  int64_t nnzb;
  int64_t *rp, *ci;         /* block structure at upload */
  fc3d_hip_problem* h;
  unsigned long last_use;
} resident_entry;

static resident_entry g_cache[RESIDENT_MAX];
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static unsigned long g_tick = 0;
static int64_t g_hits = 0, g_creates = 0, g_value_uploads = 0;

static version_t storage_version(const NumericsMatrix* M) {
  switch (M->storageType) {
    case NM_DENSE: return M->version.number;
    case NM_SPARSE_BLOCK: return M->matrix1 ? M->matrix1->version.number : 0;
    case NM_SPARSE: {
      version_t v = 0;
      if (M->matrix2)
        for (int k = 0; k < 5; k++)
          if (M->matrix2->versions[k].number > v) v = M->matrix2->versions[k].number;
      return v;
    }
    default: return 0;
  }
}

static const void* storage_ptr(const NumericsMatrix* M) {
  switch (M->storageType) {
    case NM_DENSE: return M->matrix0;
    case NM_SPARSE_BLOCK: return M->matrix1;
    case NM_SPARSE: return M->matrix2;
    default: return NULL;
  }
}

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

/* 3x3 blocks of an NM_SPARSE matrix of nc contacts (from its CSC form). */
static int sparse_to_csr(NumericsMatrix* M, int nc, int64_t** rp, int64_t** ci, double** vals) {
  CSparseMatrix* A = NM_csc(M);
  if (!A || A->m != 3 * (int64_t)nc || A->n != 3 * (int64_t)nc) return NUMERICS_ERR_INVALID_ARGUMENT;
  const int64_t nz = A->p[A->n];
  /* keys I * nc + J of every entry's block, plus every diagonal block */
  int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nz + nc + 1));
  if (!key) return NUMERICS_ERR_MEMORY;
  int64_t nk = 0;
  for (int64_t j = 0; j < A->n; j++)
    for (int64_t p = A->p[j]; p < A->p[j + 1]; p++) key[nk++] = (A->i[p] / 3) * nc + j / 3;
  for (int64_t c = 0; c < nc; c++) key[nk++] = c * nc + c;
  qsort(key, (size_t)nk, sizeof(int64_t), cmp_i64);
  int64_t nb = 0;
  for (int64_t k = 0; k < nk; k++)
    if (k == 0 || key[k] != key[k - 1]) key[nb++] = key[k];
  *rp = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
  *ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb ? nb : 1));
  *vals = (double*)calloc(9 * (size_t)(nb ? nb : 1), sizeof(double));
  if (!*rp || !*ci || !*vals) {
    free(key);
    return NUMERICS_ERR_MEMORY;
  }
  for (int64_t b = 0; b < nb; b++) {
    (*rp)[key[b] / nc + 1]++;
    (*ci)[b] = key[b] % nc;
  }
  for (int c = 0; c < nc; c++) (*rp)[c + 1] += (*rp)[c];
  /* values: the block of entry (i, j) is found by bisection in its row */
  for (int64_t j = 0; j < A->n; j++)
    for (int64_t p = A->p[j]; p < A->p[j + 1]; p++) {
      const int64_t I = A->i[p] / 3, J = j / 3;
      int64_t lo = (*rp)[I], hi = (*rp)[I + 1] - 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if ((*ci)[mid] < J) lo = mid + 1; else hi = mid;
      }
      (*vals)[9 * lo + (A->i[p] % 3) + 3 * (j % 3)] += A->x[p];
    }
  free(key);
  return 0;
}

/* CSR over 3x3 blocks of problem->M in any supported storage. */
int fc3d_problem_csr(FrictionContactProblem* problem, int64_t** rp, int64_t** ci, double** vals) {
  NumericsMatrix* M = problem->M;
  const int nc = problem->numberOfContacts;
  if (M->storageType == NM_SPARSE_BLOCK) {
    if (!M->matrix1) return NUMERICS_ERR_INVALID_ARGUMENT;
    return fc3d_sbm_to_csr(M->matrix1, rp, ci, vals);
  }
  if (M->storageType == NM_DENSE) {
    SparseBlockStructuredMatrix* S = SBM_from_dense_3x3(nc, M->matrix0);
    if (!S) return NUMERICS_ERR_INVALID_ARGUMENT;
    int rc = fc3d_sbm_to_csr(S, rp, ci, vals);
    SBM_free_all(S);
    return rc;
  }
  if (M->storageType == NM_SPARSE) {
    if (!M->matrix2) return NUMERICS_ERR_INVALID_ARGUMENT;
    return sparse_to_csr(M, nc, rp, ci, vals);
  }
  return NUMERICS_ERR_NOT_IMPLEMENTED;
}

/* Builds a fresh device copy of problem->M, q and mu. */
int fc3d_problem_to_device(FrictionContactProblem* problem, int device, fc3d_hip_problem** h) {
  int64_t *rp = NULL, *ci = NULL;
  double* vals = NULL;
  int rc = fc3d_problem_csr(problem, &rp, &ci, &vals);
  if (!rc)
    rc = fc3d_hip_problem_create(problem->numberOfContacts, rp, ci, vals, problem->q, problem->mu,
                                 device, h);
  free(rp);
  free(ci);
  free(vals);
  return rc;
}

static void entry_clear(resident_entry* e) {
  fc3d_hip_problem_destroy(e->h);
  free(e->rp);
  free(e->ci);
  memset(e, 0, sizeof *e);
}

/* Does the SBM in use have the structure e was built from? */
static int same_sbm_structure(const resident_entry* e, const SparseBlockStructuredMatrix* S) {
  const size_t n = S->filled1 ? S->filled1 - 1 : 0;
  if ((int)n != e->nc || (int64_t)S->nbblocks != e->nnzb) return 0;
  for (size_t i = 0; i <= n; i++)
    if ((int64_t)S->index1_data[i] != e->rp[i]) return 0;
  for (size_t k = 0; k < S->nbblocks; k++)
    if ((int64_t)S->index2_data[k] != e->ci[k]) return 0;
  return 1;
}

static int same_structure(const resident_entry* e, int nc, const int64_t* rp, const int64_t* ci) {
  if (nc != e->nc || rp[nc] != e->nnzb) return 0;
  return memcmp(rp, e->rp, sizeof(int64_t) * ((size_t)nc + 1)) == 0 &&
         memcmp(ci, e->ci, sizeof(int64_t) * (size_t)e->nnzb) == 0;
}

/* A device problem holding problem->M, q and mu, resident or fresh.  mode:
 * -1 fresh (token -1, destroyed by the release); 0 resident, W's values
 * refreshed; 1 resident, W's values kept while the storage's version is
 * unchanged and non-zero.  The caller owns the problem until
 * fc3d_resident_release(token). */
int fc3d_resident_acquire(FrictionContactProblem* problem, int device, int mode,
                          fc3d_hip_problem** out, int* token) {
  *out = NULL;
  *token = -1;
  if (mode < 0) return fc3d_problem_to_device(problem, device, out);
  NumericsMatrix* M = problem->M;
  const int nc = problem->numberOfContacts;
  const version_t ver = storage_version(M);
  const void* store = storage_ptr(M);
  pthread_mutex_lock(&g_lock);
  int slot = -1;
  for (int k = 0; k < RESIDENT_MAX; k++)
    if (g_cache[k].h && g_cache[k].M == M && g_cache[k].device == device &&
        g_cache[k].storage == (int)M->storageType && !g_cache[k].busy) {
      slot = k;
      break;
    }
  if (slot >= 0) g_cache[slot].busy = 1;
  pthread_mutex_unlock(&g_lock);

  int64_t *rp = NULL, *ci = NULL;
  double* vals = NULL;
  int rc = 0;
  if (slot >= 0) {
    resident_entry* e = &g_cache[slot];
    int same;
    if (M->storageType == NM_SPARSE_BLOCK && M->matrix1) {
      same = same_sbm_structure(e, M->matrix1);
    } else {
      rc = fc3d_problem_csr(problem, &rp, &ci, &vals);
      same = !rc && same_structure(e, nc, rp, ci);
    }
    if (same) {
      const int keep = mode == 1 && ver != 0 && ver == e->version && store == e->store;
      if (!keep && !vals) rc = fc3d_problem_csr(problem, &rp, &ci, &vals);
      if (!rc) rc = fc3d_hip_problem_update(e->h, keep ? NULL : vals, problem->q, problem->mu);
      free(rp);
      free(ci);
      free(vals);
      pthread_mutex_lock(&g_lock);
      if (!rc) {
        e->version = ver;
        e->store = store;
        e->last_use = ++g_tick;
        g_hits++;
        if (!keep) g_value_uploads++;
        *out = e->h;
        *token = slot;
      } else {
        e->busy = 0;
      }
      pthread_mutex_unlock(&g_lock);
      return rc;
    }
    /* another structure under the same key: replace the entry */
    pthread_mutex_lock(&g_lock);
    entry_clear(e);
    pthread_mutex_unlock(&g_lock);
    rc = 0;
  }
  if (!vals) rc = fc3d_problem_csr(problem, &rp, &ci, &vals);
  fc3d_hip_problem* h = NULL;
  if (!rc) rc = fc3d_hip_problem_create(nc, rp, ci, vals, problem->q, problem->mu, device, &h);
  free(vals);
  if (rc) {
    free(rp);
    free(ci);
    return rc;
  }
  pthread_mutex_lock(&g_lock);
  g_creates++;
  int victim = -1;
  for (int k = 0; k < RESIDENT_MAX; k++)
    if (!g_cache[k].h) { victim = k; break; }
  if (victim < 0)
    for (int k = 0; k < RESIDENT_MAX; k++)
      if (!g_cache[k].busy && (victim < 0 || g_cache[k].last_use < g_cache[victim].last_use))
        victim = k;
  if (victim >= 0) {
    resident_entry* e = &g_cache[victim];
    if (e->h) entry_clear(e);
    e->M = M;
    e->storage = (int)M->storageType;
    e->store = store;
    e->version = ver;
    e->nc = nc;
    e->device = device;
    e->nnzb = rp[nc];
    e->rp = rp;
    e->ci = ci;
    e->h = h;
    e->busy = 1;
    e->last_use = ++g_tick;
    *token = victim;
  } else {
    /* every entry is in use by another thread: this call's problem is its own */
    free(rp);
    free(ci);
  }
  pthread_mutex_unlock(&g_lock);
  *out = h;
  return 0;
}

void fc3d_resident_release(fc3d_hip_problem* h, int token) {
  if (token < 0) {
    fc3d_hip_problem_destroy(h);
    return;
  }
  pthread_mutex_lock(&g_lock);
  g_cache[token].busy = 0;
  pthread_mutex_unlock(&g_lock);
}

/* Drops every resident problem of M (NULL: all idle ones). */
void fc3d_hip_resident_clear(const NumericsMatrix* M) {
  pthread_mutex_lock(&g_lock);
  for (int k = 0; k < RESIDENT_MAX; k++)
    if (g_cache[k].h && !g_cache[k].busy && (!M || g_cache[k].M == M)) entry_clear(&g_cache[k]);
  pthread_mutex_unlock(&g_lock);
}

void fc3d_hip_resident_stats(int64_t* hits, int64_t* creates, int64_t* value_uploads) {
  pthread_mutex_lock(&g_lock);
  if (hits) *hits = g_hits;
  if (creates) *creates = g_creates;
  if (value_uploads) *value_uploads = g_value_uploads;
  pthread_mutex_unlock(&g_lock);
}
