/* sparse.c — compressed (NM_SPARSE) matrices on the host: the .dat reader
 * and writer, release, and the CSC form.
 *
 * Format of NumericsSparseMatrix.c:549-583 (NSM_write_in_file /
 * NSM_new_from_file) and CSparseMatrix.c:651-790, 883-940: the origin
 * (NSM_TRIPLET 1, NSM_CSC 2, NSM_CSR 3), a banner line, a size line
 * ("m-by-n, nzmax: .. nnz: .., 1-norm: ..", prefixed by "triplet:" or
 * "csr:"), then per column (row for CSR) a "col j : locations a to b" line
 * and one "i : value" line per entry, or for a triplet one "i j : value"
 * line per entry.  The reference prints values with %g; this writer uses
 * %.17g so that a write/read round trip is exact, and the reader takes both. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fc3d_numerics.h"

static CSparseMatrix* cs_new(int64_t m, int64_t n, int64_t nzmax, int triplet) {
  CSparseMatrix* A = (CSparseMatrix*)calloc(1, sizeof *A);
  if (!A) return NULL;
  A->m = m;
  A->n = n;
  A->nzmax = nzmax;
  A->nz = triplet ? 0 : -1;
  A->p = (int64_t*)calloc((size_t)(triplet ? (nzmax ? nzmax : 1) : n + 1), sizeof(int64_t));
  A->i = (int64_t*)calloc((size_t)(nzmax ? nzmax : 1), sizeof(int64_t));
  A->x = (double*)calloc((size_t)(nzmax ? nzmax : 1), sizeof(double));
  if (!A->p || !A->i || !A->x) {
    free(A->p);
    free(A->i);
    free(A->x);
    free(A);
    return NULL;
  }
  return A;
}

static void cs_free(CSparseMatrix* A) {
  if (!A) return;
  free(A->p);
  free(A->i);
  free(A->x);
  free(A);
}

NumericsSparseMatrix* NSM_free(NumericsSparseMatrix* A) {
  if (!A) return NULL;
  CSparseMatrix* all[] = {A->triplet, A->half_triplet, A->csc, A->trans_csc, A->csr};
  for (int k = 0; k < 5; k++)
    for (int j = 0; j < k; j++)
      if (all[j] == all[k]) all[k] = NULL; /* shared pointers freed once */
  for (int k = 0; k < 5; k++) cs_free(all[k]);
  free(A->diag_indx);
  free(A);
  return NULL;
}

/* CSparseMatrix.c:651-790 */
static CSparseMatrix* cs_read(FILE* f) {
  char line[2048];
  for (int k = 0; k < 3; k++) /* rest of the origin line, banner, size line */
    if (!fgets(line, sizeof line, f)) return NULL;
  const int triplet = strncmp(line, "triplet:", 8) == 0;
  const int csr = strncmp(line, "csr:", 4) == 0;
  long long m = 0, n = 0, nzmax = 0, nz = 0;
  const char* s = line + (triplet ? 8 : csr ? 4 : 0);
  if (sscanf(s, " %lld-by-%lld, nzmax: %lld nnz: %lld", &m, &n, &nzmax, &nz) < 3 || m < 0 ||
      n < 0 || nzmax < 0)
    return NULL;
  CSparseMatrix* A = cs_new(m, n, triplet ? (nz > nzmax ? nz : nzmax) : nzmax, triplet);
  if (!A) return NULL;
  if (triplet) {
    if (nz < 0 || nz > A->nzmax) goto fail;
    A->nz = nz;
    for (long long k = 0; k < nz; k++) {
      long long i, j;
      double v;
      if (fscanf(f, " %lld %lld : %le", &i, &j, &v) != 3 || i < 0 || i >= m || j < 0 || j >= n)
        goto fail;
      A->i[k] = i;
      A->p[k] = j;
      A->x[k] = v;
    }
    return A;
  }
  A->nz = csr ? -2 : -1;
  const long long outer = csr ? m : n, inner = csr ? n : m;
  if (csr) {
    /* p holds m + 1 row pointers */
    free(A->p);
    A->p = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
    if (!A->p) goto fail;
  }
  for (long long j = 0; j < outer; j++) {
    long long jj, a, b;
    if (!fgets(line, sizeof line, f)) goto fail;
    if (sscanf(line, " %*s %lld : locations %lld to %lld", &jj, &a, &b) != 3 || jj != j ||
        a != A->p[j] || b + 1 < a || b + 1 > nzmax)
      goto fail;
    A->p[j + 1] = b + 1;
    for (long long k = a; k <= b; k++) {
      long long i;
      double v;
      if (!fgets(line, sizeof line, f)) goto fail;
      if (sscanf(line, " %lld : %le", &i, &v) != 2 || i < 0 || i >= inner) goto fail;
      A->i[k] = i;
      A->x[k] = v;
    }
  }
  return A;
fail:
  cs_free(A);
  return NULL;
}

NumericsSparseMatrix* NSM_new_from_file(FILE* file) {
  int origin = 0;
  if (!file || fscanf(file, "%d", &origin) != 1) return NULL;
  CSparseMatrix* C = cs_read(file);
  if (!C) return NULL;
  NumericsSparseMatrix* A = (NumericsSparseMatrix*)calloc(1, sizeof *A);
  if (!A) {
    cs_free(C);
    return NULL;
  }
  if (C->nz >= 0) {
    A->triplet = C;
    A->origin = NSM_TRIPLET;
  } else if (C->nz == -2) {
    A->csr = C;
    A->origin = NSM_CSR;
  } else if (origin == NSM_HALF_TRIPLET) {
    A->half_triplet = C;
    A->origin = NSM_HALF_TRIPLET;
  } else {
    A->csc = C;
    A->origin = NSM_CSC;
  }
  return A;
}

/* cs_compress of a triplet: per-column counts, entries kept in triplet
 * order inside a column (duplicates are not summed) */
static CSparseMatrix* compress(const CSparseMatrix* T) {
  CSparseMatrix* C = cs_new(T->m, T->n, T->nz, 0);
  if (!C) return NULL;
  int64_t* w = (int64_t*)calloc((size_t)T->n + 1, sizeof(int64_t));
  if (!w) {
    cs_free(C);
    return NULL;
  }
  for (int64_t k = 0; k < T->nz; k++) w[T->p[k]]++;
  for (int64_t j = 0; j < T->n; j++) C->p[j + 1] = C->p[j] + w[j];
  for (int64_t j = 0; j < T->n; j++) w[j] = C->p[j];
  for (int64_t k = 0; k < T->nz; k++) {
    const int64_t q = w[T->p[k]]++;
    C->i[q] = T->i[k];
    C->x[q] = T->x[k];
  }
  free(w);
  return C;
}

/* cs_transpose of a compressed matrix (CSR <-> CSC) */
static CSparseMatrix* transpose(const CSparseMatrix* A, int64_t outer, int64_t inner) {
  const int64_t nnz = A->p[outer];
  CSparseMatrix* C = cs_new(inner, outer, nnz, 0);
  if (!C) return NULL;
  free(C->p);
  C->p = (int64_t*)calloc((size_t)inner + 1, sizeof(int64_t));
  int64_t* w = (int64_t*)calloc((size_t)inner + 1, sizeof(int64_t));
  if (!C->p || !w) {
    free(w);
    cs_free(C);
    return NULL;
  }
  for (int64_t p = 0; p < nnz; p++) w[A->i[p]]++;
  for (int64_t i = 0; i < inner; i++) C->p[i + 1] = C->p[i] + w[i];
  for (int64_t i = 0; i < inner; i++) w[i] = C->p[i];
  for (int64_t j = 0; j < outer; j++)
    for (int64_t p = A->p[j]; p < A->p[j + 1]; p++) {
      const int64_t q = w[A->i[p]]++;
      C->i[q] = j;
      C->x[q] = A->x[p];
    }
  free(w);
  return C;
}

CSparseMatrix* NM_csc(NumericsMatrix* A) {
  if (!A || A->storageType != NM_SPARSE || !A->matrix2) return NULL;
  NumericsSparseMatrix* S = A->matrix2;
  if (S->csc) return S->csc;
  if (S->triplet) {
    S->csc = compress(S->triplet);
  } else if (S->csr) {
    /* the CSR arrays are the CSC of the transpose */
    S->csc = transpose(S->csr, S->csr->m, S->csr->n);
    if (S->csc) {
      S->csc->m = S->csr->m;
      S->csc->n = S->csr->n;
    }
  }
  return S->csc;
}

/* CSparseMatrix.c:883-940 */
static double one_norm(const CSparseMatrix* A) {
  double best = 0.0;
  for (int64_t j = 0; j < A->n; j++) {
    double s = 0.0;
    for (int64_t p = A->p[j]; p < A->p[j + 1]; p++) s += fabs(A->x[p]);
    if (s > best) best = s;
  }
  return best;
}

int NSM_write_in_file(const NumericsSparseMatrix* S, FILE* f) {
  const CSparseMatrix* A = S->origin == NSM_TRIPLET ? S->triplet
                         : S->origin == NSM_CSR     ? S->csr
                         : S->origin == NSM_HALF_TRIPLET ? S->half_triplet
                                                         : S->csc;
  if (!A) return NUMERICS_ERR_INVALID_ARGUMENT;
  fprintf(f, "%d\n", (int)S->origin);
  fprintf(f, "CSparse-compatible listing (siconos_amd)\n");
  if (A->nz >= 0) {
    fprintf(f, "triplet: %lld-by-%lld, nzmax: %lld nnz: %lld\n", (long long)A->m,
            (long long)A->n, (long long)A->nzmax, (long long)A->nz);
    for (int64_t k = 0; k < A->nz; k++)
      fprintf(f, "    %lld %lld : %.17g\n", (long long)A->i[k], (long long)A->p[k], A->x[k]);
    return 0;
  }
  const int csr = A->nz == -2;
  const int64_t outer = csr ? A->m : A->n;
  fprintf(f, "%s%lld-by-%lld, nzmax: %lld nnz: %lld, 1-norm: %g\n", csr ? "csr: " : "",
          (long long)A->m, (long long)A->n, (long long)A->nzmax, (long long)A->p[outer],
          csr ? 0.0 : one_norm(A));
  for (int64_t j = 0; j < outer; j++) {
    fprintf(f, "    %s %lld : locations %lld to %lld\n", csr ? "row" : "col", (long long)j,
            (long long)A->p[j], (long long)A->p[j + 1] - 1);
    for (int64_t p = A->p[j]; p < A->p[j + 1]; p++)
      fprintf(f, "      %lld : %.17g\n", (long long)A->i[p], A->x[p]);
  }
  return 0;
}
