/* problem.c — FrictionContactProblem container and its ".dat" text format.
 * Format and behaviour of numerics/src/FrictionContact/FrictionContactProblem.c:67-200,
 * numerics/src/linalg/NumericsMatrix.c:1551-1700 (NM_write_in_file /
 * NM_new_from_file, dense branch) and SparseBlockMatrix.c:1619-1800
 * (SBM_write_in_file / SBM_new_from_file), for dimension 2 or 3.  Compressed
 * (NM_SPARSE) storage is read and written by sparse.c. */
#include <errno.h>
#include <stdlib.h>
#include <string.h>

#include "fc3d_numerics.h"

FrictionContactProblem* frictionContactProblem_new(void) {
  return (FrictionContactProblem*)calloc(1, sizeof(FrictionContactProblem));
}

FrictionContactProblem* frictionContactProblem_new_with_data(int dim, size_t nc, NumericsMatrix* M,
                                                             double* q, double* mu) {
  FrictionContactProblem* p = frictionContactProblem_new();
  p->dimension = dim;
  p->numberOfContacts = (int)nc;
  p->M = M;
  p->q = q;
  p->mu = mu;
  return p;
}

void frictionContactProblem_free(FrictionContactProblem* p) {
  if (!p) return;
  if (p->M) NM_free(p->M);
  free(p->q);
  free(p->mu);
  free(p);
}

static void sbm_write(const SparseBlockStructuredMatrix* m, FILE* f, int dim) {
  fprintf(f, "%zu\n", m->nbblocks);
  if (m->nbblocks == 0) return;
  fprintf(f, "%zu\n%zu\n", m->blocknumber0, m->blocknumber1);
  for (size_t i = 0; i < m->blocknumber0; i++) fprintf(f, "%zu\t", m->blocksize0[i]);
  fprintf(f, "\n");
  for (size_t i = 0; i < m->blocknumber1; i++) fprintf(f, "%zu\t", m->blocksize1[i]);
  fprintf(f, "\n%li\n%li\n", (long)m->filled1, (long)m->filled2);
  for (size_t i = 0; i < m->filled1; i++) fprintf(f, "%li\t", (long)m->index1_data[i]);
  fprintf(f, "\n");
  for (size_t i = 0; i < m->filled2; i++) fprintf(f, "%li\t", (long)m->index2_data[i]);
  fprintf(f, "\n");
  for (size_t row = 0; row + 1 < m->filled1; row++)
    for (size_t k = m->index1_data[row]; k < m->index1_data[row + 1]; k++) {
      fprintf(f, "%zu\n", k);
      for (int e = 0; e < dim * dim; e++) fprintf(f, "%32.24e\n", m->block[k][e]);
    }
}

int frictionContact_printInFile(FrictionContactProblem* p, FILE* f) {
  if (!p || !f || !p->M) return NUMERICS_ERR_INVALID_ARGUMENT;
  fprintf(f, "%d\n%d\n", p->dimension, p->numberOfContacts);
  NumericsMatrix* m = p->M;
  fprintf(f, "%d\n%d\n%d\n", m->storageType, m->size0, m->size1);
  if (m->storageType == NM_DENSE) {
    fprintf(f, "%i\t%i\n", m->size0, m->size1);
    for (int i = 0; i < m->size1 * m->size0; i++) {
      fprintf(f, "%32.24e ", m->matrix0[i]);
      if ((i + 1) % m->size1 == 0) fprintf(f, "\n");
    }
  } else if (m->storageType == NM_SPARSE_BLOCK) {
    sbm_write(m->matrix1, f, p->dimension);
  } else if (m->storageType == NM_SPARSE && m->matrix2) {
    int rc = NSM_write_in_file(m->matrix2, f);
    if (rc) return rc;
  } else {
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  for (int i = 0; i < m->size1; i++) fprintf(f, "%32.24e ", p->q[i]);
  fprintf(f, "\n");
  for (int i = 0; i < p->numberOfContacts; i++) fprintf(f, "%32.24e ", p->mu[i]);
  fprintf(f, "\n");
  return 0;
}

int frictionContact_printInFilename(FrictionContactProblem* p, char* filename) {
  FILE* f = fopen(filename, "w");
  if (!f) return errno;
  int rc = frictionContact_printInFile(p, f);
  fclose(f);
  return rc;
}

static SparseBlockStructuredMatrix* sbm_read(FILE* f, int dim) {
  SparseBlockStructuredMatrix* m = SBM_new();
  size_t v;
  if (fscanf(f, "%zu", &m->nbblocks) != 1 || m->nbblocks == 0) goto fail;
  if (fscanf(f, "%zu", &m->blocknumber0) != 1 || fscanf(f, "%zu", &m->blocknumber1) != 1) goto fail;
  m->blocksize0 = (size_t*)malloc(sizeof(size_t) * m->blocknumber0);
  m->blocksize1 = (size_t*)malloc(sizeof(size_t) * m->blocknumber1);
  for (size_t i = 0; i < m->blocknumber0; i++) {
    if (fscanf(f, "%zu", &m->blocksize0[i]) != 1 || m->blocksize0[i] != (size_t)dim * (i + 1))
      goto fail;
  }
  for (size_t i = 0; i < m->blocknumber1; i++) {
    if (fscanf(f, "%zu", &m->blocksize1[i]) != 1 || m->blocksize1[i] != (size_t)dim * (i + 1))
      goto fail;
  }
  if (fscanf(f, "%zu", &m->filled1) != 1 || fscanf(f, "%zu", &m->filled2) != 1) goto fail;
  m->index1_data = (size_t*)malloc(sizeof(size_t) * m->filled1);
  m->index2_data = (size_t*)malloc(sizeof(size_t) * (m->filled2 ? m->filled2 : 1));
  for (size_t i = 0; i < m->filled1; i++)
    if (fscanf(f, "%zu", &m->index1_data[i]) != 1) goto fail;
  for (size_t i = 0; i < m->filled2; i++)
    if (fscanf(f, "%zu", &m->index2_data[i]) != 1) goto fail;
  m->block = (double**)calloc(m->nbblocks, sizeof(double*));
  for (size_t row = 0; row + 1 < m->filled1; row++)
    for (size_t k = m->index1_data[row]; k < m->index1_data[row + 1]; k++) {
      if (k >= m->nbblocks || fscanf(f, "%zu", &v) != 1) goto fail;
      m->block[k] = (double*)malloc((size_t)(dim * dim) * sizeof(double));
      for (int e = 0; e < dim * dim; e++)
        if (fscanf(f, "%lf", &m->block[k][e]) != 1) goto fail;
    }
  return m;
fail:
  SBM_free_all(m);
  return NULL;
}

FrictionContactProblem* frictionContact_newFromFile(FILE* f) {
  if (!f) return NULL;
  int d = 0, nc = 0, st = 0, s0 = 0, s1 = 0;
  if (fscanf(f, "%d", &d) != 1 || fscanf(f, "%d", &nc) != 1) return NULL;
  if (fscanf(f, "%d", &st) != 1 || fscanf(f, "%d", &s0) != 1 || fscanf(f, "%d", &s1) != 1)
    return NULL;
  if ((d != 2 && d != 3) || nc <= 0 || s0 != d * nc || s1 != d * nc) return NULL;
  NumericsMatrix* M = NULL;
  if (st == NM_DENSE) {
    int a, b;
    if (fscanf(f, "%d", &a) != 1 || fscanf(f, "%d", &b) != 1) return NULL;
    double* data = (double*)malloc(sizeof(double) * (size_t)s0 * (size_t)s1);
    for (size_t i = 0; i < (size_t)s0 * (size_t)s1; i++)
      if (fscanf(f, "%lf", &data[i]) != 1) {
        free(data);
        return NULL;
      }
    M = NM_create_from_data(NM_DENSE, s0, s1, data);
  } else if (st == NM_SPARSE_BLOCK) {
    SparseBlockStructuredMatrix* S = sbm_read(f, d);
    if (!S) return NULL;
    M = NM_create_from_data(NM_SPARSE_BLOCK, s0, s1, S);
  } else if (st == NM_SPARSE) {
    NumericsSparseMatrix* S = NSM_new_from_file(f);
    if (!S) return NULL;
    M = NM_create_from_data(NM_SPARSE, s0, s1, S);
  } else {
    return NULL;
  }
  double* q = (double*)malloc(sizeof(double) * (size_t)s1);
  double* mu = (double*)malloc(sizeof(double) * (size_t)nc);
  for (int i = 0; i < s1; i++)
    if (fscanf(f, "%lf", &q[i]) != 1) goto fail;
  for (int i = 0; i < nc; i++)
    if (fscanf(f, "%lf", &mu[i]) != 1) goto fail;
  return frictionContactProblem_new_with_data(d, (size_t)nc, M, q, mu);
fail:
  free(q);
  free(mu);
  NM_free(M);
  return NULL;
}

FrictionContactProblem* frictionContact_new_from_filename(const char* filename) {
  FILE* f = fopen(filename, "r");
  if (!f) return NULL;
  FrictionContactProblem* p = frictionContact_newFromFile(f);
  fclose(f);
  return p;
}
