/* matrix.c — NumericsMatrix / SparseBlockStructuredMatrix containers (host).
 * Storage layout of numerics/src/linalg/SparseBlockMatrix.h:82-124 and
 * NumericsMatrix.h:73-92; products run on the GPU (fc3d_hip_gemv). */
#include <stdlib.h>
#include <string.h>

#include "fc3d_hip.h"
#include "fc3d_numerics.h"

SparseBlockStructuredMatrix* SBM_new(void) {
  return (SparseBlockStructuredMatrix*)calloc(1, sizeof(SparseBlockStructuredMatrix));
}

void SBM_free_all(SparseBlockStructuredMatrix* m) {
  if (!m) return;
  if (m->block) {
    for (size_t k = 0; k < m->nbblocks; k++) free(m->block[k]);
    free(m->block);
  }
  free(m->blocksize0);
  free(m->blocksize1);
  free(m->index1_data);
  free(m->index2_data);
  free(m->diagonal_blocks);
  free(m);
}

SparseBlockStructuredMatrix* SBM_new_3x3_from_csr(size_t nrows, const int64_t* row_ptr,
                                                  const int64_t* col_idx, const double* values) {
  if (!row_ptr || (row_ptr[nrows] > 0 && (!col_idx || !values))) return NULL;
  SparseBlockStructuredMatrix* m = SBM_new();
  size_t nb = (size_t)row_ptr[nrows];
  m->nbblocks = nb;
  m->blocknumber0 = m->blocknumber1 = nrows;
  m->blocksize0 = (size_t*)malloc(sizeof(size_t) * (nrows ? nrows : 1));
  m->blocksize1 = (size_t*)malloc(sizeof(size_t) * (nrows ? nrows : 1));
  for (size_t i = 0; i < nrows; i++) m->blocksize0[i] = m->blocksize1[i] = 3 * (i + 1);
  m->filled1 = nrows + 1;
  m->filled2 = nb;
  m->index1_data = (size_t*)malloc(sizeof(size_t) * (nrows + 1));
  m->index2_data = (size_t*)malloc(sizeof(size_t) * (nb ? nb : 1));
  m->block = (double**)malloc(sizeof(double*) * (nb ? nb : 1));
  for (size_t i = 0; i <= nrows; i++) m->index1_data[i] = (size_t)row_ptr[i];
  for (size_t k = 0; k < nb; k++) {
    m->index2_data[k] = (size_t)col_idx[k];
    m->block[k] = (double*)malloc(9 * sizeof(double));
    memcpy(m->block[k], values + 9 * k, 9 * sizeof(double));
  }
  return m;
}

size_t* SBM_compute_diagonal_block_indices(SparseBlockStructuredMatrix* M) {
  if (!M) return NULL;
  if (M->diagonal_blocks) return M->diagonal_blocks;
  M->diagonal_blocks = (size_t*)malloc(sizeof(size_t) * (M->blocknumber0 ? M->blocknumber0 : 1));
  for (size_t row = 0; row + 1 < M->filled1; ++row) {
    M->diagonal_blocks[row] = (size_t)-1;
    for (size_t k = M->index1_data[row]; k < M->index1_data[row + 1]; ++k)
      if (M->index2_data[k] == row) {
        M->diagonal_blocks[row] = k;
        break;
      }
  }
  return M->diagonal_blocks;
}

size_t SBM_diagonal_block_index(SparseBlockStructuredMatrix* M, size_t row) {
  return SBM_compute_diagonal_block_indices(M)[row];
}

SparseBlockStructuredMatrix* SBM_from_dense_3x3(int n, const double* dense) {
  /* n = number of 3x3 block rows; dense is 3n x 3n column-major */
  if (n <= 0 || !dense) return NULL;
  size_t N = 3 * (size_t)n;
  int64_t* rp = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
  int64_t cap = 16, nb = 0;
  int64_t* ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
  double* vals = (double*)malloc(sizeof(double) * 9 * (size_t)cap);
  rp[0] = 0;
  for (int I = 0; I < n; I++) {
    for (int J = 0; J < n; J++) {
      double b[9];
      int nz = 0;
      for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) {
          b[r + 3 * c] = dense[(size_t)(3 * I + r) + N * (size_t)(3 * J + c)];
          if (b[r + 3 * c] != 0.0) nz = 1;
        }
      if (!nz && I != J) continue;
      if (nb == cap) {
        cap *= 2;
        ci = (int64_t*)realloc(ci, sizeof(int64_t) * (size_t)cap);
        vals = (double*)realloc(vals, sizeof(double) * 9 * (size_t)cap);
      }
      ci[nb] = J;
      memcpy(vals + 9 * nb, b, sizeof b);
      nb++;
    }
    rp[I + 1] = nb;
  }
  SparseBlockStructuredMatrix* m = SBM_new_3x3_from_csr((size_t)n, rp, ci, vals);
  free(rp);
  free(ci);
  free(vals);
  return m;
}

NumericsMatrix* NM_create_from_data(int storageType, int size0, int size1, void* data) {
  NumericsMatrix* m = (NumericsMatrix*)calloc(1, sizeof(NumericsMatrix));
  m->storageType = (NM_types)storageType;
  m->size0 = size0;
  m->size1 = size1;
  if (storageType == NM_DENSE)
    m->matrix0 = (double*)data;
  else if (storageType == NM_SPARSE_BLOCK)
    m->matrix1 = (SparseBlockStructuredMatrix*)data;
  else if (storageType == NM_SPARSE)
    m->matrix2 = (NumericsSparseMatrix*)data;
  m->destructible = m;
  return m;
}

NumericsMatrix* NM_free(NumericsMatrix* m) {
  if (!m) return NULL;
  fc3d_hip_resident_clear(m);  /* its address may be reused by another matrix */
  free(m->matrix0);
  SBM_free_all(m->matrix1);
  NSM_free(m->matrix2);
  free(m);
  return NULL;
}

/* CSR arrays (int64) of an SBM with 3x3 blocks; contiguous values. */
int fc3d_sbm_to_csr(const SparseBlockStructuredMatrix* A, int64_t** rp, int64_t** ci,
                    double** vals) {
  size_t n = A->filled1 ? A->filled1 - 1 : 0, nb = A->nbblocks;
  *rp = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  *ci = (int64_t*)malloc(sizeof(int64_t) * (nb ? nb : 1));
  *vals = (double*)malloc(sizeof(double) * 9 * (nb ? nb : 1));
  if (!*rp || !*ci || !*vals) return NUMERICS_ERR_MEMORY;
  for (size_t i = 0; i <= n; i++) (*rp)[i] = (int64_t)A->index1_data[i];
  for (size_t k = 0; k < nb; k++) {
    (*ci)[k] = (int64_t)A->index2_data[k];
    memcpy(*vals + 9 * k, A->block[k], 9 * sizeof(double));
  }
  return 0;
}

int SBM_gemv_3x3(size_t sizeX, size_t sizeY, const SparseBlockStructuredMatrix* A, double* x,
                 double* y) {
  if (!A || !x || !y) return NUMERICS_ERR_NULL_POINTER;
  size_t n = A->filled1 ? A->filled1 - 1 : 0;
  if (sizeX != 3 * n || sizeY != 3 * n) return NUMERICS_ERR_INVALID_ARGUMENT;
  int64_t *rp, *ci;
  double* vals;
  int rc = fc3d_sbm_to_csr(A, &rp, &ci, &vals);
  double* q0 = (double*)calloc(3 * n + 1, sizeof(double));
  double* mu0 = (double*)calloc(n + 1, sizeof(double));
  fc3d_hip_problem* h = NULL;
  if (!rc) rc = fc3d_hip_problem_create((int)n, rp, ci, vals, q0, mu0, 0, &h);
  if (!rc) rc = fc3d_hip_gemv(h, x, y);
  fc3d_hip_problem_destroy(h);
  free(rp);
  free(ci);
  free(vals);
  free(q0);
  free(mu0);
  return rc;
}

int NM_prod_mv_3x3(int sizeX, int sizeY, NumericsMatrix* A, double* x, double* y) {
  if (!A) return NUMERICS_ERR_NULL_POINTER;
  if (A->storageType == NM_SPARSE_BLOCK) return SBM_gemv_3x3((size_t)sizeX, (size_t)sizeY, A->matrix1, x, y);
  if (A->storageType == NM_DENSE) {
    SparseBlockStructuredMatrix* S = SBM_from_dense_3x3(A->size0 / 3, A->matrix0);
    int rc = SBM_gemv_3x3((size_t)sizeX, (size_t)sizeY, S, x, y);
    SBM_free_all(S);
    return rc;
  }
  return NUMERICS_ERR_NOT_IMPLEMENTED;
}
