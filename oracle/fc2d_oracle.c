/* fc2d_oracle.c — CPU oracle for the fc2d NSGS path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's 2-D
 * friction-contact NSGS, which tests/ compare the GPU path against.  The product library
 * never links, loads or calls it.  Each function cites the reference
 * file:line it follows (paths relative to the Siconos tree).
 *
 * It reads the .dat problem files itself (oracle_fc2d_read), independently
 * of the product's reader, so that the two can be compared.
 *
 * Pinning (tests/test_oracle_fc2d.py):
 *   - oracle_fc2d_local_solve: closed forms for a diagonal block (take-off,
 *     stick, slide in both directions) and, for random non-symmetric blocks,
 *     the 2-D Coulomb complementarity of the result checked through the
 *     independent numpy residual; a planted slide-sign mistake is caught.
 *   - oracle_fc2d_compute_error: an independent numpy residual on dense W.
 *   - row products: numpy dense products on every storage of every fixture.
 *   - oracle_fc2d_nsgs: the reference's own test (test_fc2d_1.c with
 *     data_collection_fc2d_1.c): FC2D_NSGS at tol 1e-5, 10000 iterations,
 *     on the eleven fixtures; every run must end with info = 0 and a
 *     residual <= 1e-5 recomputed by numpy.
 * Unpinned: the reference's dense row product and gemv go through a vendor
 * BLAS (cblas_ddot / cblas_dgemv) whose summation order is not fixed; this
 * file sums in ascending column order.  Norms use a plain sqrt of the sum of
 * squares, not cblas_dnrm2.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fc3d_oracle.h"

extern int g_oracle_fc2d_mut;
int g_oracle_fc2d_mut = 0;

/* ---------------------------------------------------------------------- */
/* Reading (FrictionContactProblem.c frictionContact_newFromFile,          */
/* NumericsMatrix.c NM_new_from_file, SparseBlockMatrix.c SBM_new_from_file,*/
/* NumericsSparseMatrix.c:555-583, CSparseMatrix.c:651-790)                */
/* ---------------------------------------------------------------------- */
typedef struct {
  int nc;
  int storage;      /* 0 dense, 1 SBM 2x2, 2 sparse                       */
  int origin;       /* sparse: NSM_TRIPLET / NSM_CSC / NSM_CSR            */
  double* dense;    /* 2nc x 2nc column-major                             */
  /* SBM */
  size_t nbblocks;
  size_t* rowp;     /* nc + 1                                             */
  size_t* colb;     /* nbblocks                                           */
  double* blocks;   /* 4 per block, column-major                          */
  /* sparse: the stored form (csc/csr compressed, or triplet)              */
  int64_t m, n, nz, nzmax;
  int64_t* sp;      /* compressed: n + 1 pointers; triplet: column index */
  int64_t* si;
  double* sx;
  double* q;
  double* mu;
} fc2d_problem;

static void fc2d_free(fc2d_problem* p) {
  if (!p) return;
  free(p->dense);
  free(p->rowp);
  free(p->colb);
  free(p->blocks);
  free(p->sp);
  free(p->si);
  free(p->sx);
  free(p->q);
  free(p->mu);
  free(p);
}

static int read_sparse(FILE* f, fc2d_problem* P) {
  if (fscanf(f, "%d", &P->origin) != 1) return -1;
  char line[2048];
  /* the rest of the origin line, the CSparse banner, the size line */
  for (int k = 0; k < 3; k++)
    if (!fgets(line, sizeof line, f)) return -1;
  long long m = 0, n = 0, nzmax = 0, nz = 0;
  int triplet = strncmp(line, "triplet:", 8) == 0;
  int csr = strncmp(line, "csr:", 4) == 0;
  const char* s = line + (triplet ? 8 : csr ? 4 : 0);
  if (sscanf(s, " %lld-by-%lld, nzmax: %lld nnz: %lld", &m, &n, &nzmax, &nz) < 3) return -1;
  P->m = m;
  P->n = n;
  P->nzmax = nzmax;
  if (triplet) {
    P->nz = nz;
    P->sp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nz ? nz : 1));
    P->si = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nz ? nz : 1));
    P->sx = (double*)malloc(sizeof(double) * (size_t)(nz ? nz : 1));
    for (long long k = 0; k < nz; k++) {
      long long a, b;
      double v;
      if (fscanf(f, " %lld %lld : %le", &a, &b, &v) != 3) return -1;
      P->si[k] = a;
      P->sp[k] = b;
      P->sx[k] = v;
    }
  } else {
    P->nz = csr ? -2 : -1;
    P->sp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    P->si = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nzmax ? nzmax : 1));
    P->sx = (double*)malloc(sizeof(double) * (size_t)(nzmax ? nzmax : 1));
    for (long long j = 0; j < n; j++) {
      long long jj, a, b;
      if (!fgets(line, sizeof line, f)) return -1;
      if (sscanf(line, " %*s %lld : locations %lld to %lld", &jj, &a, &b) != 3 || jj != j)
        return -1;
      P->sp[j] = a;
      P->sp[j + 1] = b + 1;
      if (b + 1 > nzmax) return -1;
      for (long long k = a; k <= b; k++) {
        long long i;
        double v;
        if (!fgets(line, sizeof line, f)) return -1;
        if (sscanf(line, " %lld : %le", &i, &v) != 2) return -1;
        P->si[k] = i;
        P->sx[k] = v;
      }
    }
  }
  return 0;
}

static int read_sbm(FILE* f, fc2d_problem* P) {
  size_t nb, b0, b1, v;
  if (fscanf(f, "%zu %zu %zu", &nb, &b0, &b1) != 3) return -1;
  for (size_t i = 0; i < b0 + b1; i++)
    if (fscanf(f, "%zu", &v) != 1) return -1;
  size_t f1, f2;
  if (fscanf(f, "%zu %zu", &f1, &f2) != 2 || f1 != (size_t)P->nc + 1 || f2 != nb) return -1;
  P->nbblocks = nb;
  P->rowp = (size_t*)malloc(sizeof(size_t) * f1);
  P->colb = (size_t*)malloc(sizeof(size_t) * (nb ? nb : 1));
  P->blocks = (double*)malloc(sizeof(double) * 4 * (nb ? nb : 1));
  for (size_t i = 0; i < f1; i++)
    if (fscanf(f, "%zu", &P->rowp[i]) != 1) return -1;
  for (size_t i = 0; i < nb; i++)
    if (fscanf(f, "%zu", &P->colb[i]) != 1) return -1;
  for (size_t k = 0; k < nb; k++) {
    if (fscanf(f, "%zu", &v) != 1) return -1;
    for (int e = 0; e < 4; e++)
      if (fscanf(f, "%le", &P->blocks[4 * k + e]) != 1) return -1;
  }
  return 0;
}

void* oracle_fc2d_read(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return NULL;
  fc2d_problem* P = (fc2d_problem*)calloc(1, sizeof *P);
  int d, st, s0, s1;
  if (fscanf(f, "%d %d %d %d %d", &d, &P->nc, &st, &s0, &s1) != 5 || d != 2 || P->nc <= 0 ||
      s0 != 2 * P->nc || s1 != 2 * P->nc)
    goto fail;
  P->storage = st;
  if (st == 0) {
    int a, b;
    if (fscanf(f, "%d %d", &a, &b) != 2) goto fail;
    P->dense = (double*)malloc(sizeof(double) * (size_t)s0 * (size_t)s1);
    for (size_t i = 0; i < (size_t)s0 * (size_t)s1; i++)
      if (fscanf(f, "%le", &P->dense[i]) != 1) goto fail;
  } else if (st == 1) {
    if (read_sbm(f, P)) goto fail;
  } else if (st == 2) {
    if (read_sparse(f, P)) goto fail;
  } else {
    goto fail;
  }
  P->q = (double*)malloc(sizeof(double) * 2 * (size_t)P->nc);
  P->mu = (double*)malloc(sizeof(double) * (size_t)P->nc);
  for (int i = 0; i < 2 * P->nc; i++)
    if (fscanf(f, "%le", &P->q[i]) != 1) goto fail;
  for (int i = 0; i < P->nc; i++)
    if (fscanf(f, "%le", &P->mu[i]) != 1) goto fail;
  fclose(f);
  return P;
fail:
  fclose(f);
  fc2d_free(P);
  return NULL;
}

void oracle_fc2d_free(void* h) { fc2d_free((fc2d_problem*)h); }

int oracle_fc2d_info(void* h, int* out) {
  fc2d_problem* P = (fc2d_problem*)h;
  out[0] = P->nc;
  out[1] = P->storage;
  out[2] = P->origin;
  return 0;
}

/* CSC of a triplet (cs_compress: column counts, entries in triplet order) */
static void triplet_to_csc(const fc2d_problem* P, int64_t** cp, int64_t** ci, double** cx) {
  int64_t n = P->n, nz = P->nz;
  *cp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  *ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nz ? nz : 1));
  *cx = (double*)malloc(sizeof(double) * (size_t)(nz ? nz : 1));
  int64_t* w = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t k = 0; k < nz; k++) w[P->sp[k]]++;
  for (int64_t j = 0; j < n; j++) (*cp)[j + 1] = (*cp)[j] + w[j];
  for (int64_t j = 0; j < n; j++) w[j] = (*cp)[j];
  for (int64_t k = 0; k < nz; k++) {
    int64_t q = w[P->sp[k]]++;
    (*ci)[q] = P->si[k];
    (*cx)[q] = P->sx[k];
  }
  free(w);
}

/* The CSC form NM_csc gives (NumericsMatrix.c NM_csc); caller frees. */
static void sparse_csc(const fc2d_problem* P, int64_t** cp, int64_t** ci, double** cx) {
  if (P->nz >= 0) {
    triplet_to_csc(P, cp, ci, cx);
    return;
  }
  int64_t nnz = P->sp[P->n];
  *cp = (int64_t*)malloc(sizeof(int64_t) * ((size_t)P->n + 1));
  *ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  *cx = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  memcpy(*cp, P->sp, sizeof(int64_t) * ((size_t)P->n + 1));
  memcpy(*ci, P->si, sizeof(int64_t) * (size_t)nnz);
  memcpy(*cx, P->sx, sizeof(double) * (size_t)nnz);
}

/* cs_transpose: column i of the result lists row i of A, ascending column */
static void transpose(int64_t m, int64_t n, const int64_t* ap, const int64_t* ai,
                      const double* ax, int64_t** tp, int64_t** ti, double** tx) {
  int64_t nnz = ap[n];
  *tp = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
  *ti = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  *tx = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  int64_t* w = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
  for (int64_t p = 0; p < nnz; p++) w[ai[p]]++;
  for (int64_t i = 0; i < m; i++) (*tp)[i + 1] = (*tp)[i] + w[i];
  for (int64_t i = 0; i < m; i++) w[i] = (*tp)[i];
  for (int64_t j = 0; j < n; j++)
    for (int64_t p = ap[j]; p < ap[j + 1]; p++) {
      int64_t q = w[ai[p]]++;
      (*ti)[q] = j;
      (*tx)[q] = ax[p];
    }
  free(w);
}

/* Row form of a sparse problem as NM_row_prod_no_diag2 reads it
 * (NumericsMatrix.c:710-741): NM_csr for a CSR origin, else the transpose of
 * the CSC (NM_csc_trans). */
typedef struct {
  int64_t* rp;
  int64_t* ri;
  double* rx;
  /* CSC, for the diagonal block and the gemv */
  int64_t* cp;
  int64_t* ci;
  double* cx;
} sparse_rows;

static void rows_of(const fc2d_problem* P, sparse_rows* R) {
  int64_t n2 = 2 * (int64_t)P->nc;
  if (P->nz == -2) {
    /* stored CSR: rows as stored; the CSC is its transpose */
    int64_t nnz = P->sp[n2];
    R->rp = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n2 + 1));
    R->ri = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    R->rx = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    memcpy(R->rp, P->sp, sizeof(int64_t) * ((size_t)n2 + 1));
    memcpy(R->ri, P->si, sizeof(int64_t) * (size_t)nnz);
    memcpy(R->rx, P->sx, sizeof(double) * (size_t)nnz);
    transpose(n2, n2, R->rp, R->ri, R->rx, &R->cp, &R->ci, &R->cx);
    return;
  }
  sparse_csc(P, &R->cp, &R->ci, &R->cx);
  transpose(n2, n2, R->cp, R->ci, R->cx, &R->rp, &R->ri, &R->rx);
}

static void rows_free(sparse_rows* R) {
  free(R->rp);
  free(R->ri);
  free(R->rx);
  free(R->cp);
  free(R->ci);
  free(R->cx);
}

/* NM_extract_diag_block2 (NumericsMatrix.c:1799-1826; the sparse branch is
 * NSM_extract_block, NumericsSparseMatrix.c:782-820) */
static void diag_block(const fc2d_problem* P, const sparse_rows* R, int c, double* B) {
  if (P->storage == 0) {
    size_t n = 2 * (size_t)P->nc;
    const double* M = P->dense + (n + 1) * (2 * (size_t)c);
    B[0] = M[0];
    B[1] = M[1];
    M += n;
    B[2] = M[0];
    B[3] = M[1];
  } else if (P->storage == 1) {
    B[0] = B[1] = B[2] = B[3] = 0.0;
    for (size_t k = P->rowp[c]; k < P->rowp[c + 1]; k++)
      if (P->colb[k] == (size_t)c) {
        memcpy(B, P->blocks + 4 * k, 4 * sizeof(double));
        break;
      }
  } else {
    B[0] = B[1] = B[2] = B[3] = 0.0;
    int64_t r0 = 2 * (int64_t)c;
    for (int64_t j = r0; j < r0 + 2; j++)
      for (int64_t p = R->cp[j]; p < R->cp[j + 1]; p++) {
        int64_t row = R->ci[p];
        if (row >= r0) {
          if (row >= r0 + 2) break;
          B[(j - r0) * 2 + row - r0] = R->cx[p];
        }
      }
  }
}

/* op3x3.h:316-334, as written: column 0 of the block, then column 1 */
static void mvp2x2(const double* a, const double* v, double* r) {
  r[0] += a[0] * v[0];
  r[1] += a[1] * v[0];
  r[0] += a[2] * v[1];
  r[1] += a[3] * v[1];
}

/* fc2d_nsgs.c:97-113 with NM_row_prod_no_diag2 (NumericsMatrix.c:674-749):
 * q_loc = q_c + sum_{d != c} W_cd r_d; x[2c], x[2c+1] are zeroed for the
 * dense and sparse forms, the SBM form skips the diagonal block. */
static void local_q(const fc2d_problem* P, const sparse_rows* R, int c, double* x, double* y) {
  size_t in = 2 * (size_t)c, it = in + 1;
  y[0] = P->q[in];
  y[1] = P->q[it];
  if (P->storage == 1) {
    for (size_t k = P->rowp[c]; k < P->rowp[c + 1]; k++)
      if (P->colb[k] != (size_t)c) mvp2x2(P->blocks + 4 * k, x + 2 * P->colb[k], y);
    return;
  }
  double rin = x[in], rit = x[it];
  x[in] = 0.0;
  x[it] = 0.0;
  if (P->storage == 0) {
    size_t n = 2 * (size_t)P->nc;
    double d0 = 0.0, d1 = 0.0;
    for (size_t j = 0; j < n; j++) {
      d0 += P->dense[in + j * n] * x[j];
      d1 += P->dense[it + j * n] * x[j];
    }
    y[0] += d0;
    y[1] += d1;
  } else {
    for (int i = 0; i < 2; i++)
      for (int64_t p = R->rp[in + i]; p < R->rp[in + i + 1]; p++) y[i] += R->rx[p] * x[R->ri[p]];
  }
  x[in] = rin;
  x[it] = rit;
}

/* fc2d_nsgs.c:247-284 */
void oracle_fc2d_local_solve(const double* W, double D, const double* q, double mu, double* P) {
  const double Wnn = W[0], Wtn = W[1], Wnt = W[2], Wtt = W[3];
  if (q[0] > 0) {
    P[0] = 0;
    P[1] = 0;
  } else {
    P[0] = -(Wtt * q[0] - Wnt * q[1]) / D;
    P[1] = -(-Wtn * q[0] + Wnn * q[1]) / D;
    double muPn = mu * P[0];
    if (fabs(P[1]) > muPn) {
      int neg = P[1] + muPn < 0;
      if (g_oracle_fc2d_mut == 1) neg = !neg; /* planted: slide directions swapped */
      if (neg) {
        P[0] = -q[0] / (Wnn - mu * Wnt);
        P[1] = -mu * P[0];
      } else {
        P[0] = -q[0] / (Wnn + mu * Wnt);
        P[1] = mu * P[0];
      }
    }
  }
}

void oracle_fc2d_set_mutation(int m) { g_oracle_fc2d_mut = m; }

/* fc2d_compute_error.c:37-56 */
static void unitary_error(const double* r, const double* u, double mu, double* error) {
  double w0 = r[0] - u[0] - mu * fabs(u[1]);
  double w1 = r[1] - u[1];
  double normT = fabs(w1);
  if (mu * normT <= -w0) {
    w0 = 0.;
    w1 = 0.;
  } else if (normT > mu * w0) {
    w0 = (mu * normT + w0) / (mu * mu + 1.0);
    w1 = mu * w0 * (w1 < 0 ? -1 : w1 > 0 ? 1 : 0);
  }
  w0 = r[0] - w0;
  w1 = r[1] - w1;
  *error += w0 * w0 + w1 * w1;
}

static double nrm2(const double* x, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += x[i] * x[i];
  return sqrt(s);
}

/* w = q + W z (NM_gemv, NumericsMatrix.c:3089-3117) */
static void gemv(const fc2d_problem* P, const sparse_rows* R, const double* z, double* w) {
  size_t n = 2 * (size_t)P->nc;
  memcpy(w, P->q, sizeof(double) * n);
  if (P->storage == 0) {
    for (size_t j = 0; j < n; j++)
      for (size_t i = 0; i < n; i++) w[i] += P->dense[i + j * n] * z[j];
  } else if (P->storage == 1) {
    for (int c = 0; c < P->nc; c++)
      for (size_t k = P->rowp[c]; k < P->rowp[c + 1]; k++)
        mvp2x2(P->blocks + 4 * k, z + 2 * P->colb[k], w + 2 * (size_t)c);
  } else {
    /* CSparseMatrix_aaxpby (CSparseMatrix.c:109-140), alpha = beta = 1 */
    for (size_t j = 0; j < n; j++)
      for (int64_t p = R->cp[j]; p < R->cp[j + 1]; p++) w[R->ci[p]] += R->cx[p] * z[j];
  }
}

/* fc2d_compute_error.c:104-148 */
static double compute_error(const fc2d_problem* P, const sparse_rows* R, const double* z,
                            double* w, double norm) {
  size_t n = 2 * (size_t)P->nc;
  gemv(P, R, z, w);
  double error = 0.0;
  for (int ic = 0; ic < P->nc; ic++) unitary_error(z + 2 * ic, w + 2 * ic, P->mu[ic], &error);
  error = sqrt(error);
  double norm_r = nrm2(z, n), norm_w = nrm2(w, n);
  double scale = fmax(norm_r, fmax(norm_w, norm));
  if (fabs(scale) > DBL_EPSILON) error /= scale;
  return error;
}

int oracle_fc2d_compute_error(void* h, const double* z, double* w, double* error) {
  fc2d_problem* P = (fc2d_problem*)h;
  sparse_rows R = {0};
  if (P->storage == 2) rows_of(P, &R);
  *error = compute_error(P, &R, z, w, nrm2(P->q, 2 * (size_t)P->nc));
  rows_free(&R);
  return 0;
}

/* The matrix as dense column-major 2nc x 2nc (for the tests' numpy checks). */
int oracle_fc2d_dense(void* h, double* out, double* q, double* mu) {
  fc2d_problem* P = (fc2d_problem*)h;
  size_t n = 2 * (size_t)P->nc;
  memset(out, 0, sizeof(double) * n * n);
  if (P->storage == 0) {
    memcpy(out, P->dense, sizeof(double) * n * n);
  } else if (P->storage == 1) {
    for (int c = 0; c < P->nc; c++)
      for (size_t k = P->rowp[c]; k < P->rowp[c + 1]; k++) {
        size_t d = P->colb[k];
        const double* b = P->blocks + 4 * k;
        out[2 * c + 2 * d * n] += b[0];
        out[2 * c + 1 + 2 * d * n] += b[1];
        out[2 * c + (2 * d + 1) * n] += b[2];
        out[2 * c + 1 + (2 * d + 1) * n] += b[3];
      }
  } else {
    sparse_rows R = {0};
    rows_of(P, &R);
    for (size_t j = 0; j < n; j++)
      for (int64_t p = R.cp[j]; p < R.cp[j + 1]; p++) out[R.ci[p] + j * n] += R.cx[p];
    rows_free(&R);
  }
  memcpy(q, P->q, sizeof(double) * n);
  memcpy(mu, P->mu, sizeof(double) * (size_t)P->nc);
  return 0;
}

/* q_loc of contact c at reaction x (for the row-product pins). */
int oracle_fc2d_local_q(void* h, int c, double* x, double* y) {
  fc2d_problem* P = (fc2d_problem*)h;
  sparse_rows R = {0};
  if (P->storage == 2) rows_of(P, &R);
  local_q(P, &R, c, x, y);
  rows_free(&R);
  return 0;
}

/* fc2d_nsgs.c:297-538 without contact freezing (refused: its criteria
 * depend on the natural order) and without shuffling.  `order` (nc
 * entries, or NULL for the natural order) is the sequence of contacts in a
 * sweep.  mode: SICONOS_NSGS_ERROR_EVALUATION_LIGHT or _LIGHT_WITH_FULL_FINAL,
 * the only two fc2d_nsgs evaluates (with any other mode it sweeps itermax
 * times and reports info 1).
 * Returns info; out[0] iterations, out[1] residual. */
int oracle_fc2d_nsgs(void* h, double* reaction, double* velocity, double tol, int itermax,
                     int mode, const int* order, double* out) {
  fc2d_problem* P = (fc2d_problem*)h;
  const int nc = P->nc;
  const size_t n = 2 * (size_t)nc;
  sparse_rows R = {0};
  if (P->storage == 2) rows_of(P, &R);
  double norm_q = nrm2(P->q, n);
  double* B = (double*)malloc(sizeof(double) * 4 * (size_t)nc);
  double* D = (double*)malloc(sizeof(double) * (size_t)nc);
  for (int c = 0; c < nc; c++) {
    diag_block(P, &R, c, B + 4 * c);
    D[c] = B[4 * c] * B[4 * c + 3] - B[4 * c + 1] * B[4 * c + 2];
    if (D[c] < DBL_EPSILON) {
      /* fc2d_nsgs.c:228-245, 318-334 */
      out[0] = 0;
      out[1] = INFINITY;
      free(B);
      free(D);
      rows_free(&R);
      return 1;
    }
  }
  int iter = 0;
  double error = INFINITY, tolerance = tol;
  int not_conv = 1;
  const int light_mode = mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT ||
                         mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL;
  while (iter < itermax && not_conv) {
    ++iter;
    double light = 0.0;
    for (int k = 0; k < nc; k++) {
      const int c = order ? order[k] : k;
      const size_t pos = 2 * (size_t)c;
      double lr[2] = {reaction[pos], reaction[pos + 1]};
      double ql[2];
      local_q(P, &R, c, reaction, ql);
      oracle_fc2d_local_solve(B + 4 * c, D[c], ql, P->mu[c], lr);
      if (light_mode) {
        double x0 = reaction[pos] - lr[0], x1 = reaction[pos + 1] - lr[1];
        light += x0 * x0 + x1 * x1;
      }
      reaction[pos] = lr[0];
      reaction[pos + 1] = lr[1];
    }
    if (light_mode) {
      /* fc2d_nsgs.c:136-142 */
      error = sqrt(light);
      double norm_r = nrm2(reaction, n);
      if (fabs(norm_r) > DBL_EPSILON) error /= norm_r;
      if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT) {
        not_conv = error < tolerance ? 0 : 1;
      } else {
        /* fc2d_nsgs.c:191-228 */
        not_conv = 1;
        if (error < tolerance) {
          not_conv = 0;
          double abs_err = compute_error(P, &R, reaction, velocity, norm_q);
          if (abs_err > tol) {
            tolerance = error / abs_err * tol;
            not_conv = 1;
          }
        }
      }
    }
  }
  if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    error = compute_error(P, &R, reaction, velocity, norm_q);
    not_conv = error < tol ? 0 : 1;
  }
  out[0] = iter;
  out[1] = error;
  free(B);
  free(D);
  rows_free(&R);
  return not_conv;
}
