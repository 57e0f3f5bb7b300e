/* generator.c — seeded synthetic fc3d problems of spheres (no reference
 * counterpart: in the reference W comes from the mechanics kernel,
 * LinearOSNS.cpp:522-700, with contact Jacobians of the form of
 * NewtonEuler3DR.cpp:76-125).
 *
 * Each contact c joins bodies a and b, has unit normal n and the frame
 * R = [n; t1; t2] of orthoBaseFromVector (op3x3.h:817-858).  With lever arms
 * x_a - p_c = -r n and x_b - p_c = +r n,
 *     H_{c,a} = [ R | R S(-r n) ],   H_{c,b} = -[ R | R S(r n) ],
 * S(v) w = v x w, and W_{cd} = sum over shared bodies k of
 * H_{c,k} M_k^{-1} H_{d,k}^T with M_k^{-1} = diag(1/m I3, 1/I I3).
 * For r = 1, m = 1, I = 0.4 every diagonal block is diag(2, 7, 7).
 * q = H v_free.  All random numbers come from splitmix64 streams. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fc3d_numerics.h"

typedef struct {
  uint64_t s;
} rng_t;

static uint64_t rng_next(rng_t* r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double rng_uniform(rng_t* r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
static double rng_normal(rng_t* r) {
  /* Box-Muller on (0,1] x [0,1) */
  double u1 = 1.0 - rng_uniform(r);
  double u2 = rng_uniform(r);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}
static rng_t rng_stream(uint64_t seed, uint64_t stream) {
  rng_t r = {seed * 0xD1B54A32D192ED03ULL + stream * 0x9E3779B97F4A7C15ULL + 1ULL};
  rng_next(&r);
  return r;
}

static void ortho_base(const double* n, double* t1, double* t2) {
  double sign = copysign(1.0, n[2]);
  double a = -1.0 / (sign + n[2]);
  double b = n[0] * n[1] * a;
  t1[0] = 1.0 + sign * n[0] * n[0] * a; t1[1] = sign * b; t1[2] = -sign * n[0];
  t2[0] = b; t2[1] = sign + n[1] * n[1] * a; t2[2] = -n[1];
}

/* Per contact: translational and rotational 3x3 parts of H for both bodies,
 * row-major [3][3]. */
typedef struct {
  double Ta[9], Ra[9], Tb[9], Rb[9];
} contact_jac;

static void make_jacobian(const double* n, double radius, contact_jac* J) {
  double t1[3], t2[3];
  ortho_base(n, t1, t2);
  const double* rows[3] = {n, t1, t2};
  double va[3] = {-radius * n[0], -radius * n[1], -radius * n[2]};
  double vb[3] = {radius * n[0], radius * n[1], radius * n[2]};
  for (int i = 0; i < 3; i++) {
    const double* Ri = rows[i];
    /* S(v) = [[0,-vz,vy],[vz,0,-vx],[-vy,vx,0]];  (R S(v))_{i,j} = sum_k Ri_k S_kj */
    double Sa[9] = {0, -va[2], va[1], va[2], 0, -va[0], -va[1], va[0], 0};
    double Sb[9] = {0, -vb[2], vb[1], vb[2], 0, -vb[0], -vb[1], vb[0], 0};
    for (int j = 0; j < 3; j++) {
      double sa = 0, sb = 0;
      for (int k = 0; k < 3; k++) {
        sa += Ri[k] * Sa[3 * k + j];
        sb += Ri[k] * Sb[3 * k + j];
      }
      J->Ta[3 * i + j] = Ri[j];
      J->Ra[3 * i + j] = sa;
      J->Tb[3 * i + j] = -Ri[j];
      J->Rb[3 * i + j] = -sb;
    }
  }
}

/* Body incidence: contacts of each body, CSR. */
typedef struct {
  int* ptr;
  int* list;
} incidence;

static void build_incidence(int nb, int nc, const int* A, const int* B, incidence* inc) {
  inc->ptr = (int*)calloc((size_t)nb + 1, sizeof(int));
  inc->list = (int*)malloc(sizeof(int) * 2 * (size_t)nc);
  for (int c = 0; c < nc; c++) {
    inc->ptr[A[c] + 1]++;
    inc->ptr[B[c] + 1]++;
  }
  for (int k = 0; k < nb; k++) inc->ptr[k + 1] += inc->ptr[k];
  int* fill = (int*)malloc(sizeof(int) * (size_t)nb);
  memcpy(fill, inc->ptr, sizeof(int) * (size_t)nb);
  for (int c = 0; c < nc; c++) {
    inc->list[fill[A[c]]++] = c;
    inc->list[fill[B[c]]++] = c;
  }
  free(fill);
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* T and R parts of contact c for body k (row-major 3x3). */
static void part_for(const contact_jac* J, const int* A, int c, int k, const double** T,
                     const double** Rr) {
  if (A[c] == k) {
    *T = J[c].Ta;
    *Rr = J[c].Ra;
  } else {
    *T = J[c].Tb;
    *Rr = J[c].Rb;
  }
}

static int assemble(int nb, int nc, const int* A, const int* B, const double* normals,
                    const double* vfree, const double* mu, double minv_t, double minv_r,
                    fc3d_generated* out) {
  contact_jac* J = (contact_jac*)malloc(sizeof(contact_jac) * (size_t)nc);
  if (!J) return NUMERICS_ERR_MEMORY;
  for (int c = 0; c < nc; c++) make_jacobian(normals + 3 * c, 1.0, &J[c]);
  incidence inc;
  build_incidence(nb, nc, A, B, &inc);
  out->nc = nc;
  out->nbodies = nb;
  out->row_ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nc + 1));
  /* count: row c holds c and every other contact of bodies A[c], B[c] */
  int64_t nnz = 0;
  int* tmp = (int*)malloc(sizeof(int) * 1024);
  int tcap = 1024;
  out->row_ptr[0] = 0;
  for (int c = 0; c < nc; c++) {
    int n = (inc.ptr[A[c] + 1] - inc.ptr[A[c]]) + (inc.ptr[B[c] + 1] - inc.ptr[B[c]]) - 1;
    nnz += n;
    out->row_ptr[c + 1] = nnz;
    if (n > tcap) {
      tcap = 2 * n;
      tmp = (int*)realloc(tmp, sizeof(int) * (size_t)tcap);
    }
  }
  out->nnzb = nnz;
  out->col_idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  out->blocks = (double*)malloc(sizeof(double) * 9 * (size_t)(nnz ? nnz : 1));
  out->q = (double*)malloc(sizeof(double) * 3 * (size_t)nc);
  out->mu = (double*)malloc(sizeof(double) * (size_t)nc);
  out->body_a = (int*)malloc(sizeof(int) * (size_t)nc);
  out->body_b = (int*)malloc(sizeof(int) * (size_t)nc);
  if (!out->col_idx || !out->blocks || !out->q || !out->mu) return NUMERICS_ERR_MEMORY;
  memcpy(out->body_a, A, sizeof(int) * (size_t)nc);
  memcpy(out->body_b, B, sizeof(int) * (size_t)nc);
  memcpy(out->mu, mu, sizeof(double) * (size_t)nc);
  /* the inputs of the assembly, for callers that assemble elsewhere */
  out->H = (double*)malloc(sizeof(double) * 36 * (size_t)nc);
  out->minv = (double*)malloc(sizeof(double) * 6 * (size_t)nb);
  out->vfree = (double*)malloc(sizeof(double) * 6 * (size_t)nb);
  if (!out->H || !out->minv || !out->vfree) return NUMERICS_ERR_MEMORY;
  for (int c = 0; c < nc; c++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double* h = out->H + 36 * (size_t)c;
        h[i + 3 * j] = J[c].Ta[3 * i + j];
        h[i + 3 * (j + 3)] = J[c].Ra[3 * i + j];
        h[18 + i + 3 * j] = J[c].Tb[3 * i + j];
        h[18 + i + 3 * (j + 3)] = J[c].Rb[3 * i + j];
      }
  for (int k = 0; k < nb; k++)
    for (int j = 0; j < 6; j++) out->minv[6 * (size_t)k + j] = j < 3 ? minv_t : minv_r;
  memcpy(out->vfree, vfree, sizeof(double) * 6 * (size_t)nb);
  for (int c = 0; c < nc; c++) {
    int n = 0;
    for (int t = inc.ptr[A[c]]; t < inc.ptr[A[c] + 1]; t++) tmp[n++] = inc.list[t];
    for (int t = inc.ptr[B[c]]; t < inc.ptr[B[c] + 1]; t++)
      if (inc.list[t] != c) tmp[n++] = inc.list[t];
    qsort(tmp, (size_t)n, sizeof(int), cmp_int);
    int64_t base = out->row_ptr[c];
    for (int t = 0; t < n; t++) {
      int d = tmp[t];
      double* blk = out->blocks + 9 * (base + t);
      memset(blk, 0, 9 * sizeof(double));
      out->col_idx[base + t] = d;
      const int bodies[2] = {A[c], B[c]};
      for (int s = 0; s < 2; s++) {
        int k = bodies[s];
        if (A[d] != k && B[d] != k) continue;
        const double *Tc, *Rc, *Td, *Rd;
        part_for(J, A, c, k, &Tc, &Rc);
        part_for(J, A, d, k, &Td, &Rd);
        /* blk(i,j) += sum_l Tc(i,l) Td(j,l)/m + Rc(i,l) Rd(j,l)/I, column-major */
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            double st = 0, sr = 0;
            for (int l = 0; l < 3; l++) {
              st += Tc[3 * i + l] * Td[3 * j + l];
              sr += Rc[3 * i + l] * Rd[3 * j + l];
            }
            blk[i + 3 * j] += minv_t * st + minv_r * sr;
          }
      }
    }
    /* q_c = H_{c,a} v_a + H_{c,b} v_b */
    for (int i = 0; i < 3; i++) {
      double s = 0;
      const double* va = vfree + 6 * A[c];
      const double* vb = vfree + 6 * B[c];
      for (int l = 0; l < 3; l++) {
        s += J[c].Ta[3 * i + l] * va[l] + J[c].Ra[3 * i + l] * va[3 + l];
        s += J[c].Tb[3 * i + l] * vb[l] + J[c].Rb[3 * i + l] * vb[3 + l];
      }
      out->q[3 * c + i] = s;
    }
  }
  free(tmp);
  free(inc.ptr);
  free(inc.list);
  free(J);
  return 0;
}

int fc3d_generate_random_pairs(int nbodies, int nc, double mu_lo, double mu_hi, uint64_t seed,
                               fc3d_generated* out) {
  if (!out || nbodies < 2 || nc <= 0) return NUMERICS_ERR_INVALID_ARGUMENT;
  if ((double)nc > 0.5 * (double)nbodies * (double)(nbodies - 1))
    return NUMERICS_ERR_INVALID_ARGUMENT;
  memset(out, 0, sizeof *out);
  int* A = (int*)malloc(sizeof(int) * (size_t)nc);
  int* B = (int*)malloc(sizeof(int) * (size_t)nc);
  /* distinct unordered pairs through an open-addressing set */
  size_t cap = 1;
  while (cap < 4 * (size_t)nc) cap <<= 1;
  uint64_t* set = (uint64_t*)calloc(cap, sizeof(uint64_t));
  rng_t rp = rng_stream(seed, 1);
  for (int c = 0; c < nc;) {
    int a = (int)(rng_next(&rp) % (uint64_t)nbodies);
    int b = (int)(rng_next(&rp) % (uint64_t)nbodies);
    if (a == b) continue;
    if (a > b) { int t = a; a = b; b = t; }
    uint64_t key = ((uint64_t)a << 32 | (uint64_t)b) + 1;
    size_t h = (size_t)((key * 0x9E3779B97F4A7C15ULL) >> 20) & (cap - 1);
    int dup = 0;
    while (set[h]) {
      if (set[h] == key) { dup = 1; break; }
      h = (h + 1) & (cap - 1);
    }
    if (dup) continue;
    set[h] = key;
    A[c] = a;
    B[c] = b;
    c++;
  }
  free(set);
  double* normals = (double*)malloc(sizeof(double) * 3 * (size_t)nc);
  rng_t rn = rng_stream(seed, 2);
  for (int c = 0; c < nc; c++) {
    double x, y, z, r2;
    do {
      x = rng_normal(&rn); y = rng_normal(&rn); z = rng_normal(&rn);
      r2 = x * x + y * y + z * z;
    } while (r2 < 1e-12);
    double inv = 1.0 / sqrt(r2);
    normals[3 * c] = x * inv; normals[3 * c + 1] = y * inv; normals[3 * c + 2] = z * inv;
  }
  double* vfree = (double*)malloc(sizeof(double) * 6 * (size_t)nbodies);
  rng_t rv = rng_stream(seed, 3);
  for (int k = 0; k < 6 * nbodies; k++) vfree[k] = rng_normal(&rv);
  double* mu = (double*)malloc(sizeof(double) * (size_t)nc);
  rng_t rm = rng_stream(seed, 4);
  for (int c = 0; c < nc; c++) mu[c] = mu_lo + (mu_hi - mu_lo) * rng_uniform(&rm);
  int rc = assemble(nbodies, nc, A, B, normals, vfree, mu, 1.0, 1.0 / 0.4, out);
  free(A); free(B); free(normals); free(vfree); free(mu);
  return rc;
}

int fc3d_generate_granular_pile(int nbodies, int nc, double mu_lo, double mu_hi, uint64_t seed,
                                fc3d_generated* out) {
  if (!out || nbodies < 8 || nc <= 0) return NUMERICS_ERR_INVALID_ARGUMENT;
  memset(out, 0, sizeof *out);
  /* lattice n x n x h with n*n*h >= nbodies, bodies in lattice order */
  int n = (int)ceil(cbrt((double)nbodies));
  int h = (nbodies + n * n - 1) / (n * n);
  rng_t rj = rng_stream(seed, 11);
  double* X = (double*)malloc(sizeof(double) * 3 * (size_t)nbodies);
  for (int k = 0; k < nbodies; k++) {
    int i = k % n, j = (k / n) % n, l = k / (n * n);
    X[3 * k] = 2.0 * i + 0.05 * (2 * rng_uniform(&rj) - 1);
    X[3 * k + 1] = 2.0 * j + 0.05 * (2 * rng_uniform(&rj) - 1);
    X[3 * k + 2] = 2.0 * l + 0.05 * (2 * rng_uniform(&rj) - 1);
  }
  (void)h;
  /* candidate contacts: face and edge neighbours in the forward half */
  static const int dirs[9][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0},
                                 {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
  size_t ncand = 0, ccap = 9 * (size_t)nbodies;
  uint64_t* cand = (uint64_t*)malloc(sizeof(uint64_t) * ccap);
  for (int k = 0; k < nbodies; k++) {
    int i = k % n, j = (k / n) % n, l = k / (n * n);
    for (int d = 0; d < 9; d++) {
      int i2 = i + dirs[d][0], j2 = j + dirs[d][1], l2 = l + dirs[d][2];
      if (i2 < 0 || i2 >= n || j2 < 0 || j2 >= n || l2 < 0) continue;
      long k2 = (long)i2 + (long)n * j2 + (long)n * n * l2;
      if (k2 >= nbodies) continue;
      cand[ncand++] = ((uint64_t)k << 32) | (uint64_t)k2;
    }
  }
  if ((size_t)nc > ncand) {
    free(X);
    free(cand);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  /* partial Fisher-Yates: the first nc candidates of a seeded shuffle */
  rng_t rs = rng_stream(seed, 12);
  for (size_t t = 0; t < (size_t)nc; t++) {
    size_t u = t + (size_t)(rng_next(&rs) % (uint64_t)(ncand - t));
    uint64_t tmp = cand[t];
    cand[t] = cand[u];
    cand[u] = tmp;
  }
  /* keep spatial locality: contacts ordered by (a, b) */
  qsort(cand, (size_t)nc, sizeof(uint64_t), cmp_u64);
  int* A = (int*)malloc(sizeof(int) * (size_t)nc);
  int* B = (int*)malloc(sizeof(int) * (size_t)nc);
  double* normals = (double*)malloc(sizeof(double) * 3 * (size_t)nc);
  for (int c = 0; c < nc; c++) {
    A[c] = (int)(cand[c] >> 32);
    B[c] = (int)(cand[c] & 0xffffffffu);
    double d[3], s = 0;
    for (int e = 0; e < 3; e++) {
      d[e] = X[3 * B[c] + e] - X[3 * A[c] + e];
      s += d[e] * d[e];
    }
    s = 1.0 / sqrt(s);
    for (int e = 0; e < 3; e++) normals[3 * c + e] = d[e] * s;
  }
  double* vfree = (double*)malloc(sizeof(double) * 6 * (size_t)nbodies);
  rng_t rv = rng_stream(seed, 13), rsub = rng_stream(seed, 14);
  for (int k = 0; k < nbodies; k++) {
    double scale = rng_uniform(&rsub) < 0.1 ? 0.1 : 1.0;
    for (int e = 0; e < 6; e++) vfree[6 * k + e] = scale * rng_normal(&rv);
  }
  double* mu = (double*)malloc(sizeof(double) * (size_t)nc);
  rng_t rm = rng_stream(seed, 15);
  for (int c = 0; c < nc; c++) mu[c] = mu_lo + (mu_hi - mu_lo) * rng_uniform(&rm);
  int rc = assemble(nbodies, nc, A, B, normals, vfree, mu, 1.0, 1.0 / 0.4, out);
  free(X); free(cand); free(A); free(B); free(normals); free(vfree); free(mu);
  return rc;
}

void fc3d_generated_free(fc3d_generated* g) {
  if (!g) return;
  free(g->row_ptr); free(g->col_idx); free(g->blocks); free(g->q); free(g->mu);
  free(g->body_a); free(g->body_b);
  free(g->H); free(g->minv); free(g->vfree);
  memset(g, 0, sizeof *g);
}

FrictionContactProblem* fc3d_generated_to_problem(const fc3d_generated* g) {
  SparseBlockStructuredMatrix* S =
      SBM_new_3x3_from_csr((size_t)g->nc, g->row_ptr, g->col_idx, g->blocks);
  NumericsMatrix* M = NM_create_from_data(NM_SPARSE_BLOCK, 3 * g->nc, 3 * g->nc, S);
  double* q = (double*)malloc(sizeof(double) * 3 * (size_t)g->nc);
  double* mu = (double*)malloc(sizeof(double) * (size_t)g->nc);
  memcpy(q, g->q, sizeof(double) * 3 * (size_t)g->nc);
  memcpy(mu, g->mu, sizeof(double) * (size_t)g->nc);
  return frictionContactProblem_new_with_data(3, (size_t)g->nc, M, q, mu);
}

/* Colouring entry of the host API: runs the GPU colouring. */
int fc3d_hip_colour(int nc, const int64_t* row_ptr, const int64_t* col_idx, int* colour);

int fc3d_colour_contacts(int nc, const int64_t* row_ptr, const int64_t* col_idx, int* colour) {
  return fc3d_hip_colour(nc, row_ptr, col_idx, colour);
}

int fc3d_colouring_is_valid(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                            const int* colour) {
  for (int i = 0; i < nc; i++)
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      int64_t j = col_idx[k];
      if (j != i && colour[j] == colour[i]) return 0;
    }
  return 1;
}
