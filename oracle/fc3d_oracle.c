/* fc3d_oracle.c — CPU oracle for the fc3d NSGS hot path (TEST INFRASTRUCTURE).
 *
 * Plain C99 restatement of the reference algorithm.  Only tests/, smoke() and
 * bench.py's cpu_baseline leg use it, as the reference results are compared
 * against; the product never does.
 * Built with -ffp-contract=off so every product of the reference is rounded
 * before the sum it feeds, as in an IEEE-754 build of the reference C.
 *
 * What pins each part (tests/test_oracle_*.py):
 *   op3x3 (mvp/mm/gepp)      : random 3x3 systems vs numpy.linalg.solve; the
 *                              pivot branches are exercised by permuted
 *                              systems; mutation 6 must fail.
 *   projectionOnCone         : idempotence, orthogonality (x-Px).Px = 0 and
 *                              the three status cases (pattern of the
 *                              reference tools/test/test_projection.c).
 *   computeAlartCurnierSTD   : against oracle_alart_curnier_dual (independent
 *                              formulation + dual-number derivatives) on every
 *                              row of the reference fixture ACinputs.dat, plus
 *                              finite differences; mutation 1 must fail.
 *   rho split spectral norm  : closed form for W = diag(2,7,7) blocks.
 *   eig_3x3, spectral rho    : numpy.linalg.eigvalsh on random symmetric and
 *                              SPD matrices (repeated eigenvalues included),
 *                              the unsorted diagonal case; mutation 9 must
 *                              fail.
 *   fc3d_compute_error       : zero at the FC3D_Example1 closed-form solution;
 *                              independent numpy evaluation elsewhere;
 *                              mutations 2 and 3 must fail.
 *   local solvers            : closed forms for W = diag(a,b,b) (take-off,
 *                              stick, slide) and a brute-force 1-D angle search
 *                              on FrictionContact3D_1c.dat.
 *   OC_PROJ step (552)       : a brute-force solution is a fixed point of the
 *                              step; from zero with W = diag(a,b,b) the step
 *                              is the Moreau projection of -(q_n + mu|q_t|,
 *                              q_t)/a; mutations 7 and 8 must fail.
 *   fc3d_nsgs                : FC3D_Example1 closed form; residual <= tol on the
 *                              reference fixtures (data_collection_1/3 files)
 *                              under the test_nsgs_1.c option sets; sweep-order
 *                              independence inside a colour.
 *   contact freezing         : the reference's freeze collection converges
 *                              (test_nsgs_freeze_1.c); a Python replay of the
 *                              counters, criteria, reset rule and light error
 *                              from the reactions after every sweep (through
 *                              the callback); mutation 10 must fail.
 *   greedy colouring         : brute-force validity + minimality check
 *                              (each contact's colour is the smallest free one).
 * Unpinned (no fixture or import can pin them here): ULP-level agreement with
 * a vendor cblas_dnrm2 (this oracle uses sqrt of the running sum of squares),
 * PETSc's greedy colouring, libc rand() shuffles (not supported).
 */
#include "fc3d_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static int g_mut = 0;
void oracle_set_mutation(int m) { g_mut = m; }

/* work counters (not thread-safe; for single-threaded statistics only) */
static long g_cnt[8];
void oracle_counters(long* out, int reset) {
  for (int k = 0; k < 8; k++) {
    out[k] = g_cnt[k];
    if (reset) g_cnt[k] = 0;
  }
}

/* ------------------------------------------------------------------------ */
/* op3x3 — op3x3.h:342-363 (mvp3x3), 449-465 (mm3x3), 868-981 (gepp)       */
/* ------------------------------------------------------------------------ */
void oracle_mvp3x3(const double* a, const double* v, double* r) {
  if (g_mut == ORACLE_MUT_MVP_TRANSPOSED) {
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) r[i] += a[j + 3 * i] * v[j];
    return;
  }
  /* column by column: r += a(:,j) v_j */
  for (int j = 0; j < 3; j++) {
    r[0] += a[3 * j + 0] * v[j];
    r[1] += a[3 * j + 1] * v[j];
    r[2] += a[3 * j + 2] * v[j];
  }
}

void oracle_mm3x3(const double* a, const double* b, double* c) {
  /* c_ij = a_i0 b_0j + a_i1 b_1j + a_i2 b_2j, column-major */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      c[i + 3 * j] = a[i + 0] * b[0 + 3 * j] + a[i + 3] * b[1 + 3 * j] + a[i + 6] * b[2 + 3 * j];
}

int oracle_solve_3x3_gepp(const double* a, double* b) {
  double lp0, lp1, lp2, lm1, lm2, ln1, ln2, bl, bm, bn, f1, f2, s0, s1, s2, alp0;
  double a0 = a[0], a1 = a[1], a2 = a[2];
  double aa0 = fabs(a0), aa1 = fabs(a1), aa2 = fabs(a2);
  int p1 = aa0 >= aa1 ? (aa0 >= aa2 ? 0 : 2) : (aa1 >= aa2 ? 1 : 2);
  if (g_mut == ORACLE_MUT_GEPP_PIVOT) p1 = 0;
  if (p1 == 0) {
    f1 = a1 / a0; f2 = a2 / a0; lp0 = a0; alp0 = fabs(a0); lp1 = a[3]; lp2 = a[6];
    lm1 = a[4] - f1 * lp1; lm2 = a[7] - f1 * lp2; ln1 = a[5] - f2 * lp1; ln2 = a[8] - f2 * lp2;
    bl = b[0]; bm = b[1] - f1 * bl; bn = b[2] - f2 * bl;
  } else if (p1 == 1) {
    f1 = a0 / a1; f2 = a2 / a1; lp0 = a1; alp0 = fabs(a1); lp1 = a[4]; lp2 = a[7];
    lm1 = a[3] - f1 * lp1; lm2 = a[6] - f1 * lp2; ln1 = a[5] - f2 * lp1; ln2 = a[8] - f2 * lp2;
    bl = b[1]; bm = b[0] - f1 * bl; bn = b[2] - f2 * bl;
  } else {
    f1 = a0 / a2; f2 = a1 / a2; lp0 = a2; alp0 = fabs(a2); lp1 = a[5]; lp2 = a[8];
    lm1 = a[3] - f1 * lp1; lm2 = a[6] - f1 * lp2; ln1 = a[4] - f2 * lp1; ln2 = a[7] - f2 * lp2;
    bl = b[2]; bm = b[0] - f1 * bl; bn = b[1] - f2 * bl;
  }
  if (alp0 <= DBL_EPSILON) return 1;
  double alm1 = fabs(lm1), aln1 = fabs(ln1);
  if (alm1 >= aln1) {
    if (alm1 < DBL_EPSILON) return 1;
    f1 = ln1 / lm1;
    s2 = (bn - f1 * bm) / (ln2 - f1 * lm2);
    s1 = (bm - lm2 * s2) / lm1;
  } else {
    if (aln1 < DBL_EPSILON) return 1;
    f1 = lm1 / ln1;
    s2 = (bm - f1 * bn) / (lm2 - f1 * ln2);
    s1 = (bn - ln2 * s2) / ln1;
  }
  s0 = (bl - s1 * lp1 - s2 * lp2) / lp0;
  b[0] = s0; b[1] = s1; b[2] = s2;
  return 0;
}

/* op3x3.h:817-858 (Duff et al. branchless basis) */
int oracle_orthoBaseFromVector(double* n, double* t1, double* t2) {
  double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (nn == 0.) return 1;
  n[0] /= nn; n[1] /= nn; n[2] /= nn;
  double sign = copysign(1.0, n[2]);
  double a = -1.0 / (sign + n[2]);
  double b = n[0] * n[1] * a;
  t1[0] = 1.0 + sign * n[0] * n[0] * a; t1[1] = sign * b; t1[2] = -sign * n[0];
  t2[0] = b; t2[1] = sign + n[1] * n[1] * a; t2[2] = -n[1];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* projectionOnCone.c:32-73                                                 */
/* ------------------------------------------------------------------------ */
unsigned oracle_projectionOnCone(double* r, double mu) {
  double nt = sqrt(r[1] * r[1] + r[2] * r[2]);
  if (mu * nt <= -r[0]) {
    r[0] = 0.0; r[1] = 0.0; r[2] = 0.0;
    return 0; /* PROJCONE_DUAL */
  } else if (nt <= mu * r[0]) {
    return 1; /* PROJCONE_INSIDE */
  } else {
    double mu2 = mu * mu;
    if (g_mut == ORACLE_MUT_CONE_SIGN)
      r[0] = (mu * nt - r[0]) / (mu2 + 1.0);
    else
      r[0] = (mu * nt + r[0]) / (mu2 + 1.0);
    r[1] = mu * r[0] * r[1] / nt;
    r[2] = mu * r[0] * r[2] / nt;
    return 2; /* PROJCONE_BOUNDARY */
  }
}

unsigned oracle_projectionOnDualCone(double* u, double mu) {
  double nt = sqrt(u[1] * u[1] + u[2] * u[2]);
  if (nt <= -mu * u[0]) {
    u[0] = 0.0; u[1] = 0.0; u[2] = 0.0;
    return 0;
  } else if (mu * nt <= u[0]) {
    return 1;
  } else {
    double mu2 = mu * mu;
    u[0] = (nt + mu * u[0]) / (mu2 + 1.0);
    u[1] = u[0] * u[1] / nt;
    u[2] = u[0] * u[2] / nt;
    u[0] = mu * u[0];
    return 2;
  }
}

/* ------------------------------------------------------------------------ */
/* Alart-Curnier — fc3d_AlartCurnier_functions.c:506-646 (STD)              */
/* A and B are column-major; entry (i,j) at [i + 3j].                        */
/* ------------------------------------------------------------------------ */
void oracle_computeAlartCurnierSTD(const double* R, const double* u, double mu,
                                   const double* rho, double* F, double* A, double* B) {
  double RhoN = rho[0], RhoT = rho[1];
  double RVN = R[0] - RhoN * u[0];
  double RVT = R[1] - RhoT * u[1];
  double RVS = R[2] - RhoT * u[2];
  double RV = sqrt(RVT * RVT + RVS * RVS);
  double Radius;
  int AB = (A != NULL) && (B != NULL);
  if (A) { A[3] = 0.; A[6] = 0.; A[1] = 0.; A[2] = 0.; }
  if (B) { B[3] = 0.; B[6] = 0.; B[1] = 0.; B[2] = 0.; }
  if (RVN > 0.0) {
    Radius = mu * RVN;
    F[0] = RhoN * u[0];
    if (AB) { A[0] = RhoN; B[0] = 0.0; }
  } else {
    Radius = 0.0;
    F[0] = R[0];
    if (AB) { A[0] = 0.0; B[0] = 1.0; }
  }
  if (RV <= Radius) {
    F[1] = RhoT * u[1];
    F[2] = RhoT * u[2];
    if (AB) {
      A[4] = RhoT; A[7] = 0.0; A[5] = 0.0; A[8] = RhoT;
      B[1] = 0.0; B[4] = 0.0; B[7] = 0.0; B[2] = 0.0; B[5] = 0.0; B[8] = 0.0;
    }
  } else if (RV > Radius) {
    if (Radius > 0) {
      double RV1 = 1.0 / RV;
      F[1] = R[1] - Radius * RVT * RV1;
      F[2] = R[2] - Radius * RVS * RV1;
      if (AB) {
        double RV3 = RV1 * RV1 * RV1;
        double GTT = RV1 - RVT * RVT * RV3;
        double GTS = -RVT * RVS * RV3;
        double GST = GTS;
        double GSS = RV1 - RVS * RVS * RV3;
        A[1] = mu * RhoN * RVT * RV1;
        if (g_mut == ORACLE_MUT_AC_JACOBIAN_SIGN) A[1] = -A[1];
        A[2] = mu * RhoN * RVS * RV1;
        A[4] = GTT * RhoT * Radius;
        A[7] = GTS * RhoT * Radius;
        A[5] = GST * RhoT * Radius;
        A[8] = GSS * RhoT * Radius;
        B[1] = -mu * RVT * RV1;
        B[4] = 1.0 - GTT * Radius;
        B[7] = -GTS * Radius;
        B[2] = -mu * RVS * RV1;
        B[5] = -GST * Radius;
        B[8] = 1.0 - GSS * Radius;
      }
    } else {
      F[1] = R[1];
      F[2] = R[2];
      if (AB) {
        A[4] = 0.0; A[7] = 0.0; A[5] = 0.0; A[8] = 0.0;
        B[1] = 0.0; B[4] = 1.0; B[7] = 0.0; B[2] = 0.0; B[5] = 0.0; B[8] = 1.0;
      }
    }
  }
}

/* fc3d_AlartCurnier_functions.c:649-753 (Christensen-Pang radius mu*R[0]) */
void oracle_computeAlartCurnierJeanMoreau(const double* R, const double* u, double mu,
                                          const double* rho, double* F, double* A,
                                          double* B) {
  double RhoN = rho[0], RhoT = rho[1];
  double RVN = R[0] - RhoN * u[0];
  double RVT = R[1] - RhoT * u[1];
  double RVS = R[2] - RhoT * u[2];
  double RV = sqrt(RVT * RVT + RVS * RVS);
  double Radius = mu * R[0];
  int AB = (A != NULL) && (B != NULL);
  if (RVN >= 0.0) {
    F[0] = RhoN * u[0];
    if (AB) { A[0] = RhoN; B[0] = 0.0; }
  } else {
    F[0] = R[0];
    if (AB) { A[0] = 0.0; B[0] = 1.0; }
  }
  if (RV < Radius || RV < 1e-20) {
    F[1] = RhoT * u[1];
    F[2] = RhoT * u[2];
    if (AB) {
      A[4] = RhoT; A[7] = 0.0; A[5] = 0.0; A[8] = RhoT;
      B[1] = 0.0; B[4] = 0.0; B[7] = 0.0; B[2] = 0.0; B[5] = 0.0; B[8] = 0.0;
    }
  } else {
    double RV1 = 1.0 / RV;
    F[1] = R[1] - Radius * RVT * RV1;
    F[2] = R[2] - Radius * RVS * RV1;
    if (AB) {
      double RV3 = RV1 * RV1 * RV1;
      double GTT = RV1 - RVT * RVT * RV3;
      double GTS = -RVT * RVS * RV3;
      double GST = GTS;
      double GSS = RV1 - RVS * RVS * RV3;
      A[4] = GTT * RhoT * Radius;
      A[7] = GTS * RhoT * Radius;
      A[5] = GST * RhoT * Radius;
      A[8] = GSS * RhoT * Radius;
      B[1] = -mu * RVT * RV1;
      B[4] = 1.0 - GTT * Radius;
      B[7] = -GTS * Radius;
      B[2] = -mu * RVS * RV1;
      B[5] = -GST * Radius;
      B[8] = 1.0 - GSS * Radius;
    }
  }
}

/* ---- independent dual-number formulation -------------------------------- */
typedef struct { double v; double d[6]; } dual6; /* d/du0..u2, d/dr0..r2 */

static dual6 d_const(double c) { dual6 x; x.v = c; memset(x.d, 0, sizeof x.d); return x; }
static dual6 d_var(double c, int k) { dual6 x = d_const(c); x.d[k] = 1.0; return x; }
static dual6 d_add(dual6 a, dual6 b) { for (int k = 0; k < 6; k++) a.d[k] += b.d[k]; a.v += b.v; return a; }
static dual6 d_sub(dual6 a, dual6 b) { for (int k = 0; k < 6; k++) a.d[k] -= b.d[k]; a.v -= b.v; return a; }
static dual6 d_scale(dual6 a, double s) { for (int k = 0; k < 6; k++) a.d[k] *= s; a.v *= s; return a; }
static dual6 d_mul(dual6 a, dual6 b) {
  dual6 c; c.v = a.v * b.v;
  for (int k = 0; k < 6; k++) c.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return c;
}
static dual6 d_div(dual6 a, dual6 b) {
  dual6 c; c.v = a.v / b.v;
  for (int k = 0; k < 6; k++) c.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
  return c;
}
static dual6 d_sqrt(dual6 a) {
  dual6 c; c.v = sqrt(a.v);
  for (int k = 0; k < 6; k++) c.d[k] = a.d[k] / (2.0 * c.v);
  return c;
}

void oracle_alart_curnier_dual(const double* R, const double* u, double mu, const double* rho,
                               double* F, double* A, double* B) {
  dual6 U[3], Rr[3];
  for (int i = 0; i < 3; i++) { U[i] = d_var(u[i], i); Rr[i] = d_var(R[i], 3 + i); }
  /* x = r - rho u, the argument of the projection */
  dual6 x0 = d_sub(Rr[0], d_scale(U[0], rho[0]));
  dual6 x1 = d_sub(Rr[1], d_scale(U[1], rho[1]));
  dual6 x2 = d_sub(Rr[2], d_scale(U[2], rho[2]));
  /* normal part: projection onto R+ */
  dual6 pn = x0.v > 0.0 ? x0 : d_const(0.0);
  /* tangential part: projection onto the disk of radius mu * pn */
  dual6 radius = d_scale(pn, mu);
  dual6 nx = d_sqrt(d_add(d_mul(x1, x1), d_mul(x2, x2)));
  dual6 p1, p2;
  if (nx.v <= radius.v) {
    p1 = x1; p2 = x2;
  } else if (radius.v > 0.0) {
    p1 = d_mul(radius, d_div(x1, nx));
    p2 = d_mul(radius, d_div(x2, nx));
  } else {
    p1 = d_const(0.0); p2 = d_const(0.0);
  }
  dual6 f[3] = {d_sub(Rr[0], pn), d_sub(Rr[1], p1), d_sub(Rr[2], p2)};
  for (int i = 0; i < 3; i++) {
    F[i] = f[i].v;
    for (int j = 0; j < 3; j++) {
      if (A) A[i + 3 * j] = f[i].d[j];
      if (B) B[i + 3 * j] = f[i].d[3 + j];
    }
  }
}

/* fc3d_AlartCurnier_functions.c:784-812 (quirk kept: sw - W21 + W12) */
void oracle_rho_split_spectral_norm(const double* W, double* rho) {
  double sw = W[1 + 1 * 3] + W[2 + 2 * 3];
  double dw = sw * sw - 4.0 * (sw - W[2 + 1 * 3] + W[1 + 2 * 3]);
  dw = dw > 0.0 ? sqrt(dw) : 0.0;
  rho[0] = 1.0 / W[0];
  rho[1] = 2.0 / (sw + dw);
  rho[2] = rho[1];
}

/* fc3d_AlartCurnier_functions.c:755-782 */
void oracle_rho_split_spectral_norm_cond(const double* W, double* rho) {
  double sw = W[1 + 1 * 3] + W[2 + 2 * 3];
  double dw = sw * sw - 4.0 * (sw - W[2 + 1 * 3] + W[1 + 2 * 3]);
  dw = dw > 0.0 ? sqrt(dw) : 0.0;
  rho[0] = 1.0 / W[0];
  rho[1] = 2.0 * (sw - dw) / ((sw + dw) * (sw + dw));
  rho[2] = rho[1];
}

/* op3x3.h:1065-1111: eigenvalues of a symmetric column-major 3x3 matrix by
 * the trigonometric method.  A diagonal matrix returns its diagonal as stored
 * (unsorted); otherwise eig[0] >= eig[1] >= eig[2]. */
void oracle_eig_3x3(const double* a, double* eig) {
  const double a00 = a[0], a10 = a[1], a20 = a[2], a01 = a[3], a11 = a[4], a21 = a[5];
  const double a02 = a[6], a12 = a[7], a22 = a[8];
  const double pi = M_PI;
  double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  if (p1 == 0) {
    eig[0] = a00;
    eig[1] = a11;
    eig[2] = a22;
    return;
  }
  double q = (a00 + a11 + a22) / 3.0;
  double p2 = (a00 - q) * (a00 - q) + (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + 2 * p1;
  double p = sqrt(p2 / 6.0);
  double b00 = (1 / p) * (a00 - q), b11 = (1 / p) * (a11 - q), b22 = (1 / p) * (a22 - q);
  double b01 = (1 / p) * a01, b02 = (1 / p) * a02, b10 = (1 / p) * a10;
  double b12 = (1 / p) * a12, b20 = (1 / p) * a20, b21 = (1 / p) * a21;
  double det = b00 * b11 * b22 + b01 * b12 * b20 + b02 * b10 * b21 - b00 * b12 * b21 -
               b01 * b10 * b22 - b02 * b11 * b20;
  double r = det / 2.0;
  double phi;
  if (r <= -1)
    phi = pi / 3.0;
  else if (r >= 1)
    phi = 0.0;
  else
    phi = acos(r) / 3;
  if (g_mut == ORACLE_MUT_EIG_PHASE) phi = -phi;
  eig[0] = q + 2 * p * cos(phi);
  eig[2] = q + 2 * p * cos(phi + (2 * pi / 3));
  eig[1] = 3 * q - eig[0] - eig[2];
}

/* fc3d_AlartCurnier_functions.c:814-829 */
void oracle_rho_spectral_norm(const double* W, double* rho) {
  double eig[3];
  oracle_eig_3x3(W, eig);
  rho[0] = 1.0 / eig[0];
  rho[1] = rho[0];
  rho[2] = rho[0];
}

/* ------------------------------------------------------------------------ */
/* Residual — fc3d_compute_error.c:42-56, 70-122                            */
/* ------------------------------------------------------------------------ */
void oracle_unitary_compute_and_add_error(const double* r, const double* u, double mu,
                                          double* error) {
  double w[3];
  if (g_mut == ORACLE_MUT_ERROR_DROP_MU)
    w[0] = r[0] - u[0];
  else
    w[0] = r[0] - u[0] - mu * sqrt(u[1] * u[1] + u[2] * u[2]);
  w[1] = r[1] - u[1];
  w[2] = r[2] - u[2];
  oracle_projectionOnCone(w, mu);
  w[0] = r[0] - w[0];
  w[1] = r[1] - w[1];
  w[2] = r[2] - w[2];
  *error += w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
}

static double dnrm2(size_t n, const double* x) {
  /* sqrt of the running sum of squares; a vendor BLAS dnrm2 may differ in
   * the last bits (documented as unpinned). */
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += x[i] * x[i];
  return sqrt(s);
}

/* SparseBlockMatrix.c:360-419 */
void oracle_SBM_gemv_3x3(const SparseBlockStructuredMatrix* A, const double* x, double* y) {
  for (size_t row = 0; row + 1 < A->filled1; ++row) {
    for (size_t k = A->index1_data[row]; k < A->index1_data[row + 1]; ++k) {
      size_t col = A->index2_data[k];
      oracle_mvp3x3(A->block[k], &x[3 * col], &y[3 * row]);
    }
  }
}

/* SparseBlockMatrix.c:1431-1490 */
void oracle_SBM_row_prod_no_diag_3x3(const SparseBlockStructuredMatrix* A, size_t row,
                                     const double* x, double* y) {
  for (size_t k = A->index1_data[row]; k < A->index1_data[row + 1]; ++k) {
    size_t col = A->index2_data[k];
    if (col != row) oracle_mvp3x3(A->block[k], &x[3 * col], y);
  }
}

int oracle_fc3d_compute_error(FrictionContactProblem* problem, const double* z, double* w,
                              double tolerance, double norm, double* error) {
  int nc = problem->numberOfContacts;
  int n = 3 * nc;
  memcpy(w, problem->q, sizeof(double) * (size_t)n);
  oracle_SBM_gemv_3x3(problem->M->matrix1, z, w);
  *error = 0.;
  for (int ic = 0; ic < nc; ic++)
    oracle_unitary_compute_and_add_error(z + 3 * ic, w + 3 * ic, problem->mu[ic], error);
  *error = sqrt(*error);
  double norm_r = dnrm2((size_t)n, z);
  double norm_u = dnrm2((size_t)n, w);
  double scale = fmax(norm, fmax(norm_r, norm_u));
  if (fabs(scale) > DBL_EPSILON) *error /= scale;
  return *error > tolerance ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Options with reference defaults                                          */
/* ------------------------------------------------------------------------ */
static SolverOptions* options_alloc(int id, int max_iter, double tol) {
  /* SolverOptions.c:44-94 */
  SolverOptions* o = (SolverOptions*)calloc(1, sizeof(SolverOptions));
  o->solverId = id;
  o->iSize = OPTIONS_PARAM_SIZE;
  o->dSize = OPTIONS_PARAM_SIZE;
  o->iparam = (int*)calloc(OPTIONS_PARAM_SIZE, sizeof(int));
  o->dparam = (double*)calloc(OPTIONS_PARAM_SIZE, sizeof(double));
  o->iparam[SICONOS_IPARAM_MAX_ITER] = max_iter;
  o->dparam[SICONOS_DPARAM_TOL] = tol;
  o->dparam[SICONOS_DPARAM_RESIDU] = INFINITY;
  o->filterOn = true;
  o->isSet = true;
  return o;
}

SolverOptions* oracle_local_options_new(int id) {
  SolverOptions* o;
  if (id == SICONOS_FRICTION_3D_ONECONTACT_NSN) {
    /* onecontact...c:1017-1035 and registration 100 / 1e-14 */
    o = options_alloc(id, 100, 1e-14);
    o->iparam[SICONOS_FRICTION_3D_NSN_RHO_STRATEGY] =
        SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM_COND;
    o->dparam[SICONOS_FRICTION_3D_NSN_RHO] = 1.0;
    o->iparam[SICONOS_FRICTION_3D_NSN_FORMULATION] =
        SICONOS_FRICTION_3D_NSN_FORMULATION_ALARTCURNIER_STD;
    o->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH] = SICONOS_FRICTION_3D_NSN_LINESEARCH_NO;
    o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY] =
        SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_PLI_NSN_LOOP;
    o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP] = 1;
    o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER] = 10;
  } else if (id == SICONOS_FRICTION_3D_ONECONTACT_NSN_GP ||
             id == SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID) {
    /* onecontact...c:1037-1056 and registrations 100 / 1e-14 */
    o = options_alloc(id, 100, 1e-14);
    o->iparam[SICONOS_FRICTION_3D_NSN_RHO_STRATEGY] =
        SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM;
    o->dparam[SICONOS_FRICTION_3D_NSN_RHO] = 1.0;
    o->iparam[SICONOS_FRICTION_3D_NSN_FORMULATION] =
        SICONOS_FRICTION_3D_NSN_FORMULATION_ALARTCURNIER_STD;
    o->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH] =
        SICONOS_FRICTION_3D_NSN_LINESEARCH_GOLDSTEINPRICE;
    o->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH_MAX_ITER] = 10;
    o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY] =
        SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_NSN_AND_PLI_NSN_LOOP;
    o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP] = 1;
    o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER] = 100;
  } else {
    /* OC_PROJ / OC_PROJ_LI: fc3d_projection.c:917-960, 1000 / 1e-14 */
    o = options_alloc(id, 1000, 1e-14);
  }
  return o;
}

void oracle_set_local_solver(SolverOptions* nsgs, int id) {
  if (nsgs->internalSolvers[0]) oracle_options_free(nsgs->internalSolvers[0]);
  nsgs->internalSolvers[0] = oracle_local_options_new(id);
}

SolverOptions* oracle_nsgs_options_new(void) {
  /* fc3d_nsgs.c:919-950, registration 1000 / 1e-4 (:981-987) */
  SolverOptions* o = options_alloc(SICONOS_FRICTION_3D_NSGS, 1000, 1e-4);
  o->iparam[SICONOS_NSGS_ERROR_EVALUATION_TYPE] = SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL;
  o->iparam[SICONOS_FRICTION_3D_IPARAM_INTERNAL_ERROR_STRATEGY] =
      SICONOS_NSGS_INTERNAL_ERROR_STRATEGY_GIVEN_VALUE;
  o->iparam[SICONOS_NSGS_SHUFFLE] = 0;
  o->iparam[SICONOS_NSGS_SHUFFLE_SEED] = 0;
  o->iparam[SICONOS_NSGS_FREEZING_CONTACT] = 0;
  o->iparam[SICONOS_NSGS_FILTER_LOCAL_SOLUTION] = 0;
  o->iparam[SICONOS_NSGS_RELAXATION] = 0;
  o->iparam[SICONOS_NSGS_ERROR_EVALUATION_FREQUENCY] = 0;
  o->dparam[SICONOS_DPARAM_TOL] = 1e-4;
  o->dparam[SICONOS_NSGS_INTERNAL_ERROR_RATIO] = 10.0;
  o->numberOfInternalSolvers = 1;
  o->internalSolvers = (SolverOptions**)calloc(1, sizeof(SolverOptions*));
  o->internalSolvers[0] = oracle_local_options_new(SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID);
  o->iparam[SICONOS_NSGS_PRINTING_STYLE] = 1;
  return o;
}

void oracle_options_free(SolverOptions* o) {
  if (!o) return;
  for (size_t i = 0; i < o->numberOfInternalSolvers; i++) oracle_options_free(o->internalSolvers[i]);
  free(o->internalSolvers);
  free(o->iparam);
  free(o->dparam);
  free(o->dWork);
  free(o->iWork);
  free(o);
}

/* ------------------------------------------------------------------------ */
/* Local problem and local solvers                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  const double* W; /* 9, column-major */
  double q[3];
  double mu;
} local_problem;

typedef void (*ac_function)(const double*, const double*, double, const double*, double*,
                            double*, double*);

static ac_function pick_function(const SolverOptions* o) {
  return o->iparam[SICONOS_FRICTION_3D_NSN_FORMULATION] ==
                 SICONOS_FRICTION_3D_NSN_FORMULATION_JEANMOREAU_STD
             ? oracle_computeAlartCurnierJeanMoreau
             : oracle_computeAlartCurnierSTD;
}

static void local_velocity(const local_problem* lp, const double* r, double* v) {
  /* cpy3(q, v); mvp3x3(W, r, v) */
  v[0] = lp->q[0]; v[1] = lp->q[1]; v[2] = lp->q[2];
  oracle_mvp3x3(lp->W, r, v);
}

static void local_velocity_inline(const local_problem* lp, const double* r, double* v) {
  /* order of fc3d_projection.c:430-432 and onecontact...c:624-626:
   * ((W_i0 r0 + q_i) + W_i1 r1) + W_i2 r2 */
  for (int i = 0; i < 3; i++)
    v[i] = lp->W[i + 0 * 3] * r[0] + lp->q[i] + lp->W[i + 1 * 3] * r[1] + lp->W[i + 2 * 3] * r[2];
}

/* onecontact...c:856-878 */
static double compute_local_error(const local_problem* lp, const double* r) {
  double nq = dnrm2(3, lp->q);
  double v[3];
  local_velocity(lp, r, v);
  double e = 0.0;
  oracle_unitary_compute_and_add_error(r, v, lp->mu, &e);
  e = sqrt(e);
  if (fabs(nq) > DBL_EPSILON) e /= nq;
  return e;
}

/* onecontact...c:562-708 */
static int line_search_gp(const local_problem* lp, ac_function Fn, double* t_opt,
                          const double* r, const double* dR, const double* rho, int ls_max,
                          double* F, const double* A, const double* B, double* v) {
  double alpha = *t_opt, inf = 1e20, amin = 0.0, amax = inf, m1 = 0.1, m2 = 0.9;
  double q0 = 0.5 * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
  double AWpB[9], tmp[3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      AWpB[i + 3 * j] = 0.0;
      for (int k = 0; k < 3; k++) AWpB[i + 3 * j] += A[i + 3 * k] * lp->W[k + 3 * j];
      AWpB[i + 3 * j] += B[i + 3 * j];
    }
  for (int i = 0; i < 3; i++) {
    tmp[i] = 0.0;
    for (int j = 0; j < 3; j++) tmp[i] += AWpB[i + 3 * j] * dR[j];
  }
  double dqdt0 = 0.0;
  for (int i = 0; i < 3; i++) dqdt0 += F[i] * tmp[i];
  for (int it = 0; it < ls_max; it++) {
    g_cnt[4]++;
    for (int i = 0; i < 3; i++) tmp[i] = r[i] + alpha * dR[i];
    local_velocity_inline(lp, tmp, v);
    Fn(tmp, v, lp->mu, rho, F, NULL, NULL);
    double q = 0.5 * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
    double slope = (q - q0) / alpha;
    int C1 = slope >= m2 * dqdt0;
    int C2 = slope <= m1 * dqdt0;
    if (C1 && C2) { *t_opt = alpha; return 0; }
    else if (!C1) amin = alpha;
    else amax = alpha;
    if (alpha < inf) alpha = 0.5 * (amin + amax);
    else alpha = 10 * alpha;
  }
  *t_opt = alpha;
  return -1;
}

/* onecontact...c:429-559 (OC_NSN) and 710-854 (OC_NSN_GP) */
static int nsn_newton(const local_problem* lp, double* r, SolverOptions* o, const double* rho,
                      int damped) {
  ac_function Fn = pick_function(o);
  double nq = sqrt(lp->q[0] * lp->q[0] + lp->q[1] * lp->q[1] + lp->q[2] * lp->q[2]);
  double nrel = 1.0;
  if (nq > DBL_EPSILON) nrel /= nq;
  double dR[3] = {0, 0, 0}, F[3] = {0, 0, 0}, A[9] = {0}, B[9] = {0}, AWpB[9] = {0}, v[3];
  local_velocity(lp, r, v);
  int itermax = o->iparam[SICONOS_IPARAM_MAX_ITER];
  int inew = 0, info3;
  double t = 1., t_opt = 1., t_init = 1.;
  Fn(r, v, lp->mu, rho, F, NULL, NULL);
  o->dparam[SICONOS_DPARAM_RESIDU] = sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) * nrel;
  if (o->dparam[SICONOS_DPARAM_RESIDU] < o->dparam[SICONOS_DPARAM_TOL]) {
    o->iparam[SICONOS_IPARAM_ITER_DONE] = inew;
    return 0;
  }
  int ls_max = o->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH_MAX_ITER];
  g_cnt[0]++;
  for (inew = 1; inew < itermax; ++inew) {
    g_cnt[1]++;
    Fn(r, v, lp->mu, rho, F, A, B);
    oracle_mm3x3(A, lp->W, AWpB);
    for (int k = 0; k < 9; k++) AWpB[k] += B[k];
    for (int k = 0; k < 9; k++) AWpB[k] *= -1.;
    dR[0] = F[0]; dR[1] = F[1]; dR[2] = F[2];
    info3 = oracle_solve_3x3_gepp(AWpB, dR);
    if (info3) { dR[0] = 0; dR[1] = 0; dR[2] = 0; }
    if (damped) {
      if (!info3) {
        t_opt = t_init;
        line_search_gp(lp, Fn, &t_opt, r, dR, rho, ls_max, F, A, B, v);
        t = t_opt;
      }
      r[0] = r[0] + t * dR[0];
      r[1] = r[1] + t * dR[1];
      r[2] = r[2] + t * dR[2];
    } else {
      r[0] += dR[0]; r[1] += dR[1]; r[2] += dR[2];
    }
    local_velocity(lp, r, v);
    Fn(r, v, lp->mu, rho, F, NULL, NULL);
    o->dparam[SICONOS_DPARAM_RESIDU] = sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) * nrel;
    if (damped) {
      if (info3) break;
      if (o->dparam[SICONOS_DPARAM_RESIDU] < o->dparam[SICONOS_DPARAM_TOL]) {
        o->iparam[SICONOS_IPARAM_ITER_DONE] = inew;
        return 0;
      }
      if (sqrt(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]) < 10 * DBL_EPSILON) return 1;
    } else {
      if (o->dparam[SICONOS_DPARAM_RESIDU] < o->dparam[SICONOS_DPARAM_TOL]) {
        o->iparam[SICONOS_IPARAM_ITER_DONE] = inew;
        return 0;
      }
      if (sqrt(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]) < 10 * DBL_EPSILON) return 1;
      if (info3) break;
    }
  }
  o->iparam[SICONOS_IPARAM_ITER_DONE] = inew;
  return 1;
}

/* fc3d_projection.c:350-514; rho is the persistent per-contact PLI step */
static int pli_solve(const local_problem* lp, double* r, SolverOptions* o, double* rho_state) {
  double rho = *rho_state, rho_k;
  double vel[3] = {0, 0, 0}, vel_k[3], r_k[3];
  double err = 1.0, err_k, a1 = 0.0, a2 = 0.0;
  int it = 0, ls_it, success;
  double tol = o->dparam[SICONOS_DPARAM_TOL];
  double tau = 2.0 / 3.0, tauinv = 3.0 / 2.0, L = 0.9, Lmin = 0.3;
  double mu = lp->mu;
  g_cnt[2]++;
  while ((err > tol) && (it < o->iparam[SICONOS_IPARAM_MAX_ITER])) {
    it++;
    g_cnt[3]++;
    err_k = err;
    r_k[0] = r[0]; r_k[1] = r[1]; r_k[2] = r[2];
    local_velocity_inline(lp, r, vel_k);
    ls_it = 0;
    success = 0;
    rho_k = rho / tau;
    double nUT = sqrt(vel_k[1] * vel_k[1] + vel_k[2] * vel_k[2]);
    while (!success && (ls_it < 10)) {
      rho_k = rho_k * tau;
      r[0] = r_k[0] - rho_k * (vel_k[0] + mu * nUT);
      r[1] = r_k[1] - rho_k * vel_k[1];
      r[2] = r_k[2] - rho_k * vel_k[2];
      oracle_projectionOnCone(r, mu);
      local_velocity_inline(lp, r, vel);
      a1 = sqrt((vel_k[0] - vel[0]) * (vel_k[0] - vel[0]) + (vel_k[1] - vel[1]) * (vel_k[1] - vel[1]) +
                (vel_k[2] - vel[2]) * (vel_k[2] - vel[2]));
      a2 = sqrt((r_k[0] - r[0]) * (r_k[0] - r[0]) + (r_k[1] - r[1]) * (r_k[1] - r[1]) +
                (r_k[2] - r[2]) * (r_k[2] - r[2]));
      success = (rho_k * a1 <= L * a2) ? 1 : 0;
      ls_it++;
    }
    err = 0.0;
    oracle_unitary_compute_and_add_error(r, vel, mu, &err);
    if ((rho_k * a1 < Lmin * a2) && (err < err_k))
      rho = rho_k * tauinv;
    else
      rho = rho_k;
  }
  *rho_state = rho;
  o->dparam[SICONOS_DPARAM_RESIDU] = err;
  return err > tol ? 1 : 0;
}

/* fc3d_projection.c:564-603 */
static int proj_solve(const local_problem* lp, double* r) {
  double an = 1. / lp->W[g_mut == ORACLE_MUT_PROJ_RHO ? 4 : 0];
  double w[3];
  local_velocity_inline(lp, r, w);
  double nUT = sqrt(w[1] * w[1] + w[2] * w[2]);
  if (g_mut == ORACLE_MUT_PROJ_DROP_MU) nUT = 0.0;
  r[0] -= an * (w[0] + lp->mu * nUT);
  r[1] -= an * w[1];
  r[2] -= an * w[2];
  oracle_projectionOnCone(r, lp->mu);
  return 0;
}

/* onecontact...c:880-918 */
static void keep_or_discard(const local_problem* lp, double* r, double* r_bak, SolverOptions* o,
                            double* cur) {
  double e = compute_local_error(lp, r);
  double res = o->dparam[SICONOS_DPARAM_RESIDU];
  if (isnan(res) || isinf(res)) {
    r[0] = r_bak[0]; r[1] = r_bak[1]; r[2] = r_bak[2];
  } else if (e <= o->dparam[SICONOS_DPARAM_TOL] || e <= *cur) {
    *cur = e;
    r_bak[0] = r[0]; r_bak[1] = r[1]; r_bak[2] = r[2];
  } else {
    r[0] = r_bak[0]; r[1] = r_bak[1]; r[2] = r_bak[2];
  }
}

/* onecontact...c:920-1015 */
static int hybrid_solve(const local_problem* lp, double* r, SolverOptions* o, double* state) {
  int info = -1;
  double bak[3] = {r[0], r[1], r[2]};
  int max_loop = o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP];
  int newton_it = o->iparam[SICONOS_IPARAM_MAX_ITER];
  int pli_it = o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER];
  int strategy = o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY];
  double cur = compute_local_error(lp, r);
  if (!(strategy == SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_PLI_NSN_LOOP ||
        strategy == SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_NSN_AND_PLI_NSN_LOOP))
    return NUMERICS_ERR_INVALID_ARGUMENT;
  if (strategy == SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_NSN_AND_PLI_NSN_LOOP) {
    o->iparam[SICONOS_IPARAM_MAX_ITER] = newton_it;
    info = nsn_newton(lp, r, o, state + 1, 1);
    keep_or_discard(lp, r, bak, o, &cur);
    if (cur <= o->dparam[SICONOS_DPARAM_TOL]) {
      o->dparam[SICONOS_DPARAM_RESIDU] = cur;
      return info;
    }
  }
  int loop = 0;
  while (loop < max_loop && cur >= o->dparam[SICONOS_DPARAM_TOL]) {
    loop++;
    o->iparam[SICONOS_IPARAM_MAX_ITER] = pli_it;
    info = pli_solve(lp, r, o, state);
    keep_or_discard(lp, r, bak, o, &cur);
    if (cur < o->dparam[SICONOS_DPARAM_TOL]) {
      o->iparam[SICONOS_IPARAM_MAX_ITER] = newton_it;
      break;
    }
    o->iparam[SICONOS_IPARAM_MAX_ITER] = newton_it;
    info = nsn_newton(lp, r, o, state + 1, 1);
    keep_or_discard(lp, r, bak, o, &cur);
  }
  o->dparam[SICONOS_DPARAM_RESIDU] = cur;
  return info;
}

int oracle_local_state_init(const double* W, SolverOptions* o, double* state) {
  /* onecontact...c:62-187 (AC rho), fc3d_projection.c:329-341 (PLI rho = 1) */
  state[0] = 1.0;
  state[1] = state[2] = state[3] = 0.0;
  int id = o->solverId;
  if (id == SICONOS_FRICTION_3D_ONECONTACT_NSN || id == SICONOS_FRICTION_3D_ONECONTACT_NSN_GP ||
      id == SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID) {
    int s = o->iparam[SICONOS_FRICTION_3D_NSN_RHO_STRATEGY];
    if (s == SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM_COND)
      oracle_rho_split_spectral_norm_cond(W, state + 1);
    else if (s == SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM)
      oracle_rho_split_spectral_norm(W, state + 1);
    else if (s == SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPECTRAL_NORM)
      oracle_rho_spectral_norm(W, state + 1);
    else if (s == SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_CONSTANT)
      state[1] = state[2] = state[3] = o->dparam[SICONOS_FRICTION_3D_NSN_RHO];
    else
      return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  return 0;
}

int oracle_local_solve(const double* W, const double* q, double mu, double* r, SolverOptions* o,
                       double* state) {
  local_problem lp;
  lp.W = W;
  lp.q[0] = q[0]; lp.q[1] = q[1]; lp.q[2] = q[2];
  lp.mu = mu;
  int id = o->solverId;
  if (id == SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnCone) return proj_solve(&lp, r);
  if (id == SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnConeWithLocalIteration)
    return pli_solve(&lp, r, o, state);
  /* onecontact...c:221-290: take-off check, then dispatch */
  if (lp.q[0] > -DBL_EPSILON) {
    r[0] = 0.0; r[1] = 0.0; r[2] = 0.0;
    o->iparam[SICONOS_IPARAM_ITER_DONE] = 0;
    o->dparam[SICONOS_DPARAM_RESIDU] = 0.0;
    return 0;
  }
  g_cnt[5]++;
  if (id == SICONOS_FRICTION_3D_ONECONTACT_NSN) return nsn_newton(&lp, r, o, state + 1, 0);
  if (id == SICONOS_FRICTION_3D_ONECONTACT_NSN_GP) return nsn_newton(&lp, r, o, state + 1, 1);
  return hybrid_solve(&lp, r, o, state);
}

/* ------------------------------------------------------------------------ */
/* Tolerance manager — tolerance_manager.h:42-241                           */
/* ------------------------------------------------------------------------ */
typedef struct {
  double user_tol, min_tol, working_tol, local_orig, local_cur;
} tol_manager;

static void tm_init(tol_manager* tm, double user_tol, SolverOptions* lo) {
  tm->user_tol = user_tol;
  tm->working_tol = user_tol;
  tm->min_tol = DBL_EPSILON * 1e-6;
  tm->local_orig = lo ? lo->dparam[SICONOS_DPARAM_TOL] : 0.0;
  tm->local_cur = tm->local_orig;
}

static int tm_check(tol_manager* tm, SolverOptions* lo, double full, double incr) {
  if (full <= tm->user_tol) return 0;
  if (incr < 1e-15 || incr < DBL_EPSILON * 100.0) {
    if (!lo) return 1;
    double nt = fmax(tm->local_cur / 100.0, tm->min_tol);
    lo->dparam[SICONOS_DPARAM_TOL] = nt;
    tm->local_cur = nt;
    return 1;
  }
  if (full > 0.0 && incr > 0.0) {
    double ratio = incr / full;
    double nt = ratio * tm->user_tol;
    double umin = tm->user_tol * 1e-6;
    double mn = umin > tm->min_tol ? umin : tm->min_tol;
    if (nt < mn) nt = mn;
    if (nt > 0.0 && nt < tm->working_tol) tm->working_tol = nt;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* fc3d_nsgs — fc3d_nsgs.c:560-917 (general loop, contact freezing of
 * fc3d_nsgs.c:719-800 included; libc rand() shuffles are refused)          */
/* ------------------------------------------------------------------------ */
void oracle_fc3d_nsgs(FrictionContactProblem* problem, double* reaction, double* velocity,
                      int* info, SolverOptions* options, const int* order) {
  size_t nc = (size_t)problem->numberOfContacts;
  int itermax = options->iparam[SICONOS_IPARAM_MAX_ITER];
  double norm_q = dnrm2(3 * nc, problem->q);
  double omega = options->dparam[SICONOS_NSGS_RELAXATION_VALUE];
  double norm_r = 1e24;
  if (options->numberOfInternalSolvers < 1) { *info = NUMERICS_ERR_INVALID_ARGUMENT; return; }
  SolverOptions* lo = options->internalSolvers[0];
  if (options->iparam[SICONOS_NSGS_SHUFFLE] != 0) {
    *info = NUMERICS_ERR_NOT_IMPLEMENTED;
    return;
  }
  const int nfreeze = options->iparam[SICONOS_NSGS_FREEZING_CONTACT];
  int ee = options->iparam[SICONOS_NSGS_ERROR_EVALUATION_TYPE];
  if (ee < 0 || ee > 3) { *info = NUMERICS_ERR_INVALID_ARGUMENT; return; }
  SparseBlockStructuredMatrix* M = problem->M->matrix1;
  if (problem->M->storageType != NM_SPARSE_BLOCK || !M) { *info = NUMERICS_ERR_NOT_IMPLEMENTED; return; }

  tol_manager tm0;
  tm_init(&tm0, options->dparam[SICONOS_DPARAM_TOL], lo);
  double tolerance = tm0.working_tol;

  /* per-contact local solver state (the reference's dWork) */
  double* state = (double*)malloc(sizeof(double) * 4 * nc);
  double** diag = (double**)malloc(sizeof(double*) * nc);
  for (size_t c = 0; c < nc; c++) {
    diag[c] = NULL;
    for (size_t k = M->index1_data[c]; k < M->index1_data[c + 1]; k++)
      if (M->index2_data[k] == c) { diag[c] = M->block[k]; break; }
    int st = oracle_local_state_init(diag[c], lo, state + 4 * c);
    if (st) { *info = st; free(state); free(diag); return; }
  }

  /* fc3d_nsgs.c:302-313 and :712: freeze counters and the last light-error
   * term of each contact (the term a frozen contact adds again) */
  unsigned* frz = nfreeze > 0 ? (unsigned*)calloc(nc, sizeof(unsigned)) : NULL;
  double* le2 = (double*)calloc(nc, sizeof(double));
  int iter = 0, notconv = 1;
  double incr = INFINITY, full = INFINITY;
  double lr[3];
  while (iter < itermax && notconv > 0) {
    ++iter;
    double light_sum = 0.0;
    /* fc3d_nsgs.c:721-735 */
    const double ftmp = (double)nc * (double)nc / 1000.;
    const double crit1 = tolerance * tolerance / ftmp;
    const double crit2 = norm_r * norm_r / ftmp;
    if (nfreeze > 0) {
      unsigned nfrozen = 0;
      for (size_t c = 0; c < nc; c++)
        if (frz[c] > 0) nfrozen++;
      if (nfrozen >= nc - 1)
        for (size_t c = 0; c < nc; c++) frz[c] = 0;
    }
    for (size_t i = 0; i < nc; ++i) {
      size_t c = order ? (size_t)order[i] : i;
      if (nfreeze > 0 && frz[c] > 0) { /* fc3d_nsgs.c:749-756 */
        frz[c] -= 1;
        if (g_mut != ORACLE_MUT_FREEZE_NO_STALE) light_sum += le2[c];
        continue;
      }
      /* fc3d_nsgs_update: q_loc = q_c + sum_{j != c} W_cj r_j */
      double ql[3] = {problem->q[3 * c], problem->q[3 * c + 1], problem->q[3 * c + 2]};
      oracle_SBM_row_prod_no_diag_3x3(M, c, reaction, ql);
      lo->iparam[SICONOS_FRICTION_3D_CURRENT_CONTACT_NUMBER] = (int)c;
      lr[0] = reaction[3 * c]; lr[1] = reaction[3 * c + 1]; lr[2] = reaction[3 * c + 2];
      oracle_local_solve(diag[c], ql, problem->mu[c], lr, lo, state + 4 * c);
      if (options->iparam[SICONOS_NSGS_RELAXATION] == 1)
        for (int k = 0; k < 3; k++) lr[k] = omega * lr[k] + (1.0 - omega) * reaction[3 * c + k];
      double d0 = reaction[3 * c] - lr[0], d1 = reaction[3 * c + 1] - lr[1],
             d2 = reaction[3 * c + 2] - lr[2];
      le2[c] = d0 * d0 + d1 * d1 + d2 * d2;
      light_sum += le2[c];
      if (nfreeze > 0) { /* fc3d_nsgs.c:776-800 */
        double sq = lr[0] * lr[0] + lr[1] * lr[1] + lr[2] * lr[2];
        int relative = le2[c] <= crit1 * sq;
        int small = sq <= crit2;
        if ((relative || small) && iter >= 10) frz[c] = (unsigned)nfreeze;
      }
      if (options->iparam[SICONOS_NSGS_FILTER_LOCAL_SOLUTION] == 1) {
        double res = lo->dparam[SICONOS_DPARAM_RESIDU];
        if (!(isnan(res) || isinf(res) || res > 1.0)) memcpy(&reaction[3 * c], lr, sizeof lr);
      } else {
        memcpy(&reaction[3 * c], lr, sizeof lr);
      }
    }
    if (ee == SICONOS_NSGS_ERROR_EVALUATION_LIGHT) {
      incr = sqrt(light_sum);
      norm_r = dnrm2(3 * nc, reaction);
      if (fabs(norm_r) > DBL_EPSILON) incr /= norm_r;
      notconv = incr < tolerance ? 0 : 1;
    } else if (ee == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
      incr = sqrt(light_sum);
      norm_r = dnrm2(3 * nc, reaction);
      if (fabs(norm_r) > DBL_EPSILON) incr /= norm_r;
      /* determine_convergence_with_full_final: a fresh manager each call,
       * fc3d_nsgs.c:537-558 */
      if (incr >= tolerance) {
        notconv = 1;
      } else {
        tol_manager tm;
        tm_init(&tm, options->dparam[SICONOS_DPARAM_TOL], lo);
        tm.working_tol = tolerance;
        oracle_fc3d_compute_error(problem, reaction, velocity, options->dparam[SICONOS_DPARAM_TOL],
                                  norm_q, &full);
        notconv = tm_check(&tm, lo, full, incr);
        tolerance = tm.working_tol;
      }
      if (!(tolerance > 0.0)) break;
    } else {
      /* FULL / ADAPTIVE: calculateFullErrorAdaptiveInterval (:391-417) */
      int freq = options->iparam[SICONOS_NSGS_ERROR_EVALUATION_FREQUENCY];
      full = 1e24;
      if (freq > 0) {
        if (iter % freq == 0) {
          oracle_fc3d_compute_error(problem, reaction, velocity, tolerance, norm_q, &full);
          if (full > tolerance && ee == SICONOS_NSGS_ERROR_EVALUATION_ADAPTIVE)
            options->iparam[SICONOS_NSGS_ERROR_EVALUATION_FREQUENCY] *= 2;
        }
      } else {
        oracle_fc3d_compute_error(problem, reaction, velocity, tolerance, norm_q, &full);
      }
      incr = full;
      notconv = full < tolerance ? 0 : 1;
    }
    if (options->callback && options->callback->collectStatsIteration)
      options->callback->collectStatsIteration(options->callback->env, 3 * nc, reaction, velocity,
                                               incr, NULL);
  }
  if (ee == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    oracle_fc3d_compute_error(problem, reaction, velocity, tolerance, norm_q, &full);
    notconv = full < options->dparam[SICONOS_DPARAM_TOL] ? 0 : 1;
  }
  *info = notconv;
  if (iter == itermax && notconv > 0) *info = NUMERICS_ERR_MAX_ITER;
  options->dparam[SICONOS_DPARAM_RESIDU] = full;
  options->iparam[SICONOS_IPARAM_ITER_DONE] = iter;
  if (lo) lo->dparam[SICONOS_DPARAM_TOL] = tm0.local_orig;
  free(state);
  free(diag);
  free(frz);
  free(le2);
}

/* fc3d_driver.c:51-145 */
int oracle_fc3d_driver(FrictionContactProblem* problem, double* reaction, double* velocity,
                       SolverOptions* options, const int* order) {
  if (!problem || !reaction || !velocity || !options) return NUMERICS_ERR_NULL_POINTER;
  if (!options->iparam || !options->dparam) return NUMERICS_ERR_INVALID_OPTION;
  if (!problem->M || !problem->q || !problem->mu) return NUMERICS_ERR_NULL_POINTER;
  if (problem->numberOfContacts <= 0) return NUMERICS_ERR_INVALID_ARGUMENT;
  if (problem->dimension != 3) return NUMERICS_ERR_INVALID_ARGUMENT;
  options->iparam[SICONOS_IPARAM_ITER_DONE] = 0;
  options->dparam[SICONOS_DPARAM_RESIDU] = 0.0;
  int nc = problem->numberOfContacts, trivial = 1;
  for (int i = 0; i < nc; i++)
    if (problem->q[3 * i] < -DBL_EPSILON) { trivial = 0; break; }
  if (trivial) {
    for (int i = 0; i < 3 * nc; i++) { velocity[i] = problem->q[i]; reaction[i] = 0.; }
    return NUMERICS_OK;
  }
  if (options->solverId != SICONOS_FRICTION_3D_NSGS) return NUMERICS_ERR_INVALID_SOLVER;
  int info = 0;
  oracle_fc3d_nsgs(problem, reaction, velocity, &info, options, order);
  return info;
}

/* ------------------------------------------------------------------------ */
/* Greedy colouring (sequential definition)                                 */
/* ------------------------------------------------------------------------ */
int oracle_greedy_colouring(int nc, const int64_t* row_ptr, const int64_t* col_idx, int* colour) {
  int ncol = 0;
  int cap = 64;
  int* mark = (int*)malloc(sizeof(int) * (size_t)cap);
  for (int k = 0; k < cap; k++) mark[k] = -1;
  for (int i = 0; i < nc; i++) {
    int first = 1;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      int64_t j = col_idx[k];
      if (j >= i) continue;
      if (g_mut == ORACLE_MUT_COLOUR_OFF_BY_ONE && first) { first = 0; continue; }
      first = 0;
      int cj = colour[j];
      if (cj >= cap) {
        int nc2 = cap;
        while (cj >= nc2) nc2 *= 2;
        mark = (int*)realloc(mark, sizeof(int) * (size_t)nc2);
        for (int t = cap; t < nc2; t++) mark[t] = -1;
        cap = nc2;
      }
      mark[cj] = i;
    }
    int c = 0;
    while (c < cap && mark[c] == i) c++;
    if (c >= cap) {
      mark = (int*)realloc(mark, sizeof(int) * (size_t)(2 * cap));
      for (int t = cap; t < 2 * cap; t++) mark[t] = -1;
      cap *= 2;
    }
    colour[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(mark);
  return ncol;
}

/* ------------------------------------------------------------------------ */
/* Helpers                                                                  */
/* ------------------------------------------------------------------------ */
FrictionContactProblem* oracle_problem_from_csr(int nc, const int64_t* row_ptr,
                                                const int64_t* col_idx, const double* blocks,
                                                const double* q, const double* mu) {
  SparseBlockStructuredMatrix* S = (SparseBlockStructuredMatrix*)calloc(1, sizeof *S);
  size_t nb = (size_t)row_ptr[nc];
  S->nbblocks = nb;
  S->blocknumber0 = S->blocknumber1 = (size_t)nc;
  S->blocksize0 = (size_t*)malloc(sizeof(size_t) * (size_t)nc);
  S->blocksize1 = (size_t*)malloc(sizeof(size_t) * (size_t)nc);
  for (int i = 0; i < nc; i++) S->blocksize0[i] = S->blocksize1[i] = 3 * (size_t)(i + 1);
  S->filled1 = (size_t)nc + 1;
  S->filled2 = nb;
  S->index1_data = (size_t*)malloc(sizeof(size_t) * ((size_t)nc + 1));
  S->index2_data = (size_t*)malloc(sizeof(size_t) * (nb ? nb : 1));
  for (int i = 0; i <= nc; i++) S->index1_data[i] = (size_t)row_ptr[i];
  for (size_t k = 0; k < nb; k++) S->index2_data[k] = (size_t)col_idx[k];
  double* store = (double*)malloc(sizeof(double) * 9 * (nb ? nb : 1));
  memcpy(store, blocks, sizeof(double) * 9 * nb);
  S->block = (double**)malloc(sizeof(double*) * (nb ? nb : 1));
  for (size_t k = 0; k < nb; k++) S->block[k] = store + 9 * k;
  NumericsMatrix* M = (NumericsMatrix*)calloc(1, sizeof *M);
  M->storageType = NM_SPARSE_BLOCK;
  M->size0 = M->size1 = 3 * nc;
  M->matrix1 = S;
  M->destructible = M;
  FrictionContactProblem* p = (FrictionContactProblem*)calloc(1, sizeof *p);
  p->dimension = 3;
  p->numberOfContacts = nc;
  p->M = M;
  p->q = (double*)malloc(sizeof(double) * 3 * (size_t)nc);
  memcpy(p->q, q, sizeof(double) * 3 * (size_t)nc);
  p->mu = (double*)malloc(sizeof(double) * (size_t)nc);
  memcpy(p->mu, mu, sizeof(double) * (size_t)nc);
  return p;
}

void oracle_problem_free(FrictionContactProblem* p) {
  if (!p) return;
  SparseBlockStructuredMatrix* S = p->M->matrix1;
  if (S) {
    if (S->nbblocks) free(S->block[0]);
    free(S->block);
    free(S->blocksize0);
    free(S->blocksize1);
    free(S->index1_data);
    free(S->index2_data);
    free(S);
  }
  free(p->M);
  free(p->q);
  free(p->mu);
  free(p);
}

#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_solve_many(int nprob, const int* nc, const int64_t* row_ptr_off,
                      const int64_t* block_off, const int64_t* row_ptr, const int64_t* col_idx,
                      const double* blocks, const double* q, const double* mu, double* reaction,
                      double tol, int max_iter, int nthreads, int* iters, double* residuals) {
  int ok = 0;
  int64_t* q_off = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nprob + 1));
  q_off[0] = 0;
  for (int p = 0; p < nprob; p++) q_off[p + 1] = q_off[p] + nc[p];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : ok)
#endif
  for (int p = 0; p < nprob; p++) {
    const int64_t* rp = row_ptr + row_ptr_off[p];
    int64_t b0 = block_off[p];
    int n = nc[p];
    /* local row pointers are relative to the problem's first block */
    int64_t* lrp = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    for (int i = 0; i <= n; i++) lrp[i] = rp[i] - rp[0];
    FrictionContactProblem* pb = oracle_problem_from_csr(n, lrp, col_idx + b0, blocks + 9 * b0,
                                                         q + 3 * q_off[p], mu + q_off[p]);
    SolverOptions* o = oracle_nsgs_options_new();
    o->dparam[SICONOS_DPARAM_TOL] = tol;
    o->iparam[SICONOS_IPARAM_MAX_ITER] = max_iter;
    double* u = (double*)calloc(3 * (size_t)n, sizeof(double));
    double* r = reaction + 3 * q_off[p];
    int info = oracle_fc3d_driver(pb, r, u, o, NULL);
    if (iters) iters[p] = o->iparam[SICONOS_IPARAM_ITER_DONE];
    if (residuals) residuals[p] = o->dparam[SICONOS_DPARAM_RESIDU];
    if (info == 0) ok += 1;
    free(u);
    free(lrp);
    oracle_options_free(o);
    oracle_problem_free(pb);
  }
  (void)nthreads;
  free(q_off);
  return ok;
}
