// fc3d_device.hpp — per-contact device math of the fc3d NSGS path (gfx950).
//
// Every function restates one reference function; the operation order is the
// reference's, so that a build with -ffp-contract=off gives the same FP64
// results as the reference C.  T is the sweep precision (double or float).
// The 3x3 Newton factorisation runs in double for T = double and, for
// T = float, in float unless the matrix is near-degenerate.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

namespace fc3d {

template <typename T> struct num;
template <> struct num<double> { static constexpr double eps = DBL_EPSILON; };
template <> struct num<float> { static constexpr float eps = FLT_EPSILON; };

// Quotients and square roots of the local solvers: IEEE in double (the
// reference's results bit for bit); in float the hardware's reciprocal and
// square root (v_rcp_f32, v_sqrt_f32, about 1 ulp), two instructions and one
// in place of the correctly rounded sequences of about ten, which sit on the
// dependent chain of every FP32 contact step.  The FP32 sweeps are refined
// and judged in FP64, so their last bit is not part of any result.
template <typename T>
__device__ __forceinline__ T qdiv(T a, T b) { return a / b; }
template <>
__device__ __forceinline__ float qdiv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
template <typename T>
__device__ __forceinline__ T qsqrt(T x) { return sqrt(x); }
template <>
__device__ __forceinline__ float qsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// op3x3.h:342-363 — r += a v, column by column
template <typename T>
__device__ __forceinline__ void mvp3x3(const T* a, const T* v, T* r) {
#pragma unroll
  for (int j = 0; j < 3; j++) {
    r[0] += a[3 * j + 0] * v[j];
    r[1] += a[3 * j + 1] * v[j];
    r[2] += a[3 * j + 2] * v[j];
  }
}

// op3x3.h:449-465 — c = a b
template <typename T>
__device__ __forceinline__ void mm3x3(const T* a, const T* b, T* c) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      c[i + 3 * j] = a[i + 0] * b[0 + 3 * j] + a[i + 3] * b[1 + 3 * j] + a[i + 6] * b[2 + 3 * j];
}

// op3x3.h:868-981 — partial-pivot 3x3 solve, 0 on success.  The FP64 sweeps
// call it in double, as the reference does; the FP32 sweeps call it in float
// first (see nsn_newton).  The pivot test keeps the reference's absolute
// DBL_EPSILON threshold in both.  With `minpiv`, the smallest pivot magnitude
// used is returned there.
template <typename G>
__device__ __forceinline__ int solve_3x3_gepp(const G* a, G* b, G* minpiv = nullptr) {
  G lp0, lp1, lp2, lm1, lm2, ln1, ln2, bl, bm, bn, f1, f2, s0, s1, s2, alp0;
  const G a0 = a[0], a1 = a[1], a2 = a[2];
  const G aa0 = fabs(a0), aa1 = fabs(a1), aa2 = fabs(a2);
  const int p1 = aa0 >= aa1 ? (aa0 >= aa2 ? 0 : 2) : (aa1 >= aa2 ? 1 : 2);
  if (p1 == 0) {
    f1 = qdiv<G>(a1, a0); f2 = qdiv<G>(a2, a0); lp0 = a0; alp0 = fabs(a0); lp1 = a[3]; lp2 = a[6];
    lm1 = a[4] - f1 * lp1; lm2 = a[7] - f1 * lp2; ln1 = a[5] - f2 * lp1; ln2 = a[8] - f2 * lp2;
    bl = b[0]; bm = b[1] - f1 * bl; bn = b[2] - f2 * bl;
  } else if (p1 == 1) {
    f1 = qdiv<G>(a0, a1); f2 = qdiv<G>(a2, a1); lp0 = a1; alp0 = fabs(a1); lp1 = a[4]; lp2 = a[7];
    lm1 = a[3] - f1 * lp1; lm2 = a[6] - f1 * lp2; ln1 = a[5] - f2 * lp1; ln2 = a[8] - f2 * lp2;
    bl = b[1]; bm = b[0] - f1 * bl; bn = b[2] - f2 * bl;
  } else {
    f1 = qdiv<G>(a0, a2); f2 = qdiv<G>(a1, a2); lp0 = a2; alp0 = fabs(a2); lp1 = a[5]; lp2 = a[8];
    lm1 = a[3] - f1 * lp1; lm2 = a[6] - f1 * lp2; ln1 = a[4] - f2 * lp1; ln2 = a[7] - f2 * lp2;
    bl = b[2]; bm = b[0] - f1 * bl; bn = b[1] - f2 * bl;
  }
  if (alp0 <= G(DBL_EPSILON)) return 1;
  const G alm1 = fabs(lm1), aln1 = fabs(ln1);
  G p2;
  if (alm1 >= aln1) {
    if (alm1 < G(DBL_EPSILON)) return 1;
    f1 = qdiv<G>(ln1, lm1);
    p2 = ln2 - f1 * lm2;
    s2 = qdiv<G>(bn - f1 * bm, p2);
    s1 = qdiv<G>(bm - lm2 * s2, lm1);
  } else {
    if (aln1 < G(DBL_EPSILON)) return 1;
    f1 = qdiv<G>(lm1, ln1);
    p2 = lm2 - f1 * ln2;
    s2 = qdiv<G>(bm - f1 * bn, p2);
    s1 = qdiv<G>(bn - ln2 * s2, ln1);
  }
  if (minpiv) *minpiv = fmin(fmin(alp0, fmax(alm1, aln1)), fabs(p2));
  s0 = qdiv<G>(bl - s1 * lp1 - s2 * lp2, lp0);
  b[0] = s0; b[1] = s1; b[2] = s2;
  return 0;
}

// The FP32 Newton step's 3x3 solve (column-major a), by the adjugate: x =
// adj(a) b / det(a), with no data-dependent branch, so that a wave's lanes
// do not each take their own pivoting path.  Returns 1 without a solution
// when det is not finite or |det| <= 2^-24 max|a_ij|^3 (a near-singular
// matrix, which the caller solves in FP64); 0 otherwise.
__device__ __forceinline__ int solve_3x3_adjugate(const float* a, float* b, float scale) {
  const float c00 = a[4] * a[8] - a[7] * a[5], c01 = a[7] * a[2] - a[1] * a[8],
              c02 = a[1] * a[5] - a[4] * a[2];
  const float c10 = a[6] * a[5] - a[3] * a[8], c11 = a[0] * a[8] - a[6] * a[2],
              c12 = a[3] * a[2] - a[0] * a[5];
  const float c20 = a[3] * a[7] - a[6] * a[4], c21 = a[6] * a[1] - a[0] * a[7],
              c22 = a[0] * a[4] - a[3] * a[1];
  const float det = a[0] * c00 + a[3] * c01 + a[6] * c02;
  if (!(fabsf(det) > 0x1p-24f * scale * scale * scale) || !isfinite(det)) return 1;
  const float id = __builtin_amdgcn_rcpf(det);
  const float b0 = b[0], b1 = b[1], b2 = b[2];
  b[0] = (c00 * b0 + c10 * b1 + c20 * b2) * id;
  b[1] = (c01 * b0 + c11 * b1 + c21 * b2) * id;
  b[2] = (c02 * b0 + c12 * b1 + c22 * b2) * id;
  return 0;
}

// projectionOnCone.c:32-50
template <typename T>
__device__ __forceinline__ unsigned projection_on_cone(T* r, T mu) {
  const T nt = qsqrt<T>(r[1] * r[1] + r[2] * r[2]);
  if (mu * nt <= -r[0]) {
    r[0] = T(0); r[1] = T(0); r[2] = T(0);
    return 0u;
  } else if (nt <= mu * r[0]) {
    return 1u;
  } else {
    const T mu2 = mu * mu;
    r[0] = qdiv<T>(mu * nt + r[0], mu2 + T(1));
    r[1] = qdiv<T>(mu * r[0] * r[1], nt);
    r[2] = qdiv<T>(mu * r[0] * r[2], nt);
    return 2u;
  }
}

// fc3d_AlartCurnier_functions.c:506-646 (Alart-Curnier, radius mu*max(0,RVN))
template <typename T, bool AB>
__device__ __forceinline__ void ac_std(const T* R, const T* u, T mu, const T* rho, T* F, T* A,
                                       T* B) {
  const T RhoN = rho[0], RhoT = rho[1];
  const T RVN = R[0] - RhoN * u[0];
  const T RVT = R[1] - RhoT * u[1];
  const T RVS = R[2] - RhoT * u[2];
  const T RV = qsqrt<T>(RVT * RVT + RVS * RVS);
  T Radius;
  if (AB) { A[3] = T(0); A[6] = T(0); A[1] = T(0); A[2] = T(0);
            B[3] = T(0); B[6] = T(0); B[1] = T(0); B[2] = T(0); }
  if (RVN > T(0)) {
    Radius = mu * RVN;
    F[0] = RhoN * u[0];
    if (AB) { A[0] = RhoN; B[0] = T(0); }
  } else {
    Radius = T(0);
    F[0] = R[0];
    if (AB) { A[0] = T(0); B[0] = T(1); }
  }
  if (RV <= Radius) {
    F[1] = RhoT * u[1];
    F[2] = RhoT * u[2];
    if (AB) {
      A[4] = RhoT; A[7] = T(0); A[5] = T(0); A[8] = RhoT;
      B[1] = T(0); B[4] = T(0); B[7] = T(0); B[2] = T(0); B[5] = T(0); B[8] = T(0);
    }
  } else if (RV > Radius) {
    if (Radius > T(0)) {
      const T RV1 = qdiv<T>(T(1), RV);
      F[1] = R[1] - Radius * RVT * RV1;
      F[2] = R[2] - Radius * RVS * RV1;
      if (AB) {
        const T RV3 = RV1 * RV1 * RV1;
        const T GTT = RV1 - RVT * RVT * RV3;
        const T GTS = -RVT * RVS * RV3;
        const T GST = GTS;
        const T GSS = RV1 - RVS * RVS * RV3;
        A[1] = mu * RhoN * RVT * RV1;
        A[2] = mu * RhoN * RVS * RV1;
        A[4] = GTT * RhoT * Radius;
        A[7] = GTS * RhoT * Radius;
        A[5] = GST * RhoT * Radius;
        A[8] = GSS * RhoT * Radius;
        B[1] = -mu * RVT * RV1;
        B[4] = T(1) - GTT * Radius;
        B[7] = -GTS * Radius;
        B[2] = -mu * RVS * RV1;
        B[5] = -GST * Radius;
        B[8] = T(1) - GSS * Radius;
      }
    } else {
      F[1] = R[1];
      F[2] = R[2];
      if (AB) {
        A[4] = T(0); A[7] = T(0); A[5] = T(0); A[8] = T(0);
        B[1] = T(0); B[4] = T(1); B[7] = T(0); B[2] = T(0); B[5] = T(0); B[8] = T(1);
      }
    }
  }
}

// ac_std in two halves, for a Newton loop that evaluates F at each new
// iterate and needs A and B there only if it goes on: the F half keeps the
// intermediates, and the A, B half finishes from them.  The branches and
// the arithmetic are ac_std's.
template <typename T>
struct AcStdPoint {
  T RVT, RVS, RV, Radius, RV1;
  bool normal_active;  // RVN > 0
};

template <typename T>
__device__ __forceinline__ void ac_std_f(const T* R, const T* u, T mu, const T* rho, T* F,
                                         AcStdPoint<T>& p) {
  const T RhoN = rho[0], RhoT = rho[1];
  const T RVN = R[0] - RhoN * u[0];
  p.RVT = R[1] - RhoT * u[1];
  p.RVS = R[2] - RhoT * u[2];
  p.RV = qsqrt<T>(p.RVT * p.RVT + p.RVS * p.RVS);
  p.normal_active = RVN > T(0);
  p.Radius = p.normal_active ? mu * RVN : T(0);
  F[0] = p.normal_active ? RhoN * u[0] : R[0];
  p.RV1 = T(0);
  if (p.RV <= p.Radius) {
    F[1] = RhoT * u[1];
    F[2] = RhoT * u[2];
  } else if (p.RV > p.Radius) {
    if (p.Radius > T(0)) {
      p.RV1 = qdiv<T>(T(1), p.RV);
      F[1] = R[1] - p.Radius * p.RVT * p.RV1;
      F[2] = R[2] - p.Radius * p.RVS * p.RV1;
    } else {
      F[1] = R[1];
      F[2] = R[2];
    }
  }
}

template <typename T>
__device__ __forceinline__ void ac_std_ab(const AcStdPoint<T>& p, T mu, const T* rho, T* A,
                                          T* B) {
  const T RhoN = rho[0], RhoT = rho[1];
  const T RVT = p.RVT, RVS = p.RVS, RV = p.RV, Radius = p.Radius, RV1 = p.RV1;
  A[3] = T(0); A[6] = T(0); A[1] = T(0); A[2] = T(0);
  B[3] = T(0); B[6] = T(0); B[1] = T(0); B[2] = T(0);
  if (p.normal_active) {
    A[0] = RhoN; B[0] = T(0);
  } else {
    A[0] = T(0); B[0] = T(1);
  }
  if (RV <= Radius) {
    A[4] = RhoT; A[7] = T(0); A[5] = T(0); A[8] = RhoT;
    B[1] = T(0); B[4] = T(0); B[7] = T(0); B[2] = T(0); B[5] = T(0); B[8] = T(0);
  } else if (RV > Radius) {
    if (Radius > T(0)) {
      const T RV3 = RV1 * RV1 * RV1;
      const T GTT = RV1 - RVT * RVT * RV3;
      const T GTS = -RVT * RVS * RV3;
      const T GST = GTS;
      const T GSS = RV1 - RVS * RVS * RV3;
      A[1] = mu * RhoN * RVT * RV1;
      A[2] = mu * RhoN * RVS * RV1;
      A[4] = GTT * RhoT * Radius;
      A[7] = GTS * RhoT * Radius;
      A[5] = GST * RhoT * Radius;
      A[8] = GSS * RhoT * Radius;
      B[1] = -mu * RVT * RV1;
      B[4] = T(1) - GTT * Radius;
      B[7] = -GTS * Radius;
      B[2] = -mu * RVS * RV1;
      B[5] = -GST * Radius;
      B[8] = T(1) - GSS * Radius;
    } else {
      A[4] = T(0); A[7] = T(0); A[5] = T(0); A[8] = T(0);
      B[1] = T(0); B[4] = T(1); B[7] = T(0); B[2] = T(0); B[5] = T(0); B[8] = T(1);
    }
  }
}

// fc3d_AlartCurnier_functions.c:649-753 (radius mu*R[0])
template <typename T, bool AB>
__device__ __forceinline__ void ac_jeanmoreau(const T* R, const T* u, T mu, const T* rho, T* F,
                                              T* A, T* B) {
  const T RhoN = rho[0], RhoT = rho[1];
  const T RVN = R[0] - RhoN * u[0];
  const T RVT = R[1] - RhoT * u[1];
  const T RVS = R[2] - RhoT * u[2];
  const T RV = qsqrt<T>(RVT * RVT + RVS * RVS);
  const T Radius = mu * R[0];
  if (RVN >= T(0)) {
    F[0] = RhoN * u[0];
    if (AB) { A[0] = RhoN; B[0] = T(0); }
  } else {
    F[0] = R[0];
    if (AB) { A[0] = T(0); B[0] = T(1); }
  }
  if (RV < Radius || RV < T(1e-20)) {
    F[1] = RhoT * u[1];
    F[2] = RhoT * u[2];
    if (AB) {
      A[4] = RhoT; A[7] = T(0); A[5] = T(0); A[8] = RhoT;
      B[1] = T(0); B[4] = T(0); B[7] = T(0); B[2] = T(0); B[5] = T(0); B[8] = T(0);
    }
  } else {
    const T RV1 = qdiv<T>(T(1), RV);
    F[1] = R[1] - Radius * RVT * RV1;
    F[2] = R[2] - Radius * RVS * RV1;
    if (AB) {
      const T RV3 = RV1 * RV1 * RV1;
      const T GTT = RV1 - RVT * RVT * RV3;
      const T GTS = -RVT * RVS * RV3;
      const T GST = GTS;
      const T GSS = RV1 - RVS * RVS * RV3;
      A[4] = GTT * RhoT * Radius;
      A[7] = GTS * RhoT * Radius;
      A[5] = GST * RhoT * Radius;
      A[8] = GSS * RhoT * Radius;
      B[1] = -mu * RVT * RV1;
      B[4] = T(1) - GTT * Radius;
      B[7] = -GTS * Radius;
      B[2] = -mu * RVS * RV1;
      B[5] = -GST * Radius;
      B[8] = T(1) - GSS * Radius;
    }
  }
}

template <typename T, bool AB>
__device__ __forceinline__ void ac_function(int formulation, const T* R, const T* u, T mu,
                                            const T* rho, T* F, T* A, T* B) {
  if (formulation == 1)
    ac_jeanmoreau<T, AB>(R, u, mu, rho, F, A, B);
  else
    ac_std<T, AB>(R, u, mu, rho, F, A, B);
}

// op3x3.h:1065-1111 — eigenvalues of a symmetric column-major 3x3 matrix by
// the trigonometric method, in the reference's operation order.  A diagonal
// matrix returns its diagonal as stored (not sorted); otherwise eig[0] is the
// largest.  acos and cos are the device library's, so eig may differ from a
// host libm's in the last bit.
template <typename T>
__device__ __forceinline__ void eig_3x3(const T* a, T* eig) {
  const T a00 = a[0], a10 = a[1], a20 = a[2], a01 = a[3], a11 = a[4], a21 = a[5];
  const T a02 = a[6], a12 = a[7], a22 = a[8];
  const T pi = T(3.14159265358979323846);
  const T p1 = a01 * a01 + a02 * a02 + a12 * a12;
  if (p1 == T(0)) {
    eig[0] = a00;
    eig[1] = a11;
    eig[2] = a22;
    return;
  }
  const T q = (a00 + a11 + a22) / T(3);
  const T p2 = (a00 - q) * (a00 - q) + (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + T(2) * p1;
  const T p = sqrt(p2 / T(6));
  const T ip = T(1) / p;
  const T b00 = ip * (a00 - q), b11 = ip * (a11 - q), b22 = ip * (a22 - q);
  const T b01 = ip * a01, b02 = ip * a02, b10 = ip * a10, b12 = ip * a12, b20 = ip * a20;
  const T b21 = ip * a21;
  const T det = b00 * b11 * b22 + b01 * b12 * b20 + b02 * b10 * b21 - b00 * b12 * b21 -
                b01 * b10 * b22 - b02 * b11 * b20;
  const T r = det / T(2);
  T phi;
  if (r <= T(-1))
    phi = pi / T(3);
  else if (r >= T(1))
    phi = T(0);
  else
    phi = acos(r) / T(3);
  eig[0] = q + T(2) * p * cos(phi);
  eig[2] = q + T(2) * p * cos(phi + (T(2) * pi / T(3)));
  eig[1] = T(3) * q - eig[0] - eig[2];
}

// fc3d_AlartCurnier_functions.c:784-812 (quirk sw - W21 + W12 kept),
// 755-782 (_cond variant) and 814-829 (spectral norm: 1/eig[0] on all three);
// strategy ids of FrictionContact_options.h:329-340.
template <typename T>
__device__ __forceinline__ void compute_rho(int strategy, T rho_constant, const T* W, T* rho) {
  if (strategy == 0) {
    rho[0] = rho[1] = rho[2] = rho_constant;
    return;
  }
  if (strategy == 1) {
    T eig[3];
    eig_3x3<T>(W, eig);
    rho[0] = qdiv<T>(T(1), eig[0]);
    rho[1] = rho[0];
    rho[2] = rho[0];
    return;
  }
  const T sw = W[1 + 1 * 3] + W[2 + 2 * 3];
  T dw = sw * sw - T(4) * (sw - W[2 + 1 * 3] + W[1 + 2 * 3]);
  dw = dw > T(0) ? qsqrt<T>(dw) : T(0);
  rho[0] = qdiv<T>(T(1), W[0]);
  if (strategy == 2)
    rho[1] = qdiv<T>(T(2) * (sw - dw), (sw + dw) * (sw + dw));
  else
    rho[1] = qdiv<T>(T(2), sw + dw);
  rho[2] = rho[1];
}

// fc3d_compute_error.c:42-56 — returns the squared local term
template <typename T>
__device__ __forceinline__ T unitary_error(const T* r, const T* u, T mu) {
  T w[3];
  w[0] = r[0] - u[0] - mu * qsqrt<T>(u[1] * u[1] + u[2] * u[2]);
  w[1] = r[1] - u[1];
  w[2] = r[2] - u[2];
  projection_on_cone<T>(w, mu);
  w[0] = r[0] - w[0];
  w[1] = r[1] - w[1];
  w[2] = r[2] - w[2];
  return w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
}

template <typename T>
struct LocalProblem {
  T W[9];
  T q[3];
  T mu;
};

template <typename T>
__device__ __forceinline__ void local_velocity(const LocalProblem<T>& lp, const T* r, T* v) {
  v[0] = lp.q[0]; v[1] = lp.q[1]; v[2] = lp.q[2];
  mvp3x3<T>(lp.W, r, v);
}

// order of fc3d_projection.c:430-432 and onecontact...c:624-626
template <typename T>
__device__ __forceinline__ void local_velocity_inline(const LocalProblem<T>& lp, const T* r,
                                                      T* v) {
#pragma unroll
  for (int i = 0; i < 3; i++)
    v[i] = lp.W[i + 0 * 3] * r[0] + lp.q[i] + lp.W[i + 1 * 3] * r[1] + lp.W[i + 2 * 3] * r[2];
}

// onecontact...c:856-878
template <typename T>
__device__ __forceinline__ T compute_local_error(const LocalProblem<T>& lp, const T* r) {
  const T nq = qsqrt<T>(lp.q[0] * lp.q[0] + lp.q[1] * lp.q[1] + lp.q[2] * lp.q[2]);
  T v[3];
  local_velocity<T>(lp, r, v);
  T e = qsqrt<T>(unitary_error<T>(r, v, lp.mu));
  if (fabs(nq) > num<T>::eps) e = qdiv<T>(e, nq);
  return e;
}

struct LocalParams {
  int solver;          // 565, 551, 550, 553, 552
  int newton_max_iter; // local iparam[0]
  int formulation;
  int ls_max_iter;
  int hybrid_strategy;
  int hybrid_max_loop;
  int hybrid_max_iter;
};

// onecontact...c:562-708
template <typename T>
__device__ void line_search_gp(const LocalProblem<T>& lp, int formulation, T* t_opt, const T* r,
                               const T* dR, const T* rho, int ls_max, T* F, const T* A,
                               const T* B, T* v) {
  T alpha = *t_opt;
  const T inf = T(1e20);
  T amin = T(0), amax = inf;
  const T m1 = T(0.1), m2 = T(0.9);
  const T q0 = T(0.5) * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
  T AWpB[9], tmp[3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < 3; k++) s += A[i + 3 * k] * lp.W[k + 3 * j];
      AWpB[i + 3 * j] = s + B[i + 3 * j];
    }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    T s = T(0);
#pragma unroll
    for (int j = 0; j < 3; j++) s += AWpB[i + 3 * j] * dR[j];
    tmp[i] = s;
  }
  T dqdt0 = T(0);
#pragma unroll
  for (int i = 0; i < 3; i++) dqdt0 += F[i] * tmp[i];
  for (int it = 0; it < ls_max; it++) {
#pragma unroll
    for (int i = 0; i < 3; i++) tmp[i] = r[i] + alpha * dR[i];
    local_velocity_inline<T>(lp, tmp, v);
    ac_function<T, false>(formulation, tmp, v, lp.mu, rho, F, nullptr, nullptr);
    const T q = T(0.5) * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
    const T slope = qdiv<T>(q - q0, alpha);
    const bool C1 = slope >= m2 * dqdt0;
    const bool C2 = slope <= m1 * dqdt0;
    if (C1 && C2) {
      *t_opt = alpha;
      return;
    } else if (!C1) {
      amin = alpha;
    } else {
      amax = alpha;
    }
    if (alpha < inf)
      alpha = T(0.5) * (amin + amax);
    else
      alpha = T(10) * alpha;
  }
  *t_opt = alpha;
}

// The Newton direction of onecontact...c:500-521 (and 786-805): solve
// -(A W + B) dR = F; dR = 0 when the matrix is singular (the returned info).
template <typename T>
__device__ __forceinline__ int newton_direction(const T* A, const T* B, const T* W, const T* F,
                                                T* dR) {
  T AW[9];
  mm3x3<T>(A, W, AW);
  T M[9], d[3];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    T x = AW[k] + B[k];
    x = x * T(-1);
    M[k] = x;
  }
  d[0] = F[0]; d[1] = F[1]; d[2] = F[2];
  int info3;
  if constexpr (sizeof(T) == 8) {
    info3 = solve_3x3_gepp<T>(M, d);
  } else {
    // FP32 adjugate solve, redone in FP64 for a near-degenerate matrix
    T scale = T(0);
#pragma unroll
    for (int k = 0; k < 9; k++) scale = fmax(scale, fabs(M[k]));
    info3 = solve_3x3_adjugate(M, d, scale);
    if (info3) {
      double M64[9], d64[3] = {double(F[0]), double(F[1]), double(F[2])};
#pragma unroll
      for (int k = 0; k < 9; k++) M64[k] = double(M[k]);
      info3 = solve_3x3_gepp<double>(M64, d64);
      d[0] = T(d64[0]); d[1] = T(d64[1]); d[2] = T(d64[2]);
    }
  }
  if (info3) {
    dR[0] = T(0); dR[1] = T(0); dR[2] = T(0);
  } else {
    dR[0] = T(d[0]); dR[1] = T(d[1]); dR[2] = T(d[2]);
  }
  return info3;
}

// nsn_newton's undamped loop (onecontact...c:429-559) for the FP32 sweeps
// with the STD formulation: the same steps, except that the A and B of an
// iteration (the evaluation at :498) are finished from the intermediates of
// the F evaluation that tested the same iterate (:480, :528), instead of a
// second evaluation of the function there.
template <typename T>
__device__ int nsn_newton_std_fp32(const LocalProblem<T>& lp, T* r, const LocalParams& P, T tol,
                                   const T* rho, T* residual) {
  const T nq = qsqrt<T>(lp.q[0] * lp.q[0] + lp.q[1] * lp.q[1] + lp.q[2] * lp.q[2]);
  T nrel = T(1);
  if (nq > num<T>::eps) nrel = qdiv<T>(nrel, nq);
  T dR[3] = {T(0), T(0), T(0)}, F[3], v[3];
  T A[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  T B[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  AcStdPoint<T> pt;
  local_velocity<T>(lp, r, v);
  ac_std_f<T>(r, v, lp.mu, rho, F, pt);
  *residual = qsqrt<T>(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) * nrel;
  if (*residual < tol) return 0;
  for (int inew = 1; inew < P.newton_max_iter; ++inew) {
    ac_std_ab<T>(pt, lp.mu, rho, A, B);
    const int info3 = newton_direction<T>(A, B, lp.W, F, dR);
    r[0] += dR[0]; r[1] += dR[1]; r[2] += dR[2];
    local_velocity<T>(lp, r, v);
    ac_std_f<T>(r, v, lp.mu, rho, F, pt);
    *residual = qsqrt<T>(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) * nrel;
    const T ndR = qsqrt<T>(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]);
    if (*residual < tol) return 0;
    if (ndR < T(10) * num<T>::eps) return 1;
    if (info3) break;
  }
  return 1;
}

// onecontact...c:429-559 (damped = false) and 710-854 (damped = true)
template <typename T>
__device__ int nsn_newton(const LocalProblem<T>& lp, T* r, const LocalParams& P, T tol,
                          const T* rho, bool damped, T* residual) {
  if constexpr (sizeof(T) == 4) {
    if (!damped && P.formulation != 1)
      return nsn_newton_std_fp32<T>(lp, r, P, tol, rho, residual);
  }
  const T nq = qsqrt<T>(lp.q[0] * lp.q[0] + lp.q[1] * lp.q[1] + lp.q[2] * lp.q[2]);
  T nrel = T(1);
  if (nq > num<T>::eps) nrel = qdiv<T>(nrel, nq);
  // A and B start at zero: the JeanMoreau function leaves the coupling
  // entries untouched (fc3d_AlartCurnier_functions.c:649-753)
  T dR[3] = {T(0), T(0), T(0)}, F[3], v[3];
  T A[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  T B[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  local_velocity<T>(lp, r, v);
  T t = T(1);
  ac_function<T, false>(P.formulation, r, v, lp.mu, rho, F, nullptr, nullptr);
  *residual = qsqrt<T>(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) * nrel;
  if (*residual < tol) return 0;
  for (int inew = 1; inew < P.newton_max_iter; ++inew) {
    ac_function<T, true>(P.formulation, r, v, lp.mu, rho, F, A, B);
    T AW[9];
    mm3x3<T>(A, lp.W, AW);
    T M[9], d[3];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      T x = AW[k] + B[k];
      x = x * T(-1);
      M[k] = x;
    }
    d[0] = F[0]; d[1] = F[1]; d[2] = F[2];
    int info3;
    if constexpr (sizeof(T) == 8) {
      info3 = solve_3x3_gepp<T>(M, d);
    } else {
      // FP32 adjugate solve, redone in FP64 for a near-degenerate matrix
      T scale = T(0);
#pragma unroll
      for (int k = 0; k < 9; k++) scale = fmax(scale, fabs(M[k]));
      info3 = solve_3x3_adjugate(M, d, scale);
      if (info3) {
        double M64[9], d64[3] = {double(F[0]), double(F[1]), double(F[2])};
#pragma unroll
        for (int k = 0; k < 9; k++) M64[k] = double(M[k]);
        info3 = solve_3x3_gepp<double>(M64, d64);
        d[0] = T(d64[0]); d[1] = T(d64[1]); d[2] = T(d64[2]);
      }
    }
    if (info3) {
      dR[0] = T(0); dR[1] = T(0); dR[2] = T(0);
    } else {
      dR[0] = T(d[0]); dR[1] = T(d[1]); dR[2] = T(d[2]);
    }
    if (damped) {
      if (!info3) {
        T t_opt = T(1);
        line_search_gp<T>(lp, P.formulation, &t_opt, r, dR, rho, P.ls_max_iter, F, A, B, v);
        t = t_opt;
      }
      r[0] = r[0] + t * dR[0];
      r[1] = r[1] + t * dR[1];
      r[2] = r[2] + t * dR[2];
    } else {
      r[0] += dR[0]; r[1] += dR[1]; r[2] += dR[2];
    }
    local_velocity<T>(lp, r, v);
    ac_function<T, false>(P.formulation, r, v, lp.mu, rho, F, nullptr, nullptr);
    *residual = qsqrt<T>(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) * nrel;
    const T ndR = qsqrt<T>(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]);
    if (damped) {
      if (info3) break;
      if (*residual < tol) return 0;
      if (ndR < T(10) * num<T>::eps) return 1;
    } else {
      if (*residual < tol) return 0;
      if (ndR < T(10) * num<T>::eps) return 1;
      if (info3) break;
    }
  }
  return 1;
}

// fc3d_projection.c:350-514; *rho_state is the persistent per-contact step
template <typename T>
__device__ int pli_solve(const LocalProblem<T>& lp, T* r, int max_iter, T tol, T* rho_state,
                         T* residual) {
  T rho = *rho_state, rho_k = rho;
  T vel[3] = {T(0), T(0), T(0)}, vel_k[3], r_k[3];
  T err = T(1), err_k, a1 = T(0), a2 = T(0);
  int it = 0;
  const T tau = T(2.0 / 3.0), tauinv = T(3.0 / 2.0), L = T(0.9), Lmin = T(0.3);
  const T mu = lp.mu;
  while ((err > tol) && (it < max_iter)) {
    it++;
    err_k = err;
    r_k[0] = r[0]; r_k[1] = r[1]; r_k[2] = r[2];
    local_velocity_inline<T>(lp, r, vel_k);
    int ls_it = 0;
    bool success = false;
    rho_k = qdiv<T>(rho, tau);
    const T nUT = qsqrt<T>(vel_k[1] * vel_k[1] + vel_k[2] * vel_k[2]);
    while (!success && (ls_it < 10)) {
      rho_k = rho_k * tau;
      r[0] = r_k[0] - rho_k * (vel_k[0] + mu * nUT);
      r[1] = r_k[1] - rho_k * vel_k[1];
      r[2] = r_k[2] - rho_k * vel_k[2];
      projection_on_cone<T>(r, mu);
      local_velocity_inline<T>(lp, r, vel);
      a1 = qsqrt<T>((vel_k[0] - vel[0]) * (vel_k[0] - vel[0]) +
                (vel_k[1] - vel[1]) * (vel_k[1] - vel[1]) +
                (vel_k[2] - vel[2]) * (vel_k[2] - vel[2]));
      a2 = qsqrt<T>((r_k[0] - r[0]) * (r_k[0] - r[0]) + (r_k[1] - r[1]) * (r_k[1] - r[1]) +
                (r_k[2] - r[2]) * (r_k[2] - r[2]));
      success = (rho_k * a1 <= L * a2);
      ls_it++;
    }
    err = T(0) + unitary_error<T>(r, vel, mu);
    if ((rho_k * a1 < Lmin * a2) && (err < err_k))
      rho = rho_k * tauinv;
    else
      rho = rho_k;
  }
  *rho_state = rho;
  *residual = err;
  return err > tol ? 1 : 0;
}

// fc3d_projection.c:564-603
template <typename T>
__device__ __forceinline__ int proj_solve(const LocalProblem<T>& lp, T* r) {
  const T an = qdiv<T>(T(1), lp.W[0]);
  T w[3];
  local_velocity_inline<T>(lp, r, w);
  const T nUT = qsqrt<T>(w[1] * w[1] + w[2] * w[2]);
  r[0] -= an * (w[0] + lp.mu * nUT);
  r[1] -= an * w[1];
  r[2] -= an * w[2];
  projection_on_cone<T>(r, lp.mu);
  return 0;
}

// onecontact...c:880-918
template <typename T>
__device__ __forceinline__ void keep_or_discard(const LocalProblem<T>& lp, T* r, T* bak, T tol,
                                                T residual, T* cur) {
  const T e = compute_local_error<T>(lp, r);
  if (isnan(residual) || isinf(residual)) {
    r[0] = bak[0]; r[1] = bak[1]; r[2] = bak[2];
  } else if (e <= tol || e <= *cur) {
    *cur = e;
    bak[0] = r[0]; bak[1] = r[1]; bak[2] = r[2];
  } else {
    r[0] = bak[0]; r[1] = bak[1]; r[2] = bak[2];
  }
}

// onecontact...c:920-1015
template <typename T>
__device__ int hybrid_solve(const LocalProblem<T>& lp, T* r, const LocalParams& P, T tol,
                            T* rho_pli, const T* rho, T* residual) {
  int info = -1;
  T bak[3] = {r[0], r[1], r[2]};
  T cur = compute_local_error<T>(lp, r);
  T res = T(0);
  if (P.hybrid_strategy == 2) {
    info = nsn_newton<T>(lp, r, P, tol, rho, true, &res);
    keep_or_discard<T>(lp, r, bak, tol, res, &cur);
    if (cur <= tol) {
      *residual = cur;
      return info;
    }
  }
  int loop = 0;
  while (loop < P.hybrid_max_loop && cur >= tol) {
    loop++;
    info = pli_solve<T>(lp, r, P.hybrid_max_iter, tol, rho_pli, &res);
    keep_or_discard<T>(lp, r, bak, tol, res, &cur);
    if (cur < tol) break;
    info = nsn_newton<T>(lp, r, P, tol, rho, true, &res);
    keep_or_discard<T>(lp, r, bak, tol, res, &cur);
  }
  *residual = cur;
  return info;
}

// onecontact...c:221-290 (take-off, then dispatch) and the projection solvers.
// state_pli: PLI step; rho: AC parameters; residual: SOLVER_RESIDUAL afterwards.
template <typename T>
__device__ int local_solve(const LocalProblem<T>& lp, T* r, const LocalParams& P, T tol,
                           T* state_pli, const T* rho, T* residual) {
  if (P.solver == 552) {
    // the projection solver never sets SOLVER_RESIDUAL, which keeps the
    // INFINITY given by solver_options_initialize (SolverOptions.c:74)
    *residual = T(INFINITY);
    return proj_solve<T>(lp, r);
  }
  if (P.solver == 553) return pli_solve<T>(lp, r, P.newton_max_iter, tol, state_pli, residual);
  if (lp.q[0] > -num<T>::eps) {
    r[0] = T(0); r[1] = T(0); r[2] = T(0);
    *residual = T(0);
    return 0;
  }
  if (P.solver == 550) return nsn_newton<T>(lp, r, P, tol, rho, false, residual);
  if (P.solver == 551) return nsn_newton<T>(lp, r, P, tol, rho, true, residual);
  return hybrid_solve<T>(lp, r, P, tol, state_pli, rho, residual);
}

}  // namespace fc3d
