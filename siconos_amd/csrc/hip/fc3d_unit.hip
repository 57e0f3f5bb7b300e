// fc3d_unit.hip — per-contact kernels over independent contacts: the
// Alart-Curnier function, the cone projection and one local solve per
// contact.  They expose the device functions of fc3d_device.hpp through the
// C-ABI so each can be checked on its own.
#include <vector>

#include "fc3d_device.hpp"
#include "fc3d_internal.hpp"

namespace fc3d {

__global__ void k_alart_curnier(int n, const double* R, const double* U, const double* mu,
                                const double* rho, int formulation, double* F, double* A,
                                double* B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[3], u[3], rh[3], f[3], a[9], b[9];
  for (int k = 0; k < 3; k++) { r[k] = R[3 * i + k]; u[k] = U[3 * i + k]; rh[k] = rho[3 * i + k]; }
  for (int k = 0; k < 9; k++) { a[k] = 0.0; b[k] = 0.0; }
  ac_function<double, true>(formulation, r, u, mu[i], rh, f, a, b);
  for (int k = 0; k < 3; k++) F[3 * i + k] = f[k];
  for (int k = 0; k < 9; k++) { A[9 * i + k] = a[k]; B[9 * i + k] = b[k]; }
}

__global__ void k_projection(int n, double* r, const double* mu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
  projection_on_cone<double>(x, mu[i]);
  r[3 * i] = x[0]; r[3 * i + 1] = x[1]; r[3 * i + 2] = x[2];
}

struct UnitLocal {
  LocalParams lp;
  double tol;
  int rho_strategy;
  double rho_constant;
};

// T = float: the FP32 sweeps' local solve, on the inputs rounded to float;
// results and state are widened back.
template <typename T>
__global__ void k_local_solve(int n, const double* W, const double* q, const double* mu,
                              double* r, UnitLocal P, double* state, int init_state, int* info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  LocalProblem<T> lp;
  for (int k = 0; k < 9; k++) lp.W[k] = T(W[9 * i + k]);
  for (int k = 0; k < 3; k++) lp.q[k] = T(q[3 * i + k]);
  lp.mu = T(mu[i]);
  T st[4];
  if (init_state) {
    st[0] = T(1);
    compute_rho<T>(P.rho_strategy, T(P.rho_constant), lp.W, st + 1);
  } else {
    for (int k = 0; k < 4; k++) st[k] = T(state[4 * i + k]);
  }
  T x[3] = {T(r[3 * i]), T(r[3 * i + 1]), T(r[3 * i + 2])};
  T res = T(0);
  const int inf = local_solve<T>(lp, x, P.lp, T(P.tol), st, st + 1, &res);
  for (int k = 0; k < 3; k++) r[3 * i + k] = double(x[k]);
  for (int k = 0; k < 4; k++) state[4 * i + k] = double(st[k]);
  info[i] = inf;
}

template <typename T>
static int upload(T** d, const T* h, size_t n) {
  FC3D_HIP_TRY(hipMalloc(d, sizeof(T) * (n ? n : 1)));
  if (n && h) FC3D_HIP_TRY(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace fc3d

using namespace fc3d;

extern "C" {

int fc3d_hip_alart_curnier(int n, const double* R, const double* U, const double* mu,
                           const double* rho, int formulation, double* F, double* A, double* B) {
  if (n < 0) return NUMERICS_ERR_INVALID_ARGUMENT;
  if (n == 0) return 0;
  int rc = use_device(0);
  if (rc) return rc;
  double *dR, *dU, *dmu, *drho, *dF, *dA, *dB;
  if ((rc = upload(&dR, R, 3 * (size_t)n)) || (rc = upload(&dU, U, 3 * (size_t)n)) ||
      (rc = upload(&dmu, mu, (size_t)n)) || (rc = upload(&drho, rho, 3 * (size_t)n)) ||
      (rc = upload<double>(&dF, nullptr, 3 * (size_t)n)) ||
      (rc = upload<double>(&dA, nullptr, 9 * (size_t)n)) ||
      (rc = upload<double>(&dB, nullptr, 9 * (size_t)n)))
    return rc;
  hipLaunchKernelGGL(k_alart_curnier, dim3((n + 127) / 128), dim3(128), 0, 0, n, dR, dU, dmu,
                     drho, formulation, dF, dA, dB);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipMemcpy(F, dF, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost));
  FC3D_HIP_TRY(hipMemcpy(A, dA, sizeof(double) * 9 * (size_t)n, hipMemcpyDeviceToHost));
  FC3D_HIP_TRY(hipMemcpy(B, dB, sizeof(double) * 9 * (size_t)n, hipMemcpyDeviceToHost));
  hipFree(dR); hipFree(dU); hipFree(dmu); hipFree(drho); hipFree(dF); hipFree(dA); hipFree(dB);
  return 0;
}

int fc3d_hip_projection_on_cone(int n, double* r, const double* mu) {
  if (n < 0) return NUMERICS_ERR_INVALID_ARGUMENT;
  if (n == 0) return 0;
  int rc = use_device(0);
  if (rc) return rc;
  double *dr, *dmu;
  if ((rc = upload(&dr, r, 3 * (size_t)n)) || (rc = upload(&dmu, mu, (size_t)n))) return rc;
  hipLaunchKernelGGL(k_projection, dim3((n + 127) / 128), dim3(128), 0, 0, n, dr, dmu);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipMemcpy(r, dr, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost));
  hipFree(dr);
  hipFree(dmu);
  return 0;
}

int fc3d_hip_local_solve(int n, const double* W, const double* q, const double* mu, double* r,
                         const fc3d_hip_params* p, double* state, int init_state, int* info) {
  if (n < 0 || !p) return NUMERICS_ERR_INVALID_ARGUMENT;
  if (n == 0) return 0;
  int rc = use_device(0);
  if (rc) return rc;
  UnitLocal P;
  P.lp.solver = p->local_solver;
  P.lp.newton_max_iter = p->local_max_iter;
  P.lp.formulation = p->local_formulation;
  P.lp.ls_max_iter = p->ls_max_iter;
  P.lp.hybrid_strategy = p->hybrid_strategy;
  P.lp.hybrid_max_loop = p->hybrid_max_loop;
  P.lp.hybrid_max_iter = p->hybrid_max_iter;
  P.tol = p->local_tol;
  const bool fp32 = p->precision == 1;
  if (fp32) {
    if (p->fp32_local_solver) P.lp.solver = p->fp32_local_solver;
    P.tol = p->fp32_local_tol;
  }
  P.rho_strategy = p->rho_strategy;
  P.rho_constant = p->rho_constant;
  double *dW, *dq, *dmu, *dr, *dst;
  int* dinfo;
  if ((rc = upload(&dW, W, 9 * (size_t)n)) || (rc = upload(&dq, q, 3 * (size_t)n)) ||
      (rc = upload(&dmu, mu, (size_t)n)) || (rc = upload(&dr, r, 3 * (size_t)n)) ||
      (rc = upload(&dst, state, 4 * (size_t)n)) || (rc = upload<int>(&dinfo, nullptr, (size_t)n)))
    return rc;
  if (fp32)
    hipLaunchKernelGGL(k_local_solve<float>, dim3((n + 63) / 64), dim3(64), 0, 0, n, dW, dq, dmu,
                       dr, P, dst, init_state, dinfo);
  else
    hipLaunchKernelGGL(k_local_solve<double>, dim3((n + 63) / 64), dim3(64), 0, 0, n, dW, dq, dmu,
                       dr, P, dst, init_state, dinfo);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipMemcpy(r, dr, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost));
  FC3D_HIP_TRY(hipMemcpy(state, dst, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost));
  FC3D_HIP_TRY(hipMemcpy(info, dinfo, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
  hipFree(dW); hipFree(dq); hipFree(dmu); hipFree(dr); hipFree(dst); hipFree(dinfo);
  return 0;
}

}  // extern "C"
