// fc3d_noalloc.hip — fc3d_compute_error on caller-owned device arrays, with
// no allocation (fc3d_hip_compute_error_device in fc3d_hip.h).
//
// The residual of fc3d_compute_error.c:70-122: u = q + W r with each row's
// blocks in their stored order (the order SBM_gemv_3x3 adds them in,
// SparseBlockMatrix.c:360-419), the unitary error of every contact, and the
// norms of r and u.  One thread per contact, one partial sum triple per wave
// of 64 contacts, then one workgroup adds the wave partials in a fixed order.
//
// The wave partials live in `scratch`, a file-scope device array of
// NOALLOC_MAX_WAVES triples (1.5 MiB per GPU, held until the process ends), so
// a call allocates nothing.  Host code never passes the array's name to a
// kernel: its device address is taken once per device with
// hipGetSymbolAddress, with that device current, and kept in g_scratch.  Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU
// (two host threads, or two streams) race; the array keeps the last call's
// partials, and every partial is rewritten before it is read.
#include <cmath>
#include <mutex>

#include "fc3d_internal.hpp"
#include "fc3d_sweep.hpp"

namespace fc3d {

constexpr int NOALLOC_MAX_CONTACTS = 1 << 22;
constexpr int NOALLOC_MAX_WAVES = NOALLOC_MAX_CONTACTS / 64;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[3 * NOALLOC_MAX_WAVES];

static std::mutex g_scratch_lock;
static double* g_scratch[NOALLOC_MAX_DEVICES] = {};

// The scratch array's address on the current device, taken on first use.
static int scratch_on(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_lock);
  if (!g_scratch[device]) {
    void* p = nullptr;
    FC3D_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return 0;
}

// Row ranges and column indices are checked against nnzb and nc: a row with
// an entry outside them gets a NaN velocity (and so a NaN error) instead of
// reading out of bounds.
__global__ __launch_bounds__(256) void k_csr_residual(int nc, int64_t nnzb,
                                                      const int64_t* __restrict__ rp,
                                                      const int64_t* __restrict__ ci,
                                                      const double* __restrict__ blocks,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ mu,
                                                      const double* __restrict__ r,
                                                      double* __restrict__ u_out,
                                                      double* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0, r2 = 0.0, u2 = 0.0;
  if (i < nc) {
    double u[3] = {q[3 * (size_t)i], q[3 * (size_t)i + 1], q[3 * (size_t)i + 2]};
    const int64_t k0 = rp[i], k1 = rp[i + 1];
    bool bad = k0 < 0 || k1 > nnzb || k1 < k0;
    for (int64_t k = bad ? k1 : k0; k < k1; k++) {
      const int64_t j = ci[k];
      if (j < 0 || j >= nc) {
        bad = true;
        break;
      }
      double w[9];
#pragma unroll
      for (int e = 0; e < 9; e++) w[e] = blocks[9 * k + e];
      const double v[3] = {r[3 * j], r[3 * j + 1], r[3 * j + 2]};
      mvp3x3<double>(w, v, u);
    }
    if (bad) u[0] = u[1] = u[2] = NAN;
#pragma unroll
    for (int e = 0; e < 3; e++) u_out[3 * (size_t)i + e] = u[e];
    const double rv[3] = {r[3 * (size_t)i], r[3 * (size_t)i + 1], r[3 * (size_t)i + 2]};
    e2 = unitary_error<double>(rv, u, mu[i]);
    r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
    u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  }
  e2 = wave_sum(e2);
  r2 = wave_sum(r2);
  u2 = wave_sum(u2);
  if ((threadIdx.x & 63) == 0) {
    const size_t wv = (size_t)i >> 6;
    part[3 * wv] = e2;
    part[3 * wv + 1] = r2;
    part[3 * wv + 2] = u2;
  }
}

// fc3d_compute_error.c:100-115: the error relative to max(|q|, |r|, |u|).
__global__ __launch_bounds__(1024) void k_csr_residual_end(const double* __restrict__ part,
                                                           int nwaves, double norm,
                                                           double* __restrict__ error) {
  __shared__ double sh[3 * 16];
  double sums[3];
  block_sums<3>(part, nwaves, sh, sums);
  if (threadIdx.x != 0) return;
  double err = sqrt(sums[0]);
  const double scale = fmax(norm, fmax(sqrt(sums[1]), sqrt(sums[2])));
  if (fabs(scale) > DBL_EPSILON) err /= scale;
  *error = err;
}

}  // namespace fc3d

extern "C" {

int fc3d_hip_compute_error_device(int nc, int64_t nnzb, const int64_t* row_ptr,
                                  const int64_t* col_idx,
                                  const double* blocks, const double* q, const double* mu,
                                  const double* reaction, double norm, double* velocity,
                                  double* error, int device, void* stream) {
  using namespace fc3d;
  if (!row_ptr || !col_idx || !blocks || !q || !mu || !reaction || !velocity || !error)
    return NUMERICS_ERR_NULL_POINTER;
  if (nc < 1 || nc > NOALLOC_MAX_CONTACTS) {
    set_error("fc3d_hip_compute_error_device: %d contacts, outside [1, %d] (the size of the "
              "library's static scratch array)", nc, NOALLOC_MAX_CONTACTS);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (nnzb < 0) {
    set_error("fc3d_hip_compute_error_device: nnzb %lld < 0", (long long)nnzb);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (!(norm >= 0.0) || !std::isfinite(norm)) {
    set_error("fc3d_hip_compute_error_device: norm %g is not a finite value >= 0", norm);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
    set_error("fc3d_hip_compute_error_device: device %d outside [0, %d)", device,
              NOALLOC_MAX_DEVICES);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  int rc = use_device(device);
  if (rc) return rc;
  double* part = nullptr;
  if ((rc = scratch_on(device, &part))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nwaves = (nc + 63) / 64;
  hipLaunchKernelGGL(k_csr_residual, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, nc,
                     nnzb, row_ptr, col_idx, blocks, q, mu, reaction, velocity, part);
  hipLaunchKernelGGL(k_csr_residual_end, dim3(1), dim3(1024), 0, s, (const double*)part, nwaves,
                     norm, error);
  FC3D_HIP_TRY(hipGetLastError());
  return 0;
}

}  // extern "C"
