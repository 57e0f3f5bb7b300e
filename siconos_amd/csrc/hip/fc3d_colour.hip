// fc3d_colour.hip — deterministic greedy colouring of the contact graph.
//
// Definition: in index order, contact i takes the smallest colour not used by
// any lower-index neighbour j (block W_ij stored).  Sequentially this is the
// greedy colouring; here every round colours, in parallel, the contacts whose
// lower-index neighbours are all coloured.  A contact's colour depends only on
// those neighbours, so the result equals the sequential one bit for bit,
// whatever the order in which rounds see writes.  The round count is the
// longest index-decreasing path of the graph.
#include <vector>

#include "fc3d_internal.hpp"

namespace fc3d {

__global__ __launch_bounds__(256) void k_colour_round(int nc, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, int* colour,
                                                      int* uncoloured) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  volatile int* vc = colour;
  if (vc[i] >= 0) return;
  const int b = rp[i], e = rp[i + 1];
  for (int k = b; k < e; k++) {
    const int j = ci[k];
    if (j < i && vc[j] < 0) {
      atomicAdd(uncoloured, 1);
      return;
    }
  }
  // every lower neighbour is coloured: smallest free colour, 64 at a time
  for (int base = 0;; base += 64) {
    unsigned long long used = 0ull;
    for (int k = b; k < e; k++) {
      const int j = ci[k];
      if (j < i) {
        const int c = vc[j] - base;
        if (c >= 0 && c < 64) used |= 1ull << c;
      }
    }
    if (used != ~0ull) {
      vc[i] = base + __ffsll((long long)~used) - 1;
      return;
    }
  }
}

// Colours on the current device; rp/ci are device arrays (int32).
int colour_on_device(int nc, const int* d_rp, const int* d_ci, int* d_colour, hipStream_t s,
                     int* ncolours_out, std::vector<int>* host_colour) {
  int* d_cnt = nullptr;
  FC3D_HIP_TRY(hipMalloc(&d_cnt, sizeof(int)));
  FC3D_HIP_TRY(hipMemsetAsync(d_colour, 0xff, sizeof(int) * (size_t)(nc > 0 ? nc : 1), s));
  const int grid = (nc + 255) / 256;
  int rounds = 0;
  int rc = 0;
  for (;;) {
    // several rounds per host check; a round with nothing left to do is free
    for (int k = 0; k < 8; k++) {
      hipMemsetAsync(d_cnt, 0, sizeof(int), s);
      if (nc > 0) hipLaunchKernelGGL(k_colour_round, dim3(grid), dim3(256), 0, s, nc, d_rp, d_ci,
                                     d_colour, d_cnt);
      rounds++;
    }
    int left = 0;
    hipError_t e = hipMemcpyAsync(&left, d_cnt, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("colouring: %s", hipGetErrorString(e));
      rc = NUMERICS_ERR_DEVICE;
      break;
    }
    if (left == 0) break;
    if (rounds > 4 * nc + 64) {  // cannot happen for a valid graph; guards the loop
      set_error("colouring did not finish after %d rounds", rounds);
      rc = NUMERICS_ERR_UNKNOWN;
      break;
    }
  }
  hipFree(d_cnt);
  if (rc) return rc;
  host_colour->resize((size_t)nc);
  if (nc > 0)
    FC3D_HIP_TRY(hipMemcpy(host_colour->data(), d_colour, sizeof(int) * (size_t)nc,
                           hipMemcpyDeviceToHost));
  int ncol = 0;
  for (int i = 0; i < nc; i++) ncol = (*host_colour)[i] + 1 > ncol ? (*host_colour)[i] + 1 : ncol;
  *ncolours_out = ncol;
  return 0;
}

}  // namespace fc3d

extern "C" int fc3d_hip_colour(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                               int* colour) {
  if (nc < 0 || (nc > 0 && (!row_ptr || !col_idx || !colour))) {
    fc3d::set_error("fc3d_hip_colour: invalid arguments");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  const int64_t nnz = nc > 0 ? row_ptr[nc] : 0;
  if (nnz > 0x7fffffffLL) {
    fc3d::set_error("fc3d_hip_colour: %lld blocks exceed the int32 index range", (long long)nnz);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  for (int i = 0; i < nc; i++) {
    if (row_ptr[i + 1] < row_ptr[i]) {
      fc3d::set_error("fc3d_hip_colour: row pointers not monotone at %d", i);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
  }
  for (int64_t k = 0; k < nnz; k++) {
    if (col_idx[k] < 0 || col_idx[k] >= nc) {
      fc3d::set_error("fc3d_hip_colour: column index %lld out of range", (long long)col_idx[k]);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
  }
  int rc = fc3d::use_device(0);
  if (rc) return rc;
  std::vector<int> rp((size_t)nc + 1), ci((size_t)(nnz > 0 ? nnz : 1));
  for (int i = 0; i <= nc; i++) rp[(size_t)i] = (int)row_ptr[i];
  for (int64_t k = 0; k < nnz; k++) ci[(size_t)k] = (int)col_idx[k];
  int *d_rp = nullptr, *d_ci = nullptr, *d_col = nullptr;
  FC3D_HIP_TRY(hipMalloc(&d_rp, sizeof(int) * rp.size()));
  FC3D_HIP_TRY(hipMalloc(&d_ci, sizeof(int) * ci.size()));
  FC3D_HIP_TRY(hipMalloc(&d_col, sizeof(int) * (size_t)(nc > 0 ? nc : 1)));
  FC3D_HIP_TRY(hipMemcpy(d_rp, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(d_ci, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice));
  std::vector<int> hc;
  int ncol = 0;
  rc = fc3d::colour_on_device(nc, d_rp, d_ci, d_col, 0, &ncol, &hc);
  hipFree(d_rp);
  hipFree(d_ci);
  hipFree(d_col);
  if (rc) return rc;
  for (int i = 0; i < nc; i++) colour[i] = hc[(size_t)i];
  return ncol;
}
