// fc3d_assembly.hip — W = H M^-1 H^T and q = H v_free on the GPU.
//
// The reference assembles interaction blocks in its mechanics kernel
// (LinearOSNS.cpp:522-700: for two interactions sharing a dynamical system k,
// the block H_{c,k} M_k^-1 H_{d,k}^T, summed over the shared systems) and
// stores them as block CSR (BlockCSRMatrix.cpp:234-263); numerics forms the
// same product on SBM matrices (GlobalFrictionContactProblem.c:399-446).
//
// Here the block structure (which contacts share a body) is built on the
// host, deterministic and ascending per row; the values are computed on the
// device in three kernels:
//   k_asm_g       G_{c,s} = H_{c,s} M_k^-1 per contact side s (3x6),
//   k_asm_blocks  W_cd = sum over shared bodies of G_{c,s} H_{d,t}^T, one
//                 thread per block,
//   k_asm_q       q_c = sum over sides of H_{c,s} v_free,k.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fc3d_internal.hpp"

namespace fc3d {

__global__ void k_asm_g(int nc, const int* __restrict__ ba, const int* __restrict__ bb,
                        const double* __restrict__ H, const double* __restrict__ Minv, int diag,
                        double* __restrict__ G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * nc) return;
  const int c = t >> 1, side = t & 1;
  const int k = side ? bb[c] : ba[c];
  const double* h = H + 36 * (size_t)c + 18 * side;
  double* g = G + 36 * (size_t)c + 18 * side;
  if (k < 0) {
    for (int e = 0; e < 18; e++) g[e] = 0.0;
    return;
  }
  if (diag) {
    const double* m = Minv + 6 * (size_t)k;
    for (int j = 0; j < 6; j++)
      for (int i = 0; i < 3; i++) g[i + 3 * j] = h[i + 3 * j] * m[j];
  } else {
    const double* m = Minv + 36 * (size_t)k;  // 6x6 column-major
    for (int j = 0; j < 6; j++)
      for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int l = 0; l < 6; l++) s += h[i + 3 * l] * m[l + 6 * j];
        g[i + 3 * j] = s;
      }
  }
}

// G_{c,sc} H_{d,sd}^T for one pair of sides, 3x3 column-major.
__device__ inline void side_product(const double* __restrict__ g, const double* __restrict__ h,
                                    double (&s)[9]) {
  double gl[18], hl[18];
#pragma unroll
  for (int e = 0; e < 18; e++) {
    gl[e] = g[e];
    hl[e] = h[e];
  }
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double t = 0.0;
#pragma unroll
      for (int l = 0; l < 6; l++) t += gl[i + 3 * l] * hl[j + 3 * l];
      s[i + 3 * j] = t;
    }
}

// One thread per stored block W_cd: the sum over the pairs of sides (sc, sd)
// that name the same body, in the order (0,0), (0,1), (1,0), (1,1). A block
// off the diagonal usually has one such pair, so the first pair found is
// computed by every lane at its own addresses, without divergence; the
// others (the diagonal's second side, two contacts between the same two
// bodies, contacts with a body on both sides) follow in a loop.
__global__ void k_asm_blocks(int64_t nnzb, const int* __restrict__ row, const int* __restrict__ col,
                             const int* __restrict__ ba, const int* __restrict__ bb,
                             const double* __restrict__ H, const double* __restrict__ G,
                             double* __restrict__ W) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nnzb) return;
  const int c = row[b], d = col[b];
  const int kc[2] = {ba[c], bb[c]}, kd[2] = {ba[d], bb[d]};
  unsigned m = 0;
#pragma unroll
  for (int sc = 0; sc < 2; sc++)
#pragma unroll
    for (int sd = 0; sd < 2; sd++)
      if (kc[sc] >= 0 && kd[sd] == kc[sc]) m |= 1u << (2 * sc + sd);
  double acc[9];
#pragma unroll
  for (int e = 0; e < 9; e++) acc[e] = 0.0;
  while (m) {
    const int p = __builtin_ctz(m), sc = p >> 1, sd = p & 1;
    m &= m - 1;
    double s[9];
    side_product(G + 36 * (size_t)c + 18 * sc, H + 36 * (size_t)d + 18 * sd, s);
#pragma unroll
    for (int e = 0; e < 9; e++) acc[e] += s[e];
  }
#pragma unroll
  for (int e = 0; e < 9; e++) W[9 * b + e] = acc[e];
}

__global__ void k_asm_q(int nc, const int* __restrict__ ba, const int* __restrict__ bb,
                        const double* __restrict__ H, const double* __restrict__ v,
                        double* __restrict__ q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int side = 0; side < 2; side++) {
    const int k = side ? bb[c] : ba[c];
    if (k < 0) continue;
    const double* h = H + 36 * (size_t)c + 18 * side;
    const double* vk = v + 6 * (size_t)k;
    for (int l = 0; l < 6; l++)
      for (int i = 0; i < 3; i++) s[i] += h[i + 3 * l] * vk[l];
  }
  for (int i = 0; i < 3; i++) q[3 * (size_t)c + i] = s[i];
}

// RAII device buffers
struct AsmBufs {
  std::vector<void*> p;
  ~AsmBufs() {
    for (void* x : p) hipFree(x);
  }
  template <typename T>
  hipError_t get(T** out, size_t n) {
    hipError_t e = hipMalloc(out, sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) p.push_back(*out);
    return e;
  }
};

}  // namespace fc3d

using namespace fc3d;

extern "C" int fc3d_hip_assemble(int nbodies, int nc, const int* body_a, const int* body_b,
                                 const double* H, const double* Minv, int minv_diagonal,
                                 const double* vfree, int device, fc3d_generated* out,
                                 double* device_ms) {
  if (!out || !body_a || !body_b || !H || !Minv || !vfree) return NUMERICS_ERR_NULL_POINTER;
  std::memset(out, 0, sizeof *out);
  if (nbodies <= 0 || nc <= 0) {
    set_error("fc3d_hip_assemble: need nbodies > 0 and nc > 0");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  for (int c = 0; c < nc; c++)
    if (body_a[c] < -1 || body_a[c] >= nbodies || body_b[c] < -1 || body_b[c] >= nbodies) {
      set_error("fc3d_hip_assemble: contact %d has a body index outside [-1, %d)", c, nbodies);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
  int rc = use_device(device);
  if (rc) return rc;

  // block structure: row c = every contact sharing a body with c, ascending
  std::vector<int> bptr((size_t)nbodies + 1, 0), blist;
  for (int c = 0; c < nc; c++) {
    if (body_a[c] >= 0) bptr[(size_t)body_a[c] + 1]++;
    if (body_b[c] >= 0 && body_b[c] != body_a[c]) bptr[(size_t)body_b[c] + 1]++;
  }
  for (int k = 0; k < nbodies; k++) bptr[(size_t)k + 1] += bptr[(size_t)k];
  blist.resize((size_t)bptr[(size_t)nbodies]);
  {
    std::vector<int> fill(bptr.begin(), bptr.end() - 1);
    for (int c = 0; c < nc; c++) {
      if (body_a[c] >= 0) blist[(size_t)fill[(size_t)body_a[c]]++] = c;
      if (body_b[c] >= 0 && body_b[c] != body_a[c]) blist[(size_t)fill[(size_t)body_b[c]]++] = c;
    }
  }
  std::vector<int64_t> row_ptr((size_t)nc + 1, 0);
  std::vector<int> cols, rows;
  std::vector<int> tmp;
  for (int c = 0; c < nc; c++) {
    tmp.clear();
    tmp.push_back(c);  // the diagonal block is always stored
    const int ks[2] = {body_a[c], body_b[c]};
    for (int s = 0; s < 2; s++) {
      if (ks[s] < 0) continue;
      for (int t = bptr[(size_t)ks[s]]; t < bptr[(size_t)ks[s] + 1]; t++)
        tmp.push_back(blist[(size_t)t]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    cols.insert(cols.end(), tmp.begin(), tmp.end());
    rows.insert(rows.end(), tmp.size(), c);
    row_ptr[(size_t)c + 1] = (int64_t)cols.size();
  }
  const int64_t nnzb = (int64_t)cols.size();

  AsmBufs B;
  int *d_ba, *d_bb, *d_row, *d_col;
  double *d_H, *d_M, *d_v, *d_G, *d_W, *d_q;
  const size_t mlen = (size_t)nbodies * (minv_diagonal ? 6 : 36);
  FC3D_HIP_TRY(B.get(&d_ba, (size_t)nc));
  FC3D_HIP_TRY(B.get(&d_bb, (size_t)nc));
  FC3D_HIP_TRY(B.get(&d_row, (size_t)nnzb));
  FC3D_HIP_TRY(B.get(&d_col, (size_t)nnzb));
  FC3D_HIP_TRY(B.get(&d_H, 36 * (size_t)nc));
  FC3D_HIP_TRY(B.get(&d_M, mlen));
  FC3D_HIP_TRY(B.get(&d_v, 6 * (size_t)nbodies));
  FC3D_HIP_TRY(B.get(&d_G, 36 * (size_t)nc));
  FC3D_HIP_TRY(B.get(&d_W, 9 * (size_t)nnzb));
  FC3D_HIP_TRY(B.get(&d_q, 3 * (size_t)nc));
  FC3D_HIP_TRY(hipMemcpy(d_ba, body_a, sizeof(int) * (size_t)nc, hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(d_bb, body_b, sizeof(int) * (size_t)nc, hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(d_row, rows.data(), sizeof(int) * (size_t)nnzb, hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(d_col, cols.data(), sizeof(int) * (size_t)nnzb, hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(d_H, H, sizeof(double) * 36 * (size_t)nc, hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(d_M, Minv, sizeof(double) * mlen, hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(d_v, vfree, sizeof(double) * 6 * (size_t)nbodies, hipMemcpyHostToDevice));

  hipEvent_t e0, e1;
  FC3D_HIP_TRY(hipEventCreate(&e0));
  FC3D_HIP_TRY(hipEventCreate(&e1));
  FC3D_HIP_TRY(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_asm_g, dim3((unsigned)((2 * (size_t)nc + 255) / 256)), dim3(256), 0, 0, nc,
                     (const int*)d_ba, (const int*)d_bb, (const double*)d_H, (const double*)d_M,
                     minv_diagonal ? 1 : 0, d_G);
  hipLaunchKernelGGL(k_asm_blocks, dim3((unsigned)((nnzb + 255) / 256)), dim3(256), 0, 0, nnzb,
                     (const int*)d_row, (const int*)d_col, (const int*)d_ba, (const int*)d_bb,
                     (const double*)d_H, (const double*)d_G, d_W);
  hipLaunchKernelGGL(k_asm_q, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0, nc,
                     (const int*)d_ba, (const int*)d_bb, (const double*)d_H, (const double*)d_v,
                     d_q);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipEventRecord(e1, 0));
  FC3D_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (device_ms) *device_ms = ms;

  out->nc = nc;
  out->nbodies = nbodies;
  out->nnzb = nnzb;
  out->row_ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nc + 1));
  out->col_idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnzb ? nnzb : 1));
  out->blocks = (double*)malloc(sizeof(double) * 9 * (size_t)(nnzb ? nnzb : 1));
  out->q = (double*)malloc(sizeof(double) * 3 * (size_t)nc);
  if (!out->row_ptr || !out->col_idx || !out->blocks || !out->q) {
    fc3d_generated_free(out);
    set_error("fc3d_hip_assemble: out of host memory");
    return NUMERICS_ERR_MEMORY;
  }
  std::memcpy(out->row_ptr, row_ptr.data(), sizeof(int64_t) * ((size_t)nc + 1));
  for (int64_t b = 0; b < nnzb; b++) out->col_idx[b] = cols[(size_t)b];
  hipError_t e = hipMemcpy(out->blocks, d_W, sizeof(double) * 9 * (size_t)nnzb,
                           hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(out->q, d_q, sizeof(double) * 3 * (size_t)nc, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    fc3d_generated_free(out);
    set_error("fc3d_hip_assemble: read-back failed: %s", hipGetErrorString(e));
    return NUMERICS_ERR_DEVICE;
  }
  return 0;
}
