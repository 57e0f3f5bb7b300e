// fc3d_mlp.hip — the learned warm start's network (no reference counterpart).
//
// MLP 22 -> 128 -> 128 -> 128 -> 3 with GELU.  Every product runs on the
// matrix cores as v_mfma_f32_16x16x16bf16_1k (BF16 operands, FP32
// accumulation):
//   forward         one wavefront carries 16 contacts through all layers,
//                   activations in LDS (16 x 128 BF16 per wave per buffer);
//   backward        the same shape, dA = dZ W with W^T kept in HBM as BF16;
//   weight gradient one 16 x 16 tile of dW per wavefront, the minibatch as
//                   the MFMA reduction (activations and dZ are stored
//                   feature-major so that both operands are contiguous);
// then Adam on FP32 master weights and a BF16 re-cast of W and W^T.
//
// MFMA 16x16x16 operand layout (wave64, lane l, i = 0..3):
//   A (M x K):  lane holds A[l % 16][4 (l / 16) + i]
//   B (K x N):  lane holds B[4 (l / 16) + i][l % 16]
//   D (M x N):  lane holds D[4 (l / 16) + i][l % 16]
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "fc3d_internal.hpp"
#include "fc3d_mlp.hpp"

namespace fc3d {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

// FP32 -> BF16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(bf16 b) {
  return __builtin_bit_cast(float, (unsigned)b << 16);
}

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// parameter offsets in the FP32 master vector
constexpr int OW1 = 0, OB1 = OW1 + MLP_H * MLP_K1, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H,
              OW3 = OB2 + MLP_H, OB3 = OW3 + MLP_H * MLP_H, OW4 = OB3 + MLP_H,
              OB4 = OW4 + MLP_NOP * MLP_H, NPARAM = OB4 + MLP_NOP;
static_assert(NPARAM == MLP_NPARAMS, "parameter count");

struct MlpWeights {
  const bf16* w[4];   // [n][k]
  const bf16* wt[4];  // [k][n]
  const float* b[4];
};

// acc[t] = act x W^T for the 16 x 16 output tiles t = 0..NT-1, where act is
// [16][K] (LDS) and W is [16 NT][K] (rows = outputs).
template <int K, int NT>
__device__ __forceinline__ void wave_dense(const bf16* act, const bf16* __restrict__ W, f32x4* acc) {
  const int l = threadIdx.x & 63, row = l & 15, kq = l >> 4;
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 16) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(&act[row * K + k0 + 4 * kq]);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const bf16x4 b =
          *reinterpret_cast<const bf16x4*>(&W[(size_t)(t * 16 + row) * K + k0 + 4 * kq]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
}

struct TrainBufs {
  int bsz;              // rows of the pass (multiple of 64)
  float* Z[3];          // pre-activations [bsz][128]
  bf16* At[4];          // feature-major layer inputs: At[0] [32][bsz], At[l] [128][bsz]
  float* out;           // [bsz][16]
  bf16* dZt[4];         // feature-major dZ: [128][bsz] for l = 0..2, [16][bsz] for l = 3
  const float* target;  // [bsz][3]
  float* loss_part;     // per-wave sum of squared errors
};

// Forward pass, 16 contacts per wave, 4 waves per block; X is [>= 64 blocks][32]
// standardised BF16 inputs.  Writes y[n][3] and, when training, what the
// backward kernels need.
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_forward(int n, const bf16* __restrict__ X,
                                                     MlpWeights w, float* __restrict__ y,
                                                     TrainBufs tb) {
  __shared__ __attribute__((aligned(16))) bf16 buf[4][2][16 * MLP_H];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int m0 = (blockIdx.x * 4 + wave) * 16;
  bf16* a0 = buf[wave][0];
  bf16* a1 = buf[wave][1];
  for (int e = l; e < 16 * MLP_K1; e += 64) {
    const int m = e / MLP_K1, k = e % MLP_K1;
    const bf16 x = X[(size_t)(m0 + m) * MLP_K1 + k];
    a0[m * MLP_K1 + k] = x;
    if (TRAIN) tb.At[0][(size_t)k * tb.bsz + m0 + m] = x;
  }
  __syncthreads();
  f32x4 acc[8];
  const int col = l & 15, rq = l >> 4;
  for (int layer = 0; layer < 3; layer++) {
    const bf16* in = (layer & 1) ? a1 : a0;
    bf16* outb = (layer & 1) ? a0 : a1;
    if (layer == 0)
      wave_dense<MLP_K1, 8>(in, w.w[0], acc);
    else
      wave_dense<MLP_H, 8>(in, w.w[layer], acc);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; t++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = 4 * rq + i, nn = t * 16 + col;
        const float z = acc[t][i] + w.b[layer][nn];
        const bf16 a = f2bf(gelu(z));
        outb[m * MLP_H + nn] = a;
        if (TRAIN) {
          tb.Z[layer][(size_t)(m0 + m) * MLP_H + nn] = z;
          tb.At[layer + 1][(size_t)nn * tb.bsz + m0 + m] = a;
        }
      }
    }
    __syncthreads();
  }
  // output layer: the third hidden activation is in a1
  wave_dense<MLP_H, 1>(a1, w.w[3], acc);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = 4 * rq + i;
    const float v = acc[0][i] + w.b[3][col];
    if (m0 + m < n && col < MLP_NO) y[(size_t)(m0 + m) * MLP_NO + col] = v;
    if (TRAIN) tb.out[(size_t)(m0 + m) * MLP_NOP + col] = v;
  }
}

// Backward pass of the data path for the mean squared error over the first
// `nvalid` rows: dZ of the output layer, then dA = dZ W and dZ = dA gelu'(Z)
// down to the first layer, stored feature-major in BF16.
__global__ __launch_bounds__(256) void k_mlp_backward(int nvalid, MlpWeights w, TrainBufs tb) {
  __shared__ __attribute__((aligned(16))) bf16 buf[4][2][16 * MLP_H];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int m0 = (blockIdx.x * 4 + wave) * 16;
  const int col = l & 15, rq = l >> 4;
  bf16* d0 = buf[wave][0];
  bf16* d1 = buf[wave][1];
  const float scale = 2.0f / (float)(nvalid * MLP_NO);
  float lsum = 0.f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = 4 * rq + i;
    float g = 0.f;
    if (col < MLP_NO && m0 + m < nvalid) {
      const float diff = tb.out[(size_t)(m0 + m) * MLP_NOP + col] -
                         tb.target[(size_t)(m0 + m) * MLP_NO + col];
      g = scale * diff;
      lsum += diff * diff;
    }
    const bf16 gb = f2bf(g);
    d0[m * MLP_NOP + col] = gb;
    tb.dZt[3][(size_t)col * tb.bsz + m0 + m] = gb;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  if (l == 0) tb.loss_part[blockIdx.x * 4 + wave] = lsum;
  __syncthreads();
  f32x4 acc[8];
  // dA of the third hidden layer = dZ4 W4 (reduction over the 16 outputs)
  wave_dense<MLP_NOP, 8>(d0, w.wt[3], acc);
  __syncthreads();
  bf16* cur = d1;
  for (int layer = 2; layer >= 0; layer--) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = 4 * rq + i, nn = t * 16 + col;
        const float z = tb.Z[layer][(size_t)(m0 + m) * MLP_H + nn];
        const bf16 g = f2bf(acc[t][i] * gelu_grad(z));
        cur[m * MLP_H + nn] = g;
        tb.dZt[layer][(size_t)nn * tb.bsz + m0 + m] = g;
      }
    }
    __syncthreads();
    if (layer == 0) break;
    // dA of layer `layer - 1`'s output = dZ_layer W_layer
    wave_dense<MLP_H, 8>(cur, w.wt[layer], acc);
    __syncthreads();
    cur = (cur == d1) ? d0 : d1;
  }
}

// dW[n][k] = sum_m dZ[m][n] A[m][k]: one 16 x 16 tile per wave.
__global__ __launch_bounds__(256) void k_mlp_wgrad(int bsz, const bf16* __restrict__ dZt,
                                                   const bf16* __restrict__ At, int kin,
                                                   float* __restrict__ dW, int ntiles) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= ntiles) return;
  const int ktiles = kin / 16;
  const int n0 = (tile / ktiles) * 16, k0 = (tile % ktiles) * 16;
  const int r = l & 15, q = l >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int m0 = 0; m0 < bsz; m0 += 16) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(&dZt[(size_t)(n0 + r) * bsz + m0 + 4 * q]);
    const bf16x4 b = *reinterpret_cast<const bf16x4*>(&At[(size_t)(k0 + r) * bsz + m0 + 4 * q]);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) dW[(size_t)(n0 + 4 * q + i) * kin + k0 + r] = acc[i];
}

// db[n] = sum_m dZ[m][n], one wave per output, fixed-order lane sums
__global__ __launch_bounds__(64) void k_mlp_bgrad(int bsz, const bf16* __restrict__ dZt,
                                                  float* __restrict__ db) {
  const int nn = blockIdx.x, l = threadIdx.x;
  float s = 0.f;
  for (int m = l; m < bsz; m += 64) s += bf2f(dZt[(size_t)nn * bsz + m]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (l == 0) db[nn] = s;
}

__global__ void k_adam(int np, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 copies of one weight matrix, [n][k] and transposed [k][n]
__global__ void k_mlp_cast(const float* __restrict__ p, bf16* __restrict__ w, bf16* __restrict__ wt,
                           int nout, int kin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout * kin) return;
  const int nn = i / kin, k = i % kin;
  const bf16 v = f2bf(p[i]);
  w[i] = v;
  wt[(size_t)k * nout + nn] = v;
}

// X[npad][32] = bf16((f - mean) / std), zero for padded columns and rows >= n
__global__ void k_mlp_inputs(int n, int npad, const float* __restrict__ f,
                             const float* __restrict__ mean, const float* __restrict__ stdv,
                             bf16* __restrict__ X) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)npad * MLP_K1) return;
  const int m = (int)(i / MLP_K1), k = (int)(i % MLP_K1);
  X[i] = (k < MLP_NF && m < n) ? f2bf((f[(size_t)m * MLP_NF + k] - mean[k]) / stdv[k]) : (bf16)0;
}

// minibatch rows: bx[r] = X[idx[r]], bt[r] = T[idx[r]]
__global__ void k_mlp_gather(int batch, const int* __restrict__ idx, const bf16* __restrict__ X,
                             const float* __restrict__ T, bf16* __restrict__ bx,
                             float* __restrict__ bt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * MLP_K1) return;
  const int r = i / MLP_K1, k = i % MLP_K1;
  const int src = idx[r];
  bx[i] = X[(size_t)src * MLP_K1 + k];
  if (k < MLP_NO) bt[(size_t)r * MLP_NO + k] = T[(size_t)src * MLP_NO + k];
}

}  // namespace fc3d

using namespace fc3d;

struct fc3d_hip_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  bf16* w[4] = {nullptr, nullptr, nullptr, nullptr};
  bf16* wt[4] = {nullptr, nullptr, nullptr, nullptr};
  float *mean = nullptr, *stdv = nullptr;
  long adam_t = 0;
  // training workspace for passes of up to ws_bsz rows
  int ws_bsz = 0;
  float* Z[3] = {nullptr, nullptr, nullptr};
  bf16* At[4] = {nullptr, nullptr, nullptr, nullptr};
  bf16* dZt[4] = {nullptr, nullptr, nullptr, nullptr};
  float *out = nullptr, *loss_part = nullptr;
};

namespace fc3d {

static const int kDims[4][2] = {{MLP_H, MLP_K1}, {MLP_H, MLP_H}, {MLP_H, MLP_H}, {MLP_NOP, MLP_H}};
static const int kWOff[4] = {OW1, OW2, OW3, OW4};
static const int kBOff[4] = {OB1, OB2, OB3, OB4};
static fc3d_hip_mlp* g_default_model = nullptr;

static MlpWeights mlp_weights(const fc3d_hip_mlp* m) {
  MlpWeights w;
  for (int k = 0; k < 4; k++) {
    w.w[k] = m->w[k];
    w.wt[k] = m->wt[k];
    w.b[k] = m->p + kBOff[k];
  }
  return w;
}

static TrainBufs no_train() {
  TrainBufs t;
  std::memset(&t, 0, sizeof t);
  return t;
}

static int mlp_cast(fc3d_hip_mlp* m) {
  for (int k = 0; k < 4; k++) {
    const int n = kDims[k][0] * kDims[k][1];
    hipLaunchKernelGGL(k_mlp_cast, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->p + kWOff[k],
                       m->w[k], m->wt[k], kDims[k][0], kDims[k][1]);
  }
  FC3D_HIP_TRY(hipGetLastError());
  return 0;
}

static void free_workspace(fc3d_hip_mlp* m) {
  for (int k = 0; k < 3; k++) {
    hipFree(m->Z[k]);
    m->Z[k] = nullptr;
  }
  for (int k = 0; k < 4; k++) {
    hipFree(m->At[k]);
    hipFree(m->dZt[k]);
    m->At[k] = m->dZt[k] = nullptr;
  }
  hipFree(m->out);
  hipFree(m->loss_part);
  m->out = m->loss_part = nullptr;
  m->ws_bsz = 0;
}

static int mlp_workspace(fc3d_hip_mlp* m, int bsz) {
  if (m->ws_bsz >= bsz) return 0;
  FC3D_HIP_TRY(hipStreamSynchronize(m->stream));
  free_workspace(m);
  for (int k = 0; k < 3; k++) FC3D_HIP_TRY(hipMalloc(&m->Z[k], sizeof(float) * bsz * MLP_H));
  FC3D_HIP_TRY(hipMalloc(&m->At[0], sizeof(bf16) * bsz * MLP_K1));
  for (int k = 1; k < 4; k++) FC3D_HIP_TRY(hipMalloc(&m->At[k], sizeof(bf16) * bsz * MLP_H));
  for (int k = 0; k < 3; k++) FC3D_HIP_TRY(hipMalloc(&m->dZt[k], sizeof(bf16) * bsz * MLP_H));
  FC3D_HIP_TRY(hipMalloc(&m->dZt[3], sizeof(bf16) * bsz * MLP_NOP));
  FC3D_HIP_TRY(hipMalloc(&m->out, sizeof(float) * bsz * MLP_NOP));
  FC3D_HIP_TRY(hipMalloc(&m->loss_part, sizeof(float) * (bsz / 16)));
  m->ws_bsz = bsz;
  return 0;
}

// Forward + backward over `bsz` rows (the first `nvalid` carry targets):
// gradients into m->g, per-wave loss partials into m->loss_part.
static int mlp_grad_pass(fc3d_hip_mlp* m, int nvalid, int bsz, const bf16* X, const float* T) {
  TrainBufs tb;
  tb.bsz = bsz;
  for (int k = 0; k < 3; k++) tb.Z[k] = m->Z[k];
  for (int k = 0; k < 4; k++) {
    tb.At[k] = m->At[k];
    tb.dZt[k] = m->dZt[k];
  }
  tb.out = m->out;
  tb.target = T;
  tb.loss_part = m->loss_part;
  const MlpWeights w = mlp_weights(m);
  const int blocks = bsz / 64;
  hipLaunchKernelGGL(k_mlp_forward<true>, dim3(blocks), dim3(256), 0, m->stream, 0, X, w,
                     (float*)nullptr, tb);
  hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(256), 0, m->stream, nvalid, w, tb);
  for (int k = 0; k < 4; k++) {
    const int nout = kDims[k][0], kin = kDims[k][1];
    const int ntiles = (nout / 16) * (kin / 16);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((ntiles + 3) / 4), dim3(256), 0, m->stream, bsz,
                       m->dZt[k], m->At[k], kin, m->g + kWOff[k], ntiles);
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(nout), dim3(64), 0, m->stream, bsz, m->dZt[k],
                       m->g + kBOff[k]);
  }
  FC3D_HIP_TRY(hipGetLastError());
  return 0;
}

static double sum_loss(fc3d_hip_mlp* m, int bsz, int* rc) {
  std::vector<float> lp((size_t)bsz / 16);
  hipError_t e = hipMemcpyAsync(lp.data(), m->loss_part, sizeof(float) * lp.size(),
                                hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    set_error("fc3d_hip_mlp: loss read-back failed: %s", hipGetErrorString(e));
    *rc = NUMERICS_ERR_DEVICE;
    return 0.0;
  }
  double s = 0.0;
  for (float v : lp) s += v;
  return s;
}

int mlp_device(const fc3d_hip_mlp* m) { return m->device; }

int mlp_forward_device(fc3d_hip_mlp* m, int n, const float* feat, bf16* X, float* y,
                       hipStream_t s) {
  const int npad = (n + 63) / 64 * 64;
  hipLaunchKernelGGL(k_mlp_inputs, dim3((unsigned)(((size_t)npad * MLP_K1 + 255) / 256)),
                     dim3(256), 0, s, n, npad, feat, m->mean, m->stdv, X);
  hipLaunchKernelGGL(k_mlp_forward<false>, dim3(npad / 64), dim3(256), 0, s, n, X,
                     mlp_weights(m), y, no_train());
  FC3D_HIP_TRY(hipGetLastError());
  return 0;
}

// RAII list of device buffers freed on every return path
struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) hipFree(q);
  }
  template <typename T>
  hipError_t get(T** out, size_t n) {
    hipError_t e = hipMalloc(out, sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) p.push_back(*out);
    return e;
  }
};

}  // namespace fc3d

extern "C" {

int fc3d_hip_mlp_create(uint64_t seed, int device, fc3d_hip_mlp** out) {
  if (!out) return NUMERICS_ERR_NULL_POINTER;
  *out = nullptr;
  int rc = use_device(device);
  if (rc) return rc;
  fc3d_hip_mlp* m = new fc3d_hip_mlp();
  m->device = device;
  auto fail = [&](int code) {
    fc3d_hip_mlp_destroy(m);
    return code;
  };
#define MLP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      set_error("fc3d_hip_mlp_create: %s -> %s", #expr, hipGetErrorString(e_));            \
      return fail(NUMERICS_ERR_DEVICE);                                                    \
    }                                                                                      \
  } while (0)
  MLP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  MLP_TRY(hipMalloc(&m->p, sizeof(float) * NPARAM));
  MLP_TRY(hipMalloc(&m->g, sizeof(float) * NPARAM));
  MLP_TRY(hipMalloc(&m->m1, sizeof(float) * NPARAM));
  MLP_TRY(hipMalloc(&m->m2, sizeof(float) * NPARAM));
  MLP_TRY(hipMalloc(&m->mean, sizeof(float) * MLP_NF));
  MLP_TRY(hipMalloc(&m->stdv, sizeof(float) * MLP_NF));
  for (int k = 0; k < 4; k++) {
    MLP_TRY(hipMalloc(&m->w[k], sizeof(bf16) * kDims[k][0] * kDims[k][1]));
    MLP_TRY(hipMalloc(&m->wt[k], sizeof(bf16) * kDims[k][0] * kDims[k][1]));
  }
  // He-uniform weights (seeded), zero biases; padded input columns and unused
  // output rows start at zero and only ever receive zero gradients
  std::vector<float> p(NPARAM, 0.f), zeros(NPARAM, 0.f);
  std::mt19937_64 gen(seed);
  for (int k = 0; k < 4; k++) {
    const int nout = kDims[k][0], kin = kDims[k][1];
    const int fan_in = k == 0 ? MLP_NF : MLP_H;
    const float a = std::sqrt(6.0f / (float)fan_in) * (k == 3 ? 0.1f : 1.0f);
    std::uniform_real_distribution<float> U(-a, a);
    for (int nn = 0; nn < nout; nn++)
      for (int kk = 0; kk < kin; kk++) {
        const bool used = (k != 0 || kk < MLP_NF) && (k != 3 || nn < MLP_NO);
        const float v = U(gen);
        p[(size_t)kWOff[k] + (size_t)nn * kin + kk] = used ? v : 0.f;
      }
  }
  std::vector<float> ones(MLP_NF, 1.f);
  MLP_TRY(hipMemcpy(m->p, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->m1, zeros.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->m2, zeros.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->g, zeros.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->mean, zeros.data(), sizeof(float) * MLP_NF, hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->stdv, ones.data(), sizeof(float) * MLP_NF, hipMemcpyHostToDevice));
#undef MLP_TRY
  if ((rc = mlp_cast(m))) return fail(rc);
  if (hipStreamSynchronize(m->stream) != hipSuccess) return fail(NUMERICS_ERR_DEVICE);
  *out = m;
  return 0;
}

void fc3d_hip_mlp_destroy(fc3d_hip_mlp* m) {
  if (!m) return;
  hipSetDevice(m->device);
  if (m->stream) hipStreamSynchronize(m->stream);
  if (g_default_model == m) g_default_model = nullptr;
  void* ptrs[] = {m->p, m->g, m->m1, m->m2, m->mean, m->stdv};
  for (void* q : ptrs)
    if (q) hipFree(q);
  for (int k = 0; k < 4; k++) {
    hipFree(m->w[k]);
    hipFree(m->wt[k]);
  }
  free_workspace(m);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

int fc3d_hip_mlp_get_params(const fc3d_hip_mlp* m, float* params, float* mean, float* stdv) {
  if (!m) return NUMERICS_ERR_NULL_POINTER;
  int rc = use_device(m->device);
  if (rc) return rc;
  FC3D_HIP_TRY(hipStreamSynchronize(m->stream));
  if (params) FC3D_HIP_TRY(hipMemcpy(params, m->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  if (mean) FC3D_HIP_TRY(hipMemcpy(mean, m->mean, sizeof(float) * MLP_NF, hipMemcpyDeviceToHost));
  if (stdv) FC3D_HIP_TRY(hipMemcpy(stdv, m->stdv, sizeof(float) * MLP_NF, hipMemcpyDeviceToHost));
  return 0;
}

int fc3d_hip_mlp_set_params(fc3d_hip_mlp* m, const float* params, const float* mean,
                            const float* stdv) {
  if (!m) return NUMERICS_ERR_NULL_POINTER;
  if (stdv)
    for (int k = 0; k < MLP_NF; k++)
      if (!(stdv[k] > 0.f)) {
        set_error("fc3d_hip_mlp_set_params: standard deviation %d is not positive", k);
        return NUMERICS_ERR_INVALID_ARGUMENT;
      }
  int rc = use_device(m->device);
  if (rc) return rc;
  FC3D_HIP_TRY(hipStreamSynchronize(m->stream));
  if (params) FC3D_HIP_TRY(hipMemcpy(m->p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  if (mean) FC3D_HIP_TRY(hipMemcpy(m->mean, mean, sizeof(float) * MLP_NF, hipMemcpyHostToDevice));
  if (stdv) FC3D_HIP_TRY(hipMemcpy(m->stdv, stdv, sizeof(float) * MLP_NF, hipMemcpyHostToDevice));
  if ((rc = mlp_cast(m))) return rc;
  FC3D_HIP_TRY(hipStreamSynchronize(m->stream));
  return 0;
}

int fc3d_hip_mlp_forward(fc3d_hip_mlp* m, int n, const float* x, float* y) {
  if (!m || (n > 0 && (!x || !y)) || n < 0) return NUMERICS_ERR_INVALID_ARGUMENT;
  if (n == 0) return 0;
  int rc = use_device(m->device);
  if (rc) return rc;
  const size_t npad = ((size_t)n + 63) / 64 * 64;
  DevBufs B;
  float *df = nullptr, *dy = nullptr;
  bf16* dx = nullptr;
  FC3D_HIP_TRY(B.get(&df, (size_t)n * MLP_NF));
  FC3D_HIP_TRY(B.get(&dx, npad * MLP_K1));
  FC3D_HIP_TRY(B.get(&dy, (size_t)n * MLP_NO));
  FC3D_HIP_TRY(hipMemcpyAsync(df, x, sizeof(float) * (size_t)n * MLP_NF, hipMemcpyHostToDevice,
                              m->stream));
  if ((rc = mlp_forward_device(m, n, df, dx, dy, m->stream))) return rc;
  FC3D_HIP_TRY(hipMemcpyAsync(y, dy, sizeof(float) * (size_t)n * MLP_NO, hipMemcpyDeviceToHost,
                              m->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(m->stream));
  return 0;
}

int fc3d_hip_mlp_gradients(fc3d_hip_mlp* m, int n, const float* x, const float* t, float* grads,
                           float* loss) {
  if (!m || !x || !t || n <= 0 || n > 65536) {
    set_error("fc3d_hip_mlp_gradients: need 1 <= n <= 65536 rows and non-NULL x, t");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  int rc = use_device(m->device);
  if (rc) return rc;
  const int bsz = (n + 63) / 64 * 64;
  if ((rc = mlp_workspace(m, bsz))) return rc;
  std::vector<float> tpad((size_t)bsz * MLP_NO, 0.f);
  std::memcpy(tpad.data(), t, sizeof(float) * (size_t)n * MLP_NO);
  DevBufs B;
  float *df = nullptr, *dt = nullptr;
  bf16* dx = nullptr;
  FC3D_HIP_TRY(B.get(&df, (size_t)n * MLP_NF));
  FC3D_HIP_TRY(B.get(&dx, (size_t)bsz * MLP_K1));
  FC3D_HIP_TRY(B.get(&dt, (size_t)bsz * MLP_NO));
  FC3D_HIP_TRY(hipMemcpyAsync(df, x, sizeof(float) * (size_t)n * MLP_NF, hipMemcpyHostToDevice,
                              m->stream));
  FC3D_HIP_TRY(hipMemcpyAsync(dt, tpad.data(), sizeof(float) * tpad.size(), hipMemcpyHostToDevice,
                              m->stream));
  hipLaunchKernelGGL(k_mlp_inputs, dim3((unsigned)(((size_t)bsz * MLP_K1 + 255) / 256)),
                     dim3(256), 0, m->stream, n, bsz, df, m->mean, m->stdv, dx);
  if ((rc = mlp_grad_pass(m, n, bsz, dx, dt))) return rc;
  const double s = sum_loss(m, bsz, &rc);
  if (rc) return rc;
  if (grads) FC3D_HIP_TRY(hipMemcpy(grads, m->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  if (loss) *loss = (float)(s / ((double)n * MLP_NO));
  return 0;
}

int fc3d_hip_mlp_train(fc3d_hip_mlp* m, int n, const float* x, const float* t, int epochs,
                       int batch, float lr, uint64_t seed, int set_normalisation,
                       float* final_loss) {
  if (!m || !x || !t || n <= 0 || epochs < 0 || batch < 64 || batch % 64 || !(lr > 0.f)) {
    set_error("fc3d_hip_mlp_train: invalid arguments (batch must be a positive multiple of 64)");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  int rc = use_device(m->device);
  if (rc) return rc;
  FC3D_HIP_TRY(hipStreamSynchronize(m->stream));
  if (set_normalisation) {
    std::vector<double> mu(MLP_NF, 0.0), var(MLP_NF, 0.0);
    for (int i = 0; i < n; i++)
      for (int k = 0; k < MLP_NF; k++) mu[(size_t)k] += x[(size_t)i * MLP_NF + k];
    for (int k = 0; k < MLP_NF; k++) mu[(size_t)k] /= n;
    for (int i = 0; i < n; i++)
      for (int k = 0; k < MLP_NF; k++) {
        const double d = x[(size_t)i * MLP_NF + k] - mu[(size_t)k];
        var[(size_t)k] += d * d;
      }
    std::vector<float> fm(MLP_NF), fs(MLP_NF);
    for (int k = 0; k < MLP_NF; k++) {
      fm[(size_t)k] = (float)mu[(size_t)k];
      const double sd = std::sqrt(var[(size_t)k] / n);
      fs[(size_t)k] = sd > 1e-6 ? (float)sd : 1.f;
    }
    FC3D_HIP_TRY(hipMemcpy(m->mean, fm.data(), sizeof(float) * MLP_NF, hipMemcpyHostToDevice));
    FC3D_HIP_TRY(hipMemcpy(m->stdv, fs.data(), sizeof(float) * MLP_NF, hipMemcpyHostToDevice));
  }
  if ((rc = mlp_workspace(m, batch))) return rc;
  DevBufs B;
  float *df = nullptr, *dt = nullptr, *bt = nullptr;
  bf16 *dx = nullptr, *bx = nullptr;
  int* didx = nullptr;
  const int steps = std::max(1, n / batch);
  FC3D_HIP_TRY(B.get(&df, (size_t)n * MLP_NF));
  FC3D_HIP_TRY(B.get(&dx, (size_t)n * MLP_K1));
  FC3D_HIP_TRY(B.get(&dt, (size_t)n * MLP_NO));
  FC3D_HIP_TRY(B.get(&bx, (size_t)batch * MLP_K1));
  FC3D_HIP_TRY(B.get(&bt, (size_t)batch * MLP_NO));
  FC3D_HIP_TRY(B.get(&didx, (size_t)steps * batch));
  FC3D_HIP_TRY(hipMemcpyAsync(df, x, sizeof(float) * (size_t)n * MLP_NF, hipMemcpyHostToDevice,
                              m->stream));
  FC3D_HIP_TRY(hipMemcpyAsync(dt, t, sizeof(float) * (size_t)n * MLP_NO, hipMemcpyHostToDevice,
                              m->stream));
  hipLaunchKernelGGL(k_mlp_inputs, dim3((unsigned)(((size_t)n * MLP_K1 + 255) / 256)), dim3(256),
                     0, m->stream, n, n, df, m->mean, m->stdv, dx);
  std::mt19937_64 gen(seed);
  std::vector<int> perm((size_t)n), idx((size_t)steps * batch);
  for (int i = 0; i < n; i++) perm[(size_t)i] = i;
  double last = 0.0;
  const float b1 = 0.9f, b2 = 0.999f;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), gen);
    for (size_t i = 0; i < idx.size(); i++) idx[i] = perm[i % (size_t)n];
    // the stream is idle here (previous epoch ended with a synchronize)
    FC3D_HIP_TRY(hipMemcpy(didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    for (int s = 0; s < steps; s++) {
      hipLaunchKernelGGL(k_mlp_gather, dim3((batch * MLP_K1 + 255) / 256), dim3(256), 0,
                         m->stream, batch, didx + (size_t)s * batch, dx, dt, bx, bt);
      if ((rc = mlp_grad_pass(m, batch, batch, bx, bt))) return rc;
      m->adam_t++;
      const float c1 = 1.f - (float)std::pow((double)b1, (double)m->adam_t);
      const float c2 = 1.f - (float)std::pow((double)b2, (double)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, NPARAM,
                         m->p, m->g, m->m1, m->m2, lr, b1, b2, c1, c2, 1e-8f);
      if ((rc = mlp_cast(m))) return rc;
    }
    // the last minibatch's loss was computed with the weights before its step
    last = sum_loss(m, batch, &rc) / ((double)batch * MLP_NO);
    if (rc) return rc;
  }
  FC3D_HIP_TRY(hipStreamSynchronize(m->stream));
  if (final_loss) *final_loss = (float)last;
  return 0;
}

void fc3d_hip_set_default_model(fc3d_hip_mlp* m) { g_default_model = m; }
fc3d_hip_mlp* fc3d_hip_default_model(void) { return g_default_model; }

}  // extern "C"
