// fc3d_problem.hip — resident fc3d problem and the coloured NSGS solve.
//
// Layout in HBM (one problem):
//   * contacts are grouped by colour (greedy colouring, fc3d_colour.hip) and,
//     inside a colour, by decreasing number of off-diagonal blocks; each colour
//     is padded to whole slices of 64 slots, one wavefront per slice, one lane
//     per contact;
//   * the off-diagonal 3x3 blocks of a slice are stored "sliced-ELL": for the
//     k-th block of every lane, 9 planes of 64 values, so that each load of a
//     wavefront reads 512 contiguous bytes (FP64); column indices likewise,
//     one plane of 64 int32 per k; blocks keep the reference row order;
//   * the diagonal block, q and mu are stored per slice in the same plane form;
//   * reactions and velocities are 4-wide vectors per slot (32 B FP64, 16 B
//     FP32) so a neighbour gather is one aligned load.
// A colour is one launch (one pass over its slices); the sweep ends with a
// one-block kernel that reduces the per-slice partial sums in a fixed order
// and runs the reference's convergence logic on the device, so a graph of
// several sweeps replays without host round trips.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fc3d_device.hpp"
#include "fc3d_internal.hpp"

namespace fc3d {

int colour_on_device(int nc, const int* d_rp, const int* d_ci, int* d_colour, hipStream_t s,
                     int* ncolours_out, std::vector<int>* host_colour);

template <typename T> struct vec4;
template <> struct vec4<double> { using type = double4; };
template <> struct vec4<float> { using type = float4; };

struct DevParams {
  LocalParams lp;
  double tol;           // user tolerance
  double local_tol;     // local tolerance at solve start (FP64 phase)
  double fp32_local_tol;
  double fp32_tol;
  double omega;
  double rho_constant;
  int mode;             // error evaluation (NonSmoothGaussSeidel_options.h:121-140)
  int freq;             // error evaluation frequency
  int max_iter;
  int filter;
  int relaxation;
  int rho_strategy;
};

struct SolveState {
  int iter;       // sweeps done (all phases)
  int stop;       // 1: no further sweep
  int notconv;    // reference hasNotConverged
  int need_full;  // residual pending for this sweep
  int freq;       // current evaluation frequency (ADAPTIVE doubles it)
  int phase;      // 1: FP32 sweeps, 0: FP64 sweeps
  int fp32_iter;
  int pad;
  double incr;
  double full;
  double working_tol;
  double local_tol;
  double norm_q;
};

template <typename T>
struct SweepPtrs {
  const int* len;   // off-diagonal blocks of the slot's row, -1 for padding
  const int* dpos;  // position of the diagonal block in the reference row order
  const int* koff;  // per slice: first k-plane
  const int* kmax;  // per slice: number of k-planes
  const int* col;   // [kplane][64] slot of the neighbour
  const T* vals;    // [kplane][9][64]
  const T* diag;    // [slice][9][64]
  const T* q;       // [slice][3][64]
  const T* mu;      // [slot]
  typename vec4<T>::type* r;
  typename vec4<T>::type* u;
  T* pli;           // per-slot PLI step (persistent local state)
  double* part;     // per-slice partial sums
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One colour of one sweep: fc3d_nsgs.c:737-822 for the contacts of slices
// [s_begin, s_end), which share no body.
template <typename T>
__global__ __launch_bounds__(256) void k_sweep(SweepPtrs<T> P, const DevParams* __restrict__ prm,
                                               const SolveState* __restrict__ st, int s_begin,
                                               int s_end, int want_phase) {
  if (st->stop || st->phase != want_phase) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = s_begin + blockIdx.x * 4 + wave;
  if (s >= s_end) return;
  const int slot = s * 64 + lane;
  const int len = P.len[slot];
  double light = 0.0, nr2 = 0.0;
  if (len >= 0) {
    T ql[3];
#pragma unroll
    for (int e = 0; e < 3; e++) ql[e] = P.q[(size_t)(s * 3 + e) * 64 + lane];
    const int koff = P.koff[s], kmax = P.kmax[s];
    for (int k = 0; k < kmax; k++) {
      if (k < len) {
        const size_t plane = (size_t)(koff + k);
        const int j = P.col[plane * 64 + lane];
        T w[9];
#pragma unroll
        for (int e = 0; e < 9; e++) w[e] = P.vals[(plane * 9 + e) * 64 + lane];
        const typename vec4<T>::type rj = P.r[j];
        const T v[3] = {rj.x, rj.y, rj.z};
        mvp3x3<T>(w, v, ql);
      }
    }
    LocalProblem<T> lp;
#pragma unroll
    for (int e = 0; e < 9; e++) lp.W[e] = P.diag[(size_t)(s * 9 + e) * 64 + lane];
    lp.q[0] = ql[0]; lp.q[1] = ql[1]; lp.q[2] = ql[2];
    lp.mu = P.mu[slot];
    const typename vec4<T>::type ro = P.r[slot];
    const T rold[3] = {ro.x, ro.y, ro.z};
    T lr[3] = {ro.x, ro.y, ro.z};
    T rho[3];
    compute_rho<T>(prm->rho_strategy, T(prm->rho_constant), lp.W, rho);
    T pli = P.pli[slot];
    T res = T(0);
    const T tol = T(want_phase ? prm->fp32_local_tol : st->local_tol);
    local_solve<T>(lp, lr, prm->lp, tol, &pli, rho, &res);
    if (prm->relaxation) {
      const T om = T(prm->omega);
#pragma unroll
      for (int e = 0; e < 3; e++) lr[e] = om * lr[e] + (T(1) - om) * rold[e];
    }
    const T d0 = rold[0] - lr[0], d1 = rold[1] - lr[1], d2 = rold[2] - lr[2];
    light = double(d0 * d0 + d1 * d1 + d2 * d2);
    const bool accept = !(prm->filter && (isnan(res) || isinf(res) || res > T(1)));
    typename vec4<T>::type out = ro;
    if (accept) { out.x = lr[0]; out.y = lr[1]; out.z = lr[2]; }
    P.r[slot] = out;
    P.pli[slot] = pli;
    nr2 = double(out.x) * double(out.x) + double(out.y) * double(out.y) +
          double(out.z) * double(out.z);
    if (want_phase == 0) nr2 = double(out.x * out.x + out.y * out.y + out.z * out.z);
  }
  light = wave_sum(light);
  nr2 = wave_sum(nr2);
  if (lane == 0) {
    P.part[2 * (size_t)s] = light;
    P.part[2 * (size_t)s + 1] = nr2;
  }
}

// fc3d_compute_error.c:70-104 per slot: u = q + W r (blocks in the reference
// row order, the diagonal at its own position), then the local term.
__global__ __launch_bounds__(256) void k_residual(SweepPtrs<double> P, const SolveState* st,
                                                  int nslices, int force, const double* x_in,
                                                  int gemv_only) {
  if (!force && !st->need_full) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + wave;
  if (s >= nslices) return;
  const int slot = s * 64 + lane;
  const int len = P.len[slot];
  double e2 = 0.0, r2 = 0.0, u2 = 0.0;
  if (len >= 0) {
    double u[3];
    if (gemv_only) {
      const double4 y0 = P.u[slot];
      u[0] = y0.x; u[1] = y0.y; u[2] = y0.z;
    } else {
#pragma unroll
      for (int e = 0; e < 3; e++) u[e] = P.q[(size_t)(s * 3 + e) * 64 + lane];
    }
    const double4* X = gemv_only ? reinterpret_cast<const double4*>(x_in) : P.r;
    const double4 rs = X[slot];
    const double rv[3] = {rs.x, rs.y, rs.z};
    const int dpos = P.dpos[slot];
    const int koff = P.koff[s], kmax = P.kmax[s];
    double dg[9];
#pragma unroll
    for (int e = 0; e < 9; e++) dg[e] = P.diag[(size_t)(s * 9 + e) * 64 + lane];
    for (int k = 0; k < kmax; k++) {
      if (k == dpos) mvp3x3<double>(dg, rv, u);
      if (k < len) {
        const size_t plane = (size_t)(koff + k);
        const int j = P.col[plane * 64 + lane];
        double w[9];
#pragma unroll
        for (int e = 0; e < 9; e++) w[e] = P.vals[(plane * 9 + e) * 64 + lane];
        const double4 rj = X[j];
        const double v[3] = {rj.x, rj.y, rj.z};
        mvp3x3<double>(w, v, u);
      }
    }
    if (dpos >= kmax) mvp3x3<double>(dg, rv, u);
    P.u[slot] = make_double4(u[0], u[1], u[2], 0.0);
    if (!gemv_only) {
      e2 = unitary_error<double>(rv, u, P.mu[slot]);
      r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
      u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    }
  }
  if (gemv_only) return;
  e2 = wave_sum(e2);
  r2 = wave_sum(r2);
  u2 = wave_sum(u2);
  if (lane == 0) {
    P.part[3 * (size_t)s] = e2;
    P.part[3 * (size_t)s + 1] = r2;
    P.part[3 * (size_t)s + 2] = u2;
  }
}

// Fixed-order sum of `n` strided partials with one block of 1024 threads.
__device__ double block_sum(const double* part, int n, int stride, int off, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) v += part[(size_t)i * stride + off];
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < 16; w++) t += sh[w];
  return t;  // valid on thread 0
}

// End of a sweep: light error (fc3d_nsgs.c:90-96) and the convergence test of
// the chosen evaluation mode (fc3d_nsgs.c:826-863).
__global__ __launch_bounds__(1024) void k_finalize_sweep(const DevParams* prm, SolveState* st,
                                                         const double* part, int nslices,
                                                         int want_phase) {
  __shared__ double sh[16];
  if (st->stop || st->phase != want_phase) return;
  const double light = block_sum(part, nslices, 2, 0, sh);
  const double nr2 = block_sum(part, nslices, 2, 1, sh);
  if (threadIdx.x != 0) return;
  st->iter += 1;
  double incr = sqrt(light);
  const double nr = sqrt(nr2);
  if (fabs(nr) > DBL_EPSILON) incr /= nr;
  st->incr = incr;
  if (want_phase == 1) {
    // FP32 phase: hand over to FP64 sweeps once the FP32 light error is small
    st->fp32_iter += 1;
    st->need_full = 0;
    if (incr < prm->fp32_tol || st->iter >= prm->max_iter) st->stop = 1;
    return;
  }
  const int mode = prm->mode;
  if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT) {
    st->notconv = incr < st->working_tol ? 0 : 1;
    st->need_full = 0;
  } else if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    if (incr >= st->working_tol) {
      st->notconv = 1;
      st->need_full = 0;
    } else {
      st->need_full = 1;
    }
  } else {
    const int f = st->freq;
    st->need_full = f > 0 ? (st->iter % f == 0) : 1;
    if (!st->need_full) {
      st->full = 1e24;
      st->incr = 1e24;
      st->notconv = 1;
    }
  }
  if (!st->need_full && (st->notconv == 0 || st->iter >= prm->max_iter)) st->stop = 1;
}

// fc3d_compute_error.c:100-115 and, unless `plain`, the tolerance manager
// (tolerance_manager.h:203-225 through fc3d_nsgs.c:482-558).
__global__ __launch_bounds__(1024) void k_finalize_residual(const DevParams* prm, SolveState* st,
                                                            const double* part, int nslices,
                                                            int plain) {
  __shared__ double sh[16];
  if (!plain && !st->need_full) return;
  const double e2 = block_sum(part, nslices, 3, 0, sh);
  const double r2 = block_sum(part, nslices, 3, 1, sh);
  const double u2 = block_sum(part, nslices, 3, 2, sh);
  if (threadIdx.x != 0) return;
  double err = sqrt(e2);
  const double scale = fmax(st->norm_q, fmax(sqrt(r2), sqrt(u2)));
  if (fabs(scale) > DBL_EPSILON) err /= scale;
  st->full = err;
  if (plain) return;
  st->need_full = 0;
  const double user = prm->tol;
  if (prm->mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    const double incr = st->incr;
    if (err <= user) {
      st->notconv = 0;
    } else if (incr < 1e-15 || incr < DBL_EPSILON * 100.0) {
      st->local_tol = fmax(st->local_tol / 100.0, DBL_EPSILON * 1e-6);
      st->notconv = 1;
    } else {
      if (err > 0.0 && incr > 0.0) {
        double nt = incr / err * user;
        const double umin = user * 1e-6, fmin_ = DBL_EPSILON * 1e-6;
        const double mn = umin > fmin_ ? umin : fmin_;
        if (nt < mn) nt = mn;
        if (nt > 0.0 && nt < st->working_tol) st->working_tol = nt;
      }
      st->notconv = 1;
    }
    if (!(st->working_tol > 0.0)) st->stop = 1;
  } else {
    st->incr = err;
    st->notconv = err < st->working_tol ? 0 : 1;
    if (prm->mode == SICONOS_NSGS_ERROR_EVALUATION_ADAPTIVE && st->freq > 0 &&
        err > st->working_tol)
      st->freq *= 2;
  }
  if (st->notconv == 0 || st->iter >= prm->max_iter) st->stop = 1;
}

template <typename T>
__global__ void k_fill(T* p, T v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename S, typename D>
__global__ void k_convert(const S* a, D* b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = D(a[i]);
}

__global__ void k_r64_to_r32(const double4* a, float4* b, const double* pli, float* pli32,
                             size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 x = a[i];
  b[i] = make_float4(float(x.x), float(x.y), float(x.z), 0.f);
  pli32[i] = float(pli[i]);
}

__global__ void k_r32_to_r64(const float4* a, double4* b, const float* pli32, double* pli,
                             size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = a[i];
  b[i] = make_double4(double(x.x), double(x.y), double(x.z), 0.0);
  pli[i] = double(pli32[i]);
}

}  // namespace fc3d

using namespace fc3d;

struct fc3d_hip_problem {
  int device = 0;
  int nc = 0;
  int64_t nnzb = 0;
  int ncolours = 0;
  int nslices = 0;
  int nslots = 0;
  int64_t nplanes = 0;
  double norm_q = 0.0;
  hipStream_t stream = nullptr;
  std::vector<int> colour;            // per original contact
  std::vector<int> colour_slice_ptr;  // ncolours + 1
  std::vector<int> slot2orig;         // -1 for padding
  std::vector<int> orig2slot;
  std::vector<double> colour_bytes64, colour_bytes32;  // algorithmic bytes per colour pass
  // device
  int *len = nullptr, *dpos = nullptr, *koff = nullptr, *kmax = nullptr, *col = nullptr;
  double *vals = nullptr, *diag = nullptr, *q = nullptr, *mu = nullptr, *pli = nullptr;
  double4 *r = nullptr, *u = nullptr, *x = nullptr;
  float *vals32 = nullptr, *diag32 = nullptr, *q32 = nullptr, *mu32 = nullptr, *pli32 = nullptr;
  float4* r32 = nullptr;
  double *part = nullptr, *part2 = nullptr;
  DevParams* prm = nullptr;
  SolveState* st = nullptr;
  // captured sweep chunks: [phase] -> graph of `graph_sweeps` sweeps
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_sweeps[2] = {0, 0};
};

namespace fc3d {

template <typename T>
static int dev_alloc(T** p, size_t n) {
  FC3D_HIP_TRY(hipMalloc(p, sizeof(T) * (n ? n : 1)));
  return 0;
}
template <typename T>
static int dev_upload(T** p, const std::vector<T>& h) {
  int rc = dev_alloc(p, h.size());
  if (rc) return rc;
  if (!h.empty()) FC3D_HIP_TRY(hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return 0;
}

static SweepPtrs<double> ptrs64(fc3d_hip_problem* h, double* part) {
  SweepPtrs<double> P;
  P.len = h->len; P.dpos = h->dpos; P.koff = h->koff; P.kmax = h->kmax; P.col = h->col;
  P.vals = h->vals; P.diag = h->diag; P.q = h->q; P.mu = h->mu; P.r = h->r; P.u = h->u;
  P.pli = h->pli; P.part = part;
  return P;
}
static SweepPtrs<float> ptrs32(fc3d_hip_problem* h) {
  SweepPtrs<float> P;
  P.len = h->len; P.dpos = h->dpos; P.koff = h->koff; P.kmax = h->kmax; P.col = h->col;
  P.vals = h->vals32; P.diag = h->diag32; P.q = h->q32; P.mu = h->mu32; P.r = h->r32;
  P.u = nullptr; P.pli = h->pli32; P.part = h->part;
  return P;
}

static int ensure_fp32(fc3d_hip_problem* h) {
  if (h->vals32) return 0;
  const size_t nv = (size_t)h->nplanes * 9 * 64, nd = (size_t)h->nslices * 9 * 64,
               nq = (size_t)h->nslices * 3 * 64, ns = (size_t)h->nslots;
  int rc;
  if ((rc = dev_alloc(&h->vals32, nv)) || (rc = dev_alloc(&h->diag32, nd)) ||
      (rc = dev_alloc(&h->q32, nq)) || (rc = dev_alloc(&h->mu32, ns)) ||
      (rc = dev_alloc(&h->r32, ns)) || (rc = dev_alloc(&h->pli32, ns)))
    return rc;
  auto conv = [&](const double* a, float* b, size_t n) {
    if (n) hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((n + 255) / 256)),
                              dim3(256), 0, h->stream, a, b, n);
  };
  conv(h->vals, h->vals32, nv);
  conv(h->diag, h->diag32, nd);
  conv(h->q, h->q32, nq);
  conv(h->mu, h->mu32, ns);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

// Enqueues one sweep of `phase` (1 FP32, 0 FP64) plus its convergence step.
static void enqueue_sweep(fc3d_hip_problem* h, int phase, hipEvent_t* ev, int nev) {
  for (int c = 0; c < h->ncolours; c++) {
    const int sb = h->colour_slice_ptr[(size_t)c], se = h->colour_slice_ptr[(size_t)c + 1];
    const int grid = (se - sb + 3) / 4;
    if (ev && 2 * c + 1 < nev) hipEventRecord(ev[2 * c], h->stream);
    if (phase)
      hipLaunchKernelGGL(k_sweep<float>, dim3(grid), dim3(256), 0, h->stream, ptrs32(h), h->prm,
                         h->st, sb, se, 1);
    else
      hipLaunchKernelGGL(k_sweep<double>, dim3(grid), dim3(256), 0, h->stream,
                         ptrs64(h, h->part), h->prm, h->st, sb, se, 0);
    if (ev && 2 * c + 1 < nev) hipEventRecord(ev[2 * c + 1], h->stream);
  }
  hipLaunchKernelGGL(k_finalize_sweep, dim3(1), dim3(1024), 0, h->stream, h->prm, h->st, h->part,
                     h->nslices, phase);
  if (!phase) {
    hipLaunchKernelGGL(k_residual, dim3((h->nslices + 3) / 4), dim3(256), 0, h->stream,
                       ptrs64(h, h->part2), h->st, h->nslices, 0, nullptr, 0);
    hipLaunchKernelGGL(k_finalize_residual, dim3(1), dim3(1024), 0, h->stream, h->prm, h->st,
                       h->part2, h->nslices, 0);
  }
}

static int build_graph(fc3d_hip_problem* h, int phase, int sweeps) {
  if (h->graph[phase] && h->graph_sweeps[phase] == sweeps) return 0;
  if (h->graph[phase]) {
    hipGraphExecDestroy(h->graph[phase]);
    h->graph[phase] = nullptr;
  }
  hipGraph_t g = nullptr;
  FC3D_HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < sweeps; k++) enqueue_sweep(h, phase, nullptr, 0);
  FC3D_HIP_TRY(hipStreamEndCapture(h->stream, &g));
  FC3D_HIP_TRY(hipGraphInstantiate(&h->graph[phase], g, nullptr, nullptr, 0));
  hipGraphDestroy(g);
  h->graph_sweeps[phase] = sweeps;
  return 0;
}

static int final_residual(fc3d_hip_problem* h) {
  hipLaunchKernelGGL(k_residual, dim3((h->nslices + 3) / 4), dim3(256), 0, h->stream,
                     ptrs64(h, h->part2), h->st, h->nslices, 1, nullptr, 0);
  hipLaunchKernelGGL(k_finalize_residual, dim3(1), dim3(1024), 0, h->stream, h->prm, h->st,
                     h->part2, h->nslices, 1);
  FC3D_HIP_TRY(hipGetLastError());
  return 0;
}

static DevParams make_params(const fc3d_hip_params* p) {
  DevParams d;
  d.lp.solver = p->local_solver;
  d.lp.newton_max_iter = p->local_max_iter;
  d.lp.formulation = p->local_formulation;
  d.lp.ls_max_iter = p->ls_max_iter;
  d.lp.hybrid_strategy = p->hybrid_strategy;
  d.lp.hybrid_max_loop = p->hybrid_max_loop;
  d.lp.hybrid_max_iter = p->hybrid_max_iter;
  d.tol = p->tol;
  d.local_tol = p->local_tol;
  d.fp32_local_tol = p->fp32_local_tol;
  d.fp32_tol = p->fp32_tol;
  d.omega = p->omega;
  d.rho_constant = p->rho_constant;
  d.mode = p->error_evaluation;
  d.freq = p->error_frequency;
  d.max_iter = p->max_iter;
  d.filter = p->filter_local_solution;
  d.relaxation = p->relaxation;
  d.rho_strategy = p->rho_strategy;
  return d;
}

}  // namespace fc3d

extern "C" {

int fc3d_hip_problem_create(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                            const double* blocks, const double* q, const double* mu, int device,
                            fc3d_hip_problem** out) {
  if (!out) return NUMERICS_ERR_NULL_POINTER;
  *out = nullptr;
  if (nc <= 0 || !row_ptr || !col_idx || !blocks || !q || !mu) {
    set_error("fc3d_hip_problem_create: invalid arguments (nc=%d)", nc);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  const int64_t nnzb = row_ptr[nc];
  if (row_ptr[0] != 0 || nnzb > 0x7fffffffLL) {
    set_error("fc3d_hip_problem_create: row_ptr[0] must be 0 and blocks < 2^31");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  // validate the structure and find each row's diagonal block
  std::vector<int> offlen((size_t)nc), dpos_row((size_t)nc);
  std::vector<int64_t> diag_block((size_t)nc);
  for (int i = 0; i < nc; i++) {
    if (row_ptr[i + 1] < row_ptr[i]) {
      set_error("fc3d_hip_problem_create: row pointers not monotone at row %d", i);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
    int64_t d = -1;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      if (col_idx[k] < 0 || col_idx[k] >= nc) {
        set_error("fc3d_hip_problem_create: column %lld out of range in row %d",
                  (long long)col_idx[k], i);
        return NUMERICS_ERR_INVALID_ARGUMENT;
      }
      if (col_idx[k] == i) {
        if (d >= 0) {
          set_error("fc3d_hip_problem_create: two diagonal blocks in row %d", i);
          return NUMERICS_ERR_INVALID_ARGUMENT;
        }
        d = k;
      }
    }
    if (d < 0) {
      set_error("fc3d_hip_problem_create: row %d has no diagonal block", i);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
    diag_block[(size_t)i] = d;
    dpos_row[(size_t)i] = (int)(d - row_ptr[i]);
    offlen[(size_t)i] = (int)(row_ptr[i + 1] - row_ptr[i] - 1);
  }
  int rc = use_device(device);
  if (rc) return rc;

  fc3d_hip_problem* h = new fc3d_hip_problem();
  h->device = device;
  h->nc = nc;
  h->nnzb = nnzb;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("fc3d_hip_problem_create: stream creation failed");
    delete h;
    return NUMERICS_ERR_DEVICE;
  }
  auto fail = [&](int code) {
    fc3d_hip_problem_destroy(h);
    return code;
  };

  // |q| in the caller's order, as dnrm2 does (fc3d_nsgs.c:571)
  double s = 0.0;
  for (int64_t i = 0; i < 3 * (int64_t)nc; i++) s += q[i] * q[i];
  h->norm_q = std::sqrt(s);

  // colour on the device
  {
    std::vector<int> rp((size_t)nc + 1), ci((size_t)(nnzb > 0 ? nnzb : 1));
    for (int i = 0; i <= nc; i++) rp[(size_t)i] = (int)row_ptr[i];
    for (int64_t k = 0; k < nnzb; k++) ci[(size_t)k] = (int)col_idx[k];
    int *d_rp = nullptr, *d_ci = nullptr, *d_col = nullptr;
    if ((rc = dev_upload(&d_rp, rp)) || (rc = dev_upload(&d_ci, ci)) ||
        (rc = dev_alloc(&d_col, (size_t)nc)))
      return fail(rc);
    rc = colour_on_device(nc, d_rp, d_ci, d_col, h->stream, &h->ncolours, &h->colour);
    hipFree(d_rp);
    hipFree(d_ci);
    hipFree(d_col);
    if (rc) return fail(rc);
  }

  // colour-major order; inside a colour, longest rows first (ties by index)
  std::vector<std::vector<int>> members((size_t)h->ncolours);
  for (int i = 0; i < nc; i++) members[(size_t)h->colour[(size_t)i]].push_back(i);
  h->colour_slice_ptr.assign((size_t)h->ncolours + 1, 0);
  for (int c = 0; c < h->ncolours; c++) {
    auto& m = members[(size_t)c];
    std::stable_sort(m.begin(), m.end(), [&](int a, int b) {
      return offlen[(size_t)a] > offlen[(size_t)b];
    });
    h->colour_slice_ptr[(size_t)c + 1] =
        h->colour_slice_ptr[(size_t)c] + (int)((m.size() + 63) / 64);
  }
  h->nslices = h->colour_slice_ptr[(size_t)h->ncolours];
  h->nslots = h->nslices * 64;
  h->slot2orig.assign((size_t)h->nslots, -1);
  h->orig2slot.assign((size_t)nc, -1);
  h->colour_bytes64.assign((size_t)h->ncolours, 0.0);
  h->colour_bytes32.assign((size_t)h->ncolours, 0.0);
  for (int c = 0; c < h->ncolours; c++) {
    const auto& m = members[(size_t)c];
    for (size_t t = 0; t < m.size(); t++) {
      const int slot = h->colour_slice_ptr[(size_t)c] * 64 + (int)t;
      h->slot2orig[(size_t)slot] = m[t];
      h->orig2slot[(size_t)m[t]] = slot;
      const double kbar = offlen[(size_t)m[t]] + 1.0;
      h->colour_bytes64[(size_t)c] += 76.0 * kbar + 108.0;
      h->colour_bytes32[(size_t)c] += 40.0 * kbar + 56.0;
    }
  }
  std::vector<int> len((size_t)h->nslots, -1), dpos((size_t)h->nslots, 0);
  std::vector<int> koff((size_t)h->nslices), kmax((size_t)h->nslices);
  int64_t planes = 0;
  for (int sl = 0; sl < h->nslices; sl++) {
    int mx = 0;
    for (int l = 0; l < 64; l++) {
      const int o = h->slot2orig[(size_t)sl * 64 + l];
      if (o >= 0) mx = std::max(mx, offlen[(size_t)o]);
    }
    koff[(size_t)sl] = (int)planes;
    kmax[(size_t)sl] = mx;
    planes += mx;
  }
  if (planes * 64 > 0x7fffffffLL) {
    set_error("fc3d_hip_problem_create: sliced layout exceeds int32 indexing");
    return fail(NUMERICS_ERR_INVALID_ARGUMENT);
  }
  h->nplanes = planes;
  std::vector<int> colv((size_t)(planes > 0 ? planes : 1) * 64, 0);
  std::vector<double> vals((size_t)(planes > 0 ? planes : 1) * 9 * 64, 0.0);
  std::vector<double> diag((size_t)h->nslices * 9 * 64, 0.0), qv((size_t)h->nslices * 3 * 64, 0.0);
  std::vector<double> muv((size_t)h->nslots, 0.0);
  for (int slot = 0; slot < h->nslots; slot++) {
    const int i = h->slot2orig[(size_t)slot];
    if (i < 0) continue;
    const int sl = slot / 64, l = slot % 64;
    len[(size_t)slot] = offlen[(size_t)i];
    dpos[(size_t)slot] = dpos_row[(size_t)i];
    int kk = 0;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      if (k == diag_block[(size_t)i]) continue;
      const size_t plane = (size_t)koff[(size_t)sl] + (size_t)kk;
      colv[plane * 64 + (size_t)l] = h->orig2slot[(size_t)col_idx[k]];
      for (int e = 0; e < 9; e++) vals[(plane * 9 + (size_t)e) * 64 + (size_t)l] = blocks[9 * k + e];
      kk++;
    }
    const int64_t d = diag_block[(size_t)i];
    for (int e = 0; e < 9; e++) diag[((size_t)sl * 9 + (size_t)e) * 64 + (size_t)l] = blocks[9 * d + e];
    for (int e = 0; e < 3; e++) qv[((size_t)sl * 3 + (size_t)e) * 64 + (size_t)l] = q[3 * i + e];
    muv[(size_t)slot] = mu[i];
  }
  if ((rc = dev_upload(&h->len, len)) || (rc = dev_upload(&h->dpos, dpos)) ||
      (rc = dev_upload(&h->koff, koff)) || (rc = dev_upload(&h->kmax, kmax)) ||
      (rc = dev_upload(&h->col, colv)) || (rc = dev_upload(&h->vals, vals)) ||
      (rc = dev_upload(&h->diag, diag)) || (rc = dev_upload(&h->q, qv)) ||
      (rc = dev_upload(&h->mu, muv)) || (rc = dev_alloc(&h->r, (size_t)h->nslots)) ||
      (rc = dev_alloc(&h->u, (size_t)h->nslots)) || (rc = dev_alloc(&h->x, (size_t)h->nslots)) ||
      (rc = dev_alloc(&h->pli, (size_t)h->nslots)) ||
      (rc = dev_alloc(&h->part, 2 * (size_t)h->nslices)) ||
      (rc = dev_alloc(&h->part2, 3 * (size_t)h->nslices)) || (rc = dev_alloc(&h->prm, 1)) ||
      (rc = dev_alloc(&h->st, 1)))
    return fail(rc);
  if (hipMemset(h->r, 0, sizeof(double4) * (size_t)h->nslots) != hipSuccess ||
      hipMemset(h->u, 0, sizeof(double4) * (size_t)h->nslots) != hipSuccess) {
    set_error("fc3d_hip_problem_create: memset failed");
    return fail(NUMERICS_ERR_DEVICE);
  }
  *out = h;
  return 0;
}

void fc3d_hip_problem_destroy(fc3d_hip_problem* h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (int k = 0; k < 2; k++)
    if (h->graph[k]) hipGraphExecDestroy(h->graph[k]);
  void* ptrs[] = {h->len,  h->dpos,   h->koff,   h->kmax, h->col,  h->vals,  h->diag,
                  h->q,    h->mu,     h->pli,    h->r,    h->u,    h->x,     h->vals32,
                  h->diag32, h->q32,  h->mu32,   h->pli32, h->r32, h->part,  h->part2,
                  h->prm,  h->st};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

int fc3d_hip_problem_num_colours(const fc3d_hip_problem* h) { return h ? h->ncolours : -1; }
int64_t fc3d_hip_problem_nnzb(const fc3d_hip_problem* h) { return h ? h->nnzb : -1; }

int fc3d_hip_problem_colours(const fc3d_hip_problem* h, int* colour) {
  if (!h || !colour) return NUMERICS_ERR_NULL_POINTER;
  std::memcpy(colour, h->colour.data(), sizeof(int) * (size_t)h->nc);
  return 0;
}

int fc3d_hip_problem_order(const fc3d_hip_problem* h, int* order) {
  if (!h || !order) return NUMERICS_ERR_NULL_POINTER;
  int k = 0;
  for (int slot = 0; slot < h->nslots; slot++)
    if (h->slot2orig[(size_t)slot] >= 0) order[k++] = h->slot2orig[(size_t)slot];
  return 0;
}

int fc3d_hip_set_reaction(fc3d_hip_problem* h, const double* reaction) {
  if (!h || !reaction) return NUMERICS_ERR_NULL_POINTER;
  int rc = use_device(h->device);
  if (rc) return rc;
  std::vector<double4> buf((size_t)h->nslots, make_double4(0, 0, 0, 0));
  for (int i = 0; i < h->nc; i++)
    buf[(size_t)h->orig2slot[(size_t)i]] =
        make_double4(reaction[3 * i], reaction[3 * i + 1], reaction[3 * i + 2], 0.0);
  FC3D_HIP_TRY(hipMemcpy(h->r, buf.data(), sizeof(double4) * buf.size(), hipMemcpyHostToDevice));
  return 0;
}

static int get_vec(const fc3d_hip_problem* h, const double4* d, double* out) {
  int rc = use_device(h->device);
  if (rc) return rc;
  std::vector<double4> buf((size_t)h->nslots);
  FC3D_HIP_TRY(hipMemcpy(buf.data(), d, sizeof(double4) * buf.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < h->nc; i++) {
    const double4 v = buf[(size_t)h->orig2slot[(size_t)i]];
    out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
  }
  return 0;
}

int fc3d_hip_get_reaction(const fc3d_hip_problem* h, double* reaction) {
  if (!h || !reaction) return NUMERICS_ERR_NULL_POINTER;
  return get_vec(h, h->r, reaction);
}

int fc3d_hip_get_velocity(const fc3d_hip_problem* h, double* velocity) {
  if (!h || !velocity) return NUMERICS_ERR_NULL_POINTER;
  return get_vec(h, h->u, velocity);
}

int fc3d_hip_compute_error(fc3d_hip_problem* h, double norm, double* error) {
  if (!h || !error) return NUMERICS_ERR_NULL_POINTER;
  int rc = use_device(h->device);
  if (rc) return rc;
  SolveState s0;
  std::memset(&s0, 0, sizeof s0);
  s0.norm_q = norm < 0 ? h->norm_q : norm;
  FC3D_HIP_TRY(hipMemcpyAsync(h->st, &s0, sizeof s0, hipMemcpyHostToDevice, h->stream));
  if ((rc = final_residual(h))) return rc;
  FC3D_HIP_TRY(hipMemcpyAsync(&s0, h->st, sizeof s0, hipMemcpyDeviceToHost, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  *error = s0.full;
  return 0;
}

int fc3d_hip_gemv(fc3d_hip_problem* h, const double* x, double* y) {
  if (!h || !x || !y) return NUMERICS_ERR_NULL_POINTER;
  int rc = use_device(h->device);
  if (rc) return rc;
  std::vector<double4> bx((size_t)h->nslots, make_double4(0, 0, 0, 0)),
      by((size_t)h->nslots, make_double4(0, 0, 0, 0));
  for (int i = 0; i < h->nc; i++) {
    const size_t s = (size_t)h->orig2slot[(size_t)i];
    bx[s] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0);
    by[s] = make_double4(y[3 * i], y[3 * i + 1], y[3 * i + 2], 0.0);
  }
  FC3D_HIP_TRY(hipMemcpy(h->x, bx.data(), sizeof(double4) * bx.size(), hipMemcpyHostToDevice));
  FC3D_HIP_TRY(hipMemcpy(h->u, by.data(), sizeof(double4) * by.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_residual, dim3((h->nslices + 3) / 4), dim3(256), 0, h->stream,
                     ptrs64(h, h->part2), h->st, h->nslices, 1,
                     reinterpret_cast<const double*>(h->x), 1);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  return get_vec(h, h->u, y);
}

int fc3d_hip_solve(fc3d_hip_problem* h, const fc3d_hip_params* p, fc3d_hip_stats* s) {
  if (!h || !p || !s) return NUMERICS_ERR_NULL_POINTER;
  const int ls = p->local_solver;
  if (!(ls == 565 || ls == 551 || ls == 550 || ls == 553 || ls == 552)) {
    set_error("fc3d_hip_solve: local solver %d not supported", ls);
    return NUMERICS_ERR_INVALID_SOLVER;
  }
  if (p->error_evaluation < 0 || p->error_evaluation > 3) {
    set_error("fc3d_hip_solve: unknown error evaluation %d", p->error_evaluation);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (p->rho_strategy != 0 && p->rho_strategy != 2 && p->rho_strategy != 3 &&
      (ls == 565 || ls == 551 || ls == 550)) {
    set_error("fc3d_hip_solve: rho strategy %d not supported", p->rho_strategy);
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  if (p->warm_start == 2) {
    set_error("fc3d_hip_solve: learned warm start needs a trained model (fc3d_hip_mlp_*)");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  int rc = use_device(h->device);
  if (rc) return rc;
  std::memset(s, 0, sizeof *s);
  s->num_colours = h->ncolours;
  s->warm_start_used = p->warm_start == 1 ? 1 : 0;

  const DevParams dp = make_params(p);
  FC3D_HIP_TRY(hipMemcpyAsync(h->prm, &dp, sizeof dp, hipMemcpyHostToDevice, h->stream));
  if (p->warm_start == 1)
    FC3D_HIP_TRY(hipMemsetAsync(h->r, 0, sizeof(double4) * (size_t)h->nslots, h->stream));
  const size_t ns = (size_t)h->nslots;
  hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, h->stream,
                     h->pli, 1.0, ns);
  const int mixed = p->precision == 1;
  if (mixed && (rc = ensure_fp32(h))) return rc;

  SolveState s0;
  std::memset(&s0, 0, sizeof s0);
  s0.stop = p->max_iter <= 0 ? 1 : 0;
  s0.notconv = 1;
  s0.freq = p->error_frequency;
  s0.phase = mixed ? 1 : 0;
  s0.incr = INFINITY;
  s0.full = INFINITY;
  s0.working_tol = p->tol;
  s0.local_tol = p->local_tol;
  s0.norm_q = h->norm_q;
  FC3D_HIP_TRY(hipMemcpyAsync(h->st, &s0, sizeof s0, hipMemcpyHostToDevice, h->stream));

  hipEvent_t t0, t1;
  FC3D_HIP_TRY(hipEventCreate(&t0));
  FC3D_HIP_TRY(hipEventCreate(&t1));
  std::vector<hipEvent_t> kev;
  if (p->instrument) {
    kev.resize(2 * (size_t)h->ncolours);
    for (auto& e : kev) FC3D_HIP_TRY(hipEventCreate(&e));
  }
  FC3D_HIP_TRY(hipEventRecord(t0, h->stream));
  const int chunk = p->sweeps_per_launch > 0 ? p->sweeps_per_launch : 1;

  for (int phase = mixed ? 1 : 0; phase >= 0; phase--) {
    if (phase == 0 && mixed) {
      // hand-over: continue in FP64 from the FP32 iterate
      hipLaunchKernelGGL(k_r32_to_r64, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                         h->stream, h->r32, h->r, h->pli32, h->pli, ns);
      SolveState cur;
      FC3D_HIP_TRY(hipMemcpyAsync(&cur, h->st, sizeof cur, hipMemcpyDeviceToHost, h->stream));
      FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
      s->fp32_sweeps = cur.fp32_iter;
      cur.phase = 0;
      cur.stop = cur.iter >= p->max_iter ? 1 : 0;
      FC3D_HIP_TRY(hipMemcpyAsync(h->st, &cur, sizeof cur, hipMemcpyHostToDevice, h->stream));
    } else if (phase == 1) {
      hipLaunchKernelGGL(k_r64_to_r32, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                         h->stream, h->r, h->r32, h->pli, h->pli32, ns);
    }
    // every replay advances the sweep counter or stops; the bound only
    // guards against a launch that silently did not run
    const long max_replays = (long)p->max_iter / (p->instrument ? 1 : chunk) + 4;
    for (long replay = 0;; replay++) {
      SolveState cur;
      if (replay > max_replays) {
        set_error("fc3d_hip_solve: sweep loop did not stop after %ld replays", replay);
        return NUMERICS_ERR_UNKNOWN;
      }
      if (p->instrument) {
        enqueue_sweep(h, phase, kev.data(), (int)kev.size());
        FC3D_HIP_TRY(hipGetLastError());
        FC3D_HIP_TRY(hipMemcpyAsync(&cur, h->st, sizeof cur, hipMemcpyDeviceToHost, h->stream));
        FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
        {
          for (int c = 0; c < h->ncolours; c++) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, kev[2 * (size_t)c], kev[2 * (size_t)c + 1]);
            s->sweep_kernel_ms += ms;
            s->sweep_kernel_launches += 1;
            s->sweep_kernel_bytes +=
                phase ? h->colour_bytes32[(size_t)c] : h->colour_bytes64[(size_t)c];
          }
        }
      } else {
        if ((rc = build_graph(h, phase, chunk))) return rc;
        FC3D_HIP_TRY(hipGraphLaunch(h->graph[phase], h->stream));
        FC3D_HIP_TRY(hipMemcpyAsync(&cur, h->st, sizeof cur, hipMemcpyDeviceToHost, h->stream));
        FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
      }
      if (cur.stop) break;
    }
  }
  // final full residual (fc3d_nsgs.c:867-875) and outputs
  SolveState fin;
  FC3D_HIP_TRY(hipMemcpyAsync(&fin, h->st, sizeof fin, hipMemcpyDeviceToHost, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  const double incr = fin.incr;
  int notconv = fin.notconv;
  double full = fin.full;
  if (p->error_evaluation == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    if ((rc = final_residual(h))) return rc;
    FC3D_HIP_TRY(hipMemcpyAsync(&fin, h->st, sizeof fin, hipMemcpyDeviceToHost, h->stream));
  }
  FC3D_HIP_TRY(hipEventRecord(t1, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  if (p->error_evaluation == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    full = fin.full;
    notconv = full < p->tol ? 0 : 1;
  }
  float ms = 0.f;
  hipEventElapsedTime(&ms, t0, t1);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  for (auto& e : kev) hipEventDestroy(e);
  s->solve_ms = ms;
  s->iterations = fin.iter;
  s->incr_error = incr;
  s->residual = full;
  s->info = notconv;
  if (fin.iter == p->max_iter && notconv > 0) s->info = NUMERICS_ERR_MAX_ITER;
  return 0;
}

}  // extern "C"
