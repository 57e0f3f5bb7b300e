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
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include <hip/hip_ext.h>

#include "fc3d_internal.hpp"
#include "fc3d_accel.hpp"
#include "fc3d_velocity.hpp"
#include "fc3d_focus.hpp"
#include "fc3d_layout.hpp"
#include "fc3d_mlp.hpp"
#include "fc3d_sweep.hpp"
#include "fc3d_fp32_kernels.hpp"

// instantiated in fc3d_fp32.hip (built with FMA contraction)
FC3D_FP32_KERNELS(extern)

namespace fc3d {

int colour_on_device(int nc, const int* d_rp, const int* d_ci, int* d_colour, hipStream_t s,
                     int* ncolours_out, std::vector<int>* host_colour);

// Contact freezing: the number of the slice's contacts frozen after this
// sweep, for the reset rule of the next one (fc3d_nsgs.c:726-735).
template <typename T>
__device__ __forceinline__ void count_frozen(const SweepPtrs<T>& P, int s, int lane, int len) {
  const bool f = len >= 0 && P.frz[s * 64 + lane] > 0;
  const unsigned long long b = __builtin_amdgcn_ballot_w64(f);
  if (lane == 0) P.fcnt[s] = __builtin_popcountll(b);
}

// One colour of one sweep: fc3d_nsgs.c:737-822 for the contacts of slices
// [s_begin, s_end), which share no body.
template <typename T, int U, int PF, int MINW>
__global__ __launch_bounds__(256, MINW) void k_sweep(SweepPtrs<T> P,
                                                     const DevParams* __restrict__ prm,
                                                     const SolveState* __restrict__ st,
                                                     int s_begin, int s_end, int want_phase) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = __builtin_amdgcn_readfirstlane(s_begin + xcd_block() * 4 + wave);
  if (s >= s_end) return;
  // The slot's own loads, the slice's plane range and the step constants are
  // issued before the stop flag is tested, so the flag's latency overlaps
  // theirs (padding slots read their in-bounds planes too); the empty asm
  // keeps the compiler from sinking them below the test.
  const int stop = st->stop, phase = st->phase;
  const StepConsts sc = step_consts<T>(prm);
  const T tol = T(want_phase ? prm->fp32_local_tol : st->local_tol);
  const int len = P.len[s * 64 + lane];
  SlotInputs<T> in;
  load_slot<T>(P, s, lane, &in);
  int koff, kmax;
  slice_planes<T>(P, s, &koff, &kmax);
  step_consts_ready(sc);
  asm volatile("" ::"v"(len), "v"(in.lp.mu), "v"(in.ro.x), "v"(in.pli), "v"(in.lp.W[8]),
               "v"(in.ql[2]), "s"(tol), "s"(koff), "s"(kmax), "s"(stop), "s"(phase));
  if (stop || phase != want_phase) return;
  double light, nr2;
  sweep_slot_loaded<T, U, PF>(P, sc, s, lane, tol, len, in, koff, kmax, &light, &nr2);
  if (P.fcnt) count_frozen(P, s, lane, len);
  light = wave_sum(light);
  nr2 = wave_sum(nr2);
  if (lane == 0) {
    P.part[2 * (size_t)s] = light;
    P.part[2 * (size_t)s + 1] = nr2;
  }
}

// One colour of one sweep with four waves per slice: wave w gathers planes
// w, w + 4, ... of the slice's 64 contacts, waves 1-3 hand their partial sums
// to wave 0 through LDS, and wave 0 finishes the contacts (local solves,
// write-back).  Four times the waves of k_sweep are in flight during the
// gather, at no extra local-solve work.  The off-diagonal sum is taken in
// another order than the reference row order, so this kernel serves only
// the FP32 sweeps.
template <typename T, int U>
__global__ __launch_bounds__(256, 2) void k_sweep_w4(SweepPtrs<T> P,
                                                     const DevParams* __restrict__ prm,
                                                     const SolveState* __restrict__ st,
                                                     int s_begin, int s_end, int want_phase) {
  __shared__ T shy[3][3][64];
  if (st->stop || st->phase != want_phase) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = s_begin + xcd_block();
  if (s >= s_end) return;
  const int len = P.len[s * 64 + lane];
  T y[3] = {T(0), T(0), T(0)};
  gather_planes<T, U>(P, s, lane, len, wave, 4, y);
  if (wave > 0)
#pragma unroll
    for (int e = 0; e < 3; e++) shy[wave - 1][e][lane] = y[e];
  __syncthreads();
  if (wave != 0) return;
  double light = 0.0, nr2 = 0.0;
  if (len >= 0) {
    SlotInputs<T> in;
    load_slot<T>(P, s, lane, &in);
#pragma unroll
    for (int e = 0; e < 3; e++)
      in.ql[e] += ((y[e] + shy[0][e][lane]) + shy[1][e][lane]) + shy[2][e][lane];
    const T tol = T(want_phase ? prm->fp32_local_tol : st->local_tol);
    finish_slot<T>(P, prm, s, lane, tol, in, &light, &nr2);
  }
  light = wave_sum(light);
  nr2 = wave_sum(nr2);
  if (lane == 0) {
    P.part[2 * (size_t)s] = light;
    P.part[2 * (size_t)s + 1] = nr2;
  }
}

// Single-launch sweep: every colour in one launch, a slice starting as soon
// as the slices holding its lower-colour neighbours have finished this sweep
// (their reactions are then final for it, and its higher-colour neighbours
// wait for it in turn, so every contact sees exactly the values of the
// colour-by-colour sweep).  Waves take slices from their XCD's queue in
// colour-major order; a wave only ever waits on slices dealt before its own,
// each held by a running wave, so the lowest unfinished slice always
// proceeds.  Completion is published with an agent-scope release (L2
// write-back) and observed with an acquire (L2 invalidate) before the
// neighbour reads.  A wait longer than DF_SPIN_LIMIT polls sets an error
// flag and proceeds, so the launch always drains.
struct DataflowState {
  int epoch;      // mark of the current sweep; only ever grows (starts at 1)
  int error;      // a dependency wait exceeded its bound
  int ticket[8];  // per-queue position of the current sweep
};

struct DFArgs {
  const int* dep_ptr;
  const int* dep_idx;
  const int* qlist;
  int* done;  // per slice: mark of the last sweep it completed
  DataflowState* df;
  int qoff[9];
};

constexpr int DF_SPIN_LIMIT = 1 << 20;

template <typename T, int U, int PF, int MINW>
__global__ __launch_bounds__(256, MINW) void k_sweep_df(SweepPtrs<T> P,
                                                        const DevParams* __restrict__ prm,
                                                        const SolveState* __restrict__ st,
                                                        DFArgs D, int want_phase) {
  if (st->stop || st->phase != want_phase) return;
  const int lane = threadIdx.x & 63;
  const int x = blockIdx.x & 7;
  const int qb = D.qoff[x], qn = D.qoff[x + 1] - qb;
  const int epoch = D.df->epoch;
  const T tol = T(want_phase ? prm->fp32_local_tol : st->local_tol);
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(&D.df->ticket[x], 1);
    t = __shfl(t, 0, 64);
    if (t >= qn) break;
    const int s = __builtin_amdgcn_readfirstlane(D.qlist[qb + t]);
    const int d0 = D.dep_ptr[s], d1 = D.dep_ptr[s + 1];
    for (int b = d0; b < d1; b += 64) {
      const int i = b + lane;
      const int dep = i < d1 ? D.dep_idx[i] : -1;
      for (int spins = 0;; spins++) {
        const bool ok = dep < 0 || __hip_atomic_load(&D.done[dep], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT) >= epoch;
        if (__builtin_amdgcn_ballot_w64(!ok) == 0) break;
        if (spins >= DF_SPIN_LIMIT) {
          if (lane == 0) atomicOr(&D.df->error, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double light, nr2;
    sweep_slot<T, U, PF>(P, prm, s, lane, tol, &light, &nr2);
    if (P.fcnt) count_frozen(P, s, lane, P.len[s * 64 + lane]);
    light = wave_sum(light);
    nr2 = wave_sum(nr2);
    if (lane == 0) {
      P.part[2 * (size_t)s] = light;
      P.part[2 * (size_t)s + 1] = nr2;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0)
      __hip_atomic_store(&D.done[s], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// After each single-launch sweep: next mark, queues rewound.
__global__ void k_df_advance(DataflowState* df) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    df->epoch += 1;
    for (int x = 0; x < 8; x++) df->ticket[x] = 0;
  }
}

// Residual pass (fc3d_compute_error.c:70-104) over all slices, or, with
// gemv_only, y += W x on the resident layout (SparseBlockMatrix.c:360-419).
__global__ __launch_bounds__(256) void k_residual(SweepPtrs<double> P, const SolveState* st,
                                                  int nslices, int force, const double* x_in,
                                                  int gemv_only) {
  if (!force && !st->need_full) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + wave;
  if (s >= nslices) return;
  if (gemv_only) {
    const int slot = s * 64 + lane;
    const int len = P.len[slot];
    if (len < 0) return;
    const double4 y0 = P.u[slot];
    double u[3] = {y0.x, y0.y, y0.z};
    const double4* X = reinterpret_cast<const double4*>(x_in);
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
    return;
  }
  double e2, r2, u2;
  residual_slot(P, s, lane, &e2, &r2, &u2);
  e2 = wave_sum(e2);
  r2 = wave_sum(r2);
  u2 = wave_sum(u2);
  if (lane == 0) {
    P.part[3 * (size_t)s] = e2;
    P.part[3 * (size_t)s + 1] = r2;
    P.part[3 * (size_t)s + 2] = u2;
  }
}

// End of a sweep: light error (fc3d_nsgs.c:90-96) and the convergence test of
// the chosen evaluation mode (fc3d_nsgs.c:826-863).  In the velocity form the
// tail launch does the same work instead, when the sweep has a tail.
// With contact freezing (fcnt not null), the frozen contacts are counted
// too: when at least nc - 1 are frozen, every counter reads 0 in the next
// sweep (fc3d_nsgs.c:726-735).
__global__ __launch_bounds__(1024) void k_finalize_sweep(const DevParams* prm, SolveState* st,
                                                         const double* part, int nslices,
                                                         int want_phase, const int* fcnt) {
  __shared__ double sh[2 * 16];
  __shared__ int nfrozen;
  if (st->stop || st->phase != want_phase) return;
  double sums[2];
  block_sums<2>(part, nslices, sh, sums);
  if (fcnt) {
    if (threadIdx.x == 0) nfrozen = 0;
    __syncthreads();
    int c = 0;
    for (int k = threadIdx.x; k < nslices; k += blockDim.x) c += fcnt[k];
    atomicAdd(&nfrozen, c);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    end_sweep_state(prm, st, sums[0], sums[1], want_phase);
    if (fcnt) st->unfreeze = nfrozen >= prm->nc - 1 ? 1 : 0;
  }
}

// fc3d_compute_error.c:100-115 and, unless `plain`, the tolerance manager
// (tolerance_manager.h:203-225 through fc3d_nsgs.c:482-558).
__global__ __launch_bounds__(1024) void k_finalize_residual(const DevParams* prm, SolveState* st,
                                                            const double* part, int nslices,
                                                            int plain) {
  __shared__ double sh[3 * 16];
  if (!plain && !st->need_full) return;
  double sums[3];
  block_sums<3>(part, nslices, sh, sums);
  const double e2 = sums[0], r2 = sums[1], u2 = sums[2];
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

// Features of the learned warm start (defined with fc3d_hip_problem_features
// in fc3d_hip.h) and, when tgt is not null, the resident reaction scaled by
// W_cc(0,0) / |q_c| as a training target; one thread per slot.
__device__ __forceinline__ void start_scales(const SweepPtrs<double>& P, int s, int lane, double* q,
                                             double* W, double* sc, double* w0) {
#pragma unroll
  for (int e = 0; e < 3; e++) q[e] = P.q[(size_t)(s * 3 + e) * 64 + lane];
#pragma unroll
  for (int e = 0; e < 9; e++) W[e] = P.diag[(size_t)(s * 9 + e) * 64 + lane];
  const double nq = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  *sc = nq > 1e-300 ? nq : 1.0;
  *w0 = W[0] > 1e-300 ? W[0] : 1.0;
}

__global__ void k_features(SweepPtrs<double> P, int nslots, float* __restrict__ feat,
                           float* __restrict__ tgt) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int s = slot >> 6, lane = slot & 63;
  float* f = feat + (size_t)slot * MLP_NF;
  const int len = P.len[slot];
  if (len < 0) {
    for (int k = 0; k < MLP_NF; k++) f[k] = 0.f;
    if (tgt)
      for (int e = 0; e < 3; e++) tgt[(size_t)slot * 3 + e] = 0.f;
    return;
  }
  double q[3], W[9], sc, w0;
  start_scales(P, s, lane, q, W, &sc, &w0);
  double qsum = 0.0, qmax = 0.0, g[3] = {0.0, 0.0, 0.0};
  int ncomp = 0;
  const int koff = P.koff[s];
  for (int k = 0; k < len; k++) {
    const size_t plane = (size_t)(koff + k);
    const int j = P.col[plane * 64 + lane];
    const int sj = j >> 6, lj = j & 63;
    const double qj0 = P.q[(size_t)(sj * 3) * 64 + lj];
    const double qj1 = P.q[(size_t)(sj * 3 + 1) * 64 + lj];
    const double qj2 = P.q[(size_t)(sj * 3 + 2) * 64 + lj];
    const double nqj = sqrt(qj0 * qj0 + qj1 * qj1 + qj2 * qj2);
    qsum += nqj;
    qmax = fmax(qmax, nqj);
    if (qj0 < 0.0) {
      ncomp++;
      const double wjj = P.diag[(size_t)(sj * 9) * 64 + lj];
      const double yj = wjj > 1e-300 ? -qj0 / wjj : 0.0;
      // first column of the column-major block W_cj times (yj, 0, 0)
#pragma unroll
      for (int e = 0; e < 3; e++) g[e] += P.vals[(plane * 9 + e) * 64 + lane] * yj;
    }
  }
  for (int e = 0; e < 3; e++) f[e] = (float)(q[e] / sc);
  for (int e = 0; e < 9; e++) f[3 + e] = (float)(W[e] / w0);
  f[12] = (float)P.mu[slot];
  f[13] = (float)log(w0);
  f[14] = (float)log(sc);
  f[15] = (float)len;
  f[16] = (float)(len > 0 ? qsum / len / sc : 0.0);
  f[17] = (float)(qmax / sc);
  for (int e = 0; e < 3; e++) f[18 + e] = (float)(g[e] / sc);
  f[21] = (float)(len > 0 ? (double)ncomp / len : 0.0);
  if (tgt) {
    const double4 r = P.r[slot];
    tgt[(size_t)slot * 3] = (float)(r.x * w0 / sc);
    tgt[(size_t)slot * 3 + 1] = (float)(r.y * w0 / sc);
    tgt[(size_t)slot * 3 + 2] = (float)(r.z * w0 / sc);
  }
}

// r = projection onto the friction cone of the network's scaled prediction
__global__ void k_apply_start(SweepPtrs<double> P, int nslots, const float* __restrict__ y) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int s = slot >> 6, lane = slot & 63;
  if (P.len[slot] < 0) {
    P.r[slot] = make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  double q[3], W[9], sc, w0;
  start_scales(P, s, lane, q, W, &sc, &w0);
  double r[3];
  for (int e = 0; e < 3; e++) {
    const double v = (double)y[(size_t)slot * 3 + e] * sc / w0;
    r[e] = isfinite(v) ? v : 0.0;
  }
  projection_on_cone<double>(r, P.mu[slot]);
  P.r[slot] = make_double4(r[0], r[1], r[2], 0.0);
}

// In-place value update for an unchanged sparsity: CSR block k goes where the
// layout put it (dst >= 0: vals plane entry, dst < 0: diagonal entry).
__global__ void k_scatter_blocks(int64_t n9, const double* __restrict__ blocks,
                                 const int64_t* __restrict__ dst, double* __restrict__ vals,
                                 double* __restrict__ diag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n9) return;
  const int64_t k = i / 9, e = i % 9;
  const int64_t d = dst[k];
  if (d >= 0)
    vals[d + e * 64] = blocks[i];
  else
    diag[(-1 - d) + e * 64] = blocks[i];
}

__global__ void k_scatter_contacts(int nc, const double* __restrict__ q,
                                   const double* __restrict__ mu, const int* __restrict__ o2s,
                                   double* __restrict__ Q, double* __restrict__ MU) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const int slot = o2s[i], s = slot >> 6, l = slot & 63;
  if (q)
    for (int e = 0; e < 3; e++) Q[(size_t)(s * 3 + e) * 64 + l] = q[3 * (size_t)i + e];
  if (mu) MU[slot] = mu[i];
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
  int* d_cps = nullptr;               // the same on the device
  std::vector<int> slot2orig;         // -1 for padding
  std::vector<int> orig2slot;
  std::vector<double> colour_bytes64, colour_bytes32;  // algorithmic bytes per colour pass
  std::vector<double> colour_nc;                       // contacts per colour
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
  int graph_schedule[2] = {0, 0};
  int tuned_schedule[2] = {0, 0};  // autotuner's choice per phase (0: not tuned)
  // learned warm start (model not owned)
  fc3d_hip_mlp* model = nullptr;
  float *feat = nullptr, *ypred = nullptr;
  unsigned short* featX = nullptr;
  double4* rbak = nullptr;
  double* plibak = nullptr;
  // Anderson acceleration (allocated on first use)
  int accel_on = 0;
  // the FP32 sweeps' local solver keeps no PLI state (550, 551, 552): their
  // kernels neither load nor store it, 1% a solve on the 1M pile
  int pli32_off = 0, graph_pli32_off[2] = {0, 0};
  int aa_nblocks = 0;
  AccelBufs<double> aa64 = {};
  AccelBufs<float> aa32 = {};
  AccelState* ast = nullptr;
  int graph_accel[2] = {0, 0};
  // single-launch sweep (schedule 7), set up on first use
  DataflowHost dfh;
  int df_ready = 0;
  int df_grid = 0;
  int *dep_ptr = nullptr, *dep_idx = nullptr, *qlist = nullptr, *done = nullptr;
  DataflowState* df = nullptr;
  // in-place updates: CSR block -> layout position, uploaded on first use
  std::vector<int64_t> block_dst;
  int64_t* d_block_dst = nullptr;
  int* d_o2s = nullptr;
  // velocity form (fc3d_hip_problem_set_bodies)
  // minv_iso: every body's diagonal M^-1 is (m_t, m_t, m_t, m_r, m_r, m_r);
  // the FP32 copy then keeps 2 values a body
  int minv_iso = 0;
  int nbodies = 0, minv_diag = 1, vel_on = 0, graph_vel[2] = {0, 0}, graph_tail[2] = {0, 0};
  int *d_ba = nullptr, *d_bb = nullptr, *d_inc_ptr = nullptr, *d_inc = nullptr;
  double *d_H = nullptr, *d_minv = nullptr, *d_vfree = nullptr;
  double* d_F = nullptr;  // frames layout (rigid-body blocks), instead of d_H
  float* d_F32 = nullptr;
  int frames = 0;
  double4* d_V = nullptr;
  float *d_H32 = nullptr, *d_minv32 = nullptr;
  float4* d_V32 = nullptr;
  std::vector<double> vel_bytes64, vel_bytes32;  // per colour
  // contact freezing (allocated on first use): per-slot counters and last
  // light-error terms, per-slice frozen counts; freeze = 0 when off
  int freeze = 0, graph_freeze[2] = {0, 0};
  double freeze_tmp = 1.0;
  int* d_frz = nullptr;
  double* d_le2 = nullptr;
  int* d_fcnt = nullptr;
  // focused passes (velocity form, allocated on first use)
  int focus_passes = 0, graph_focus[2] = {0, 0};
  int focus_capc = 0;
  FocusState* d_fs = nullptr;
  int *d_bmark = nullptr, *d_slice_colour = nullptr, *d_focus_cnt = nullptr,
      *d_focus_list = nullptr;
  float4* d_fdiff32 = nullptr;
  double4* d_fdiff64 = nullptr;
  // staged passes (diagonal M^-1, at most FOCUS_MAX_COLOURS colours)
  int focus_staged = 0;
  FocusStage<float> stage32 = {};
  FocusStage<double> stage64 = {};
  unsigned long long* d_bkey = nullptr;  // per body: round and local id (FocusStage::bkey)
  int focus_lds = 1, graph_focus_lds[2] = {0, 0};  // 0: the passes keep velocities in HBM
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
  if (h->freeze > 0) {
    P.frz = h->d_frz; P.le2 = h->d_le2; P.fcnt = h->d_fcnt; P.fst = h->st;
    P.freeze = h->freeze; P.freeze_tmp = h->freeze_tmp;
  }
  return P;
}
static SweepPtrs<float> ptrs32(fc3d_hip_problem* h) {
  SweepPtrs<float> P;
  P.len = h->len; P.dpos = h->dpos; P.koff = h->koff; P.kmax = h->kmax; P.col = h->col;
  P.vals = h->vals32; P.diag = h->diag32; P.q = h->q32; P.mu = h->mu32; P.r = h->r32;
  P.u = nullptr; P.pli = h->pli32_off ? nullptr : h->pli32; P.part = h->part;
  return P;
}

static BodyPtrs<double> bodies64(fc3d_hip_problem* h) {
  BodyPtrs<double> B;
  B.ba = h->d_ba; B.bb = h->d_bb; B.H = h->d_H; B.F = h->d_F; B.minv = h->d_minv;
  B.minv_diag = h->minv_diag;
  B.V = h->d_V; B.vfree = h->d_vfree; B.inc_ptr = h->d_inc_ptr; B.inc = h->d_inc;
  if (h->focus_passes > 0) B.fdiff = h->d_fdiff64;
  return B;
}
static BodyPtrs<float> bodies32(fc3d_hip_problem* h) {
  BodyPtrs<float> B;
  B.ba = h->d_ba; B.bb = h->d_bb; B.H = h->d_H32; B.F = h->d_F32; B.minv = h->d_minv32;
  B.minv_diag = h->minv_diag; B.minv_iso = h->minv_iso; B.V = h->d_V32; B.vfree = h->d_vfree; B.inc_ptr = h->d_inc_ptr;
  B.inc = h->d_inc;
  if (h->focus_passes > 0) B.fdiff = h->d_fdiff32;
  return B;
}

// Focused passes: buffers on first use.  Each colour's list holds up to
// focus_capc contacts (a round that lists more is skipped).
static int ensure_focus(fc3d_hip_problem* h, int mixed) {
  int rc;
  if (!h->d_fs) {
    h->focus_capc = 4096;
    std::vector<int> sc((size_t)h->nslices);
    for (int c = 0; c < h->ncolours; c++)
      for (int sl = h->colour_slice_ptr[(size_t)c]; sl < h->colour_slice_ptr[(size_t)c + 1]; sl++)
        sc[(size_t)sl] = c;
    if ((rc = dev_alloc(&h->d_fs, 1)) || (rc = dev_upload(&h->d_slice_colour, sc)) ||
        (rc = dev_alloc(&h->d_focus_cnt, (size_t)h->ncolours)) ||
        (rc = dev_alloc(&h->d_focus_list, (size_t)h->focus_capc * (size_t)h->ncolours)) ||
        (rc = dev_alloc(&h->d_fdiff64, (size_t)h->nslots)))
      return rc;
    FC3D_HIP_TRY(hipMemsetAsync(h->d_fs, 0, sizeof(FocusState), h->stream));
  }
  if (!h->d_bmark) {  // per body (set_bodies drops it with the bodies)
    if ((rc = dev_alloc(&h->d_bmark, (size_t)h->nbodies))) return rc;
    FC3D_HIP_TRY(hipMemsetAsync(h->d_bmark, 0, sizeof(int) * (size_t)h->nbodies, h->stream));
  }
  if (mixed && !h->d_fdiff32 && (rc = dev_alloc(&h->d_fdiff32, (size_t)h->nslots))) return rc;
  h->focus_staged = h->minv_diag && h->ncolours <= FOCUS_MAX_COLOURS;
  if (h->focus_staged) {
    if (!h->d_bkey) {
      if ((rc = dev_alloc(&h->d_bkey, (size_t)h->nbodies))) return rc;
      FC3D_HIP_TRY(hipMemsetAsync(h->d_bkey, 0, sizeof(unsigned long long) * (size_t)h->nbodies,
                                  h->stream));
    }
    const int N = h->focus_capc * h->ncolours;
    auto alloc_stage = [&](auto& S, auto* vt) -> int {
      using TT = std::remove_pointer_t<decltype(vt)>;
      if (S.slot) return 0;
      S.N = N;
      int r;
      if ((r = dev_alloc(&S.slot, (size_t)N)) || (r = dev_alloc(&S.ba, (size_t)N)) ||
          (r = dev_alloc(&S.bb, (size_t)N)) || (r = dev_alloc(&S.H, 36 * (size_t)N)) ||
          (r = dev_alloc(&S.W, 9 * (size_t)N)) || (r = dev_alloc(&S.mu, (size_t)N)) ||
          (r = dev_alloc(&S.minv, 12 * (size_t)N)) || (r = dev_alloc(&S.r, (size_t)N)) ||
          (r = dev_alloc(&S.pli, (size_t)N)) || (r = dev_alloc(&S.la, (size_t)N)) ||
          (r = dev_alloc(&S.lb, (size_t)N)) || (r = dev_alloc(&S.blist, 2 * (size_t)N)))
        return r;
      S.capl = 2 * N;
      (void)sizeof(TT);
      return 0;
    };
    if ((rc = alloc_stage(h->stage64, (double*)nullptr))) return rc;
    if (mixed && (rc = alloc_stage(h->stage32, (float*)nullptr))) return rc;
    h->stage64.bkey = h->d_bkey;
    h->stage32.bkey = h->d_bkey;
  }
  return 0;
}

static void free_stage_bufs(fc3d_hip_problem* h) {
  void* p[] = {h->stage64.slot, h->stage64.ba,  h->stage64.bb, h->stage64.H,  h->stage64.W,
               h->stage64.mu,   h->stage64.minv, h->stage64.r, h->stage64.pli, h->stage64.la,
               h->stage64.lb,   h->stage64.blist, h->stage32.slot, h->stage32.ba, h->stage32.bb,
               h->stage32.H,    h->stage32.W,    h->stage32.mu, h->stage32.minv, h->stage32.r,
               h->stage32.pli,  h->stage32.la,   h->stage32.lb, h->stage32.blist, h->d_bkey};
  for (void* q : p)
    if (q) hipFree(q);
}

// A focus round after the sweeps just enqueued (fc3d_focus.hpp); every kernel
// returns at once unless a round is due.
static void enqueue_focus(fc3d_hip_problem* h, int phase) {
  const unsigned nb = (unsigned)((h->nslots + 255) / 256);
  const int nbh = std::max(1, std::min(256, (h->nslots + 1023) / 1024));
  hipLaunchKernelGGL(k_focus_clear, dim3(1), dim3(256), 0, h->stream, h->prm, h->st, h->d_fs,
                     h->d_focus_cnt, h->ncolours, phase);
  if (phase)
    hipLaunchKernelGGL(k_focus_hist<float>, dim3(nbh), dim3(1024), 0, h->stream, h->prm, h->st,
                       h->d_fs, h->len, h->d_fdiff32, h->nslots, phase);
  else
    hipLaunchKernelGGL(k_focus_hist<double>, dim3(nbh), dim3(1024), 0, h->stream, h->prm, h->st,
                       h->d_fs, h->len, h->d_fdiff64, h->nslots, phase);
  hipLaunchKernelGGL(k_focus_threshold, dim3(1), dim3(FOCUS_SCAN), 0, h->stream, h->prm, h->st,
                     h->d_fs, phase);
  hipLaunchKernelGGL(k_focus_mark, dim3(nb), dim3(256), 0, h->stream, h->d_fs, h->len, h->d_ba,
                     h->d_bb, phase ? h->d_fdiff32 : nullptr, phase ? nullptr : h->d_fdiff64,
                     h->d_bmark, h->nslots);
  hipLaunchKernelGGL(k_focus_collect, dim3(nb), dim3(256), 0, h->stream, h->d_fs, h->len,
                     h->d_ba, h->d_bb, h->d_bmark, h->d_slice_colour, h->d_focus_cnt,
                     h->d_focus_list, h->focus_capc, h->nslots);
  if (h->focus_staged) {
    const dim3 gs(64), bs(256);
    if (phase) {
      BodyPtrs<float> B = bodies32(h);
      B.fdiff = nullptr;
      FocusStage<float> S = h->stage32;
      if (h->pli32_off) S.pli = nullptr;
      S.frames = h->frames;  // the FP32 passes read the frames as the sweep does
      const SweepPtrs<float> P = ptrs32(h);
      hipLaunchKernelGGL(k_focus_stage<float>, gs, bs, 0, h->stream, h->d_fs, P, B,
                         h->d_focus_cnt, h->d_focus_list, h->focus_capc, h->ncolours, S);
      hipLaunchKernelGGL(k_focus_passes_staged<float>, dim3(1), dim3(FOCUS_THREADS), 0,
                         h->stream, S, B, h->prm, h->st, h->d_fs, h->d_focus_cnt,
                         h->focus_capc, h->ncolours, h->focus_passes, phase, h->focus_lds);
      hipLaunchKernelGGL(k_focus_unstage<float>, gs, bs, 0, h->stream, h->d_fs, P,
                         h->d_focus_cnt, h->focus_capc, h->ncolours, S);
    } else {
      BodyPtrs<double> B = bodies64(h);
      B.fdiff = nullptr;
      const SweepPtrs<double> P = ptrs64(h, h->part);
      hipLaunchKernelGGL(k_focus_stage<double>, gs, bs, 0, h->stream, h->d_fs, P, B,
                         h->d_focus_cnt, h->d_focus_list, h->focus_capc, h->ncolours,
                         h->stage64);
      hipLaunchKernelGGL(k_focus_passes_staged<double>, dim3(1), dim3(FOCUS_THREADS), 0,
                         h->stream, h->stage64, B, h->prm, h->st, h->d_fs, h->d_focus_cnt,
                         h->focus_capc, h->ncolours, h->focus_passes, phase, h->focus_lds);
      hipLaunchKernelGGL(k_focus_unstage<double>, gs, bs, 0, h->stream, h->d_fs, P,
                         h->d_focus_cnt, h->focus_capc, h->ncolours, h->stage64);
    }
    return;
  }
  if (phase) {
    BodyPtrs<float> B = bodies32(h);
    B.fdiff = nullptr;
    auto k = h->frames ? k_focus_passes<float, true> : k_focus_passes<float, false>;
    hipLaunchKernelGGL(k, dim3(1), dim3(vel_tail_threads<float>()), 0,
                       h->stream, ptrs32(h), B, h->prm, h->st, h->d_fs, h->d_focus_cnt,
                       h->d_focus_list, h->focus_capc, h->ncolours, h->focus_passes, phase);
  } else {
    BodyPtrs<double> B = bodies64(h);
    B.fdiff = nullptr;
    auto k = h->frames ? k_focus_passes<double, true> : k_focus_passes<double, false>;
    hipLaunchKernelGGL(k, dim3(1), dim3(vel_tail_threads<double>()), 0,
                       h->stream, ptrs64(h, h->part), B, h->prm, h->st, h->d_fs,
                       h->d_focus_cnt, h->d_focus_list, h->focus_capc, h->ncolours,
                       h->focus_passes, phase);
  }
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
// Sweep-kernel schedules: unroll U, column prefetch PF, occupancy target.
#define FC3D_NUM_SCHEDULES 10
#define FC3D_SCHEDULE_DATAFLOW 7
#define FC3D_SCHEDULE_W4 10
static const char* const kScheduleNames[FC3D_NUM_SCHEDULES + 1] = {
    "auto", "u4/u8 occ2", "u4/u8 occ3", "u4/u8 occ4", "u4/u8 pf16 occ2", "u4/u8 pf16 occ3",
    "u2/u4 pf8 occ4", "single launch, u4/u8 occ2", "u8/u16 occ1", "u8/u16 pf16 occ1",
    "fp32: four waves per slice, u4"};

// With events, the launch goes through hipExtLaunchKernelGGL, which stamps
// them with the dispatch's own start and end (the times a kernel-trace
// profiler reports) instead of enqueuing separate marker packets.
template <int U64, int U32, int PF, int MINW>
static void launch_sweep_t(fc3d_hip_problem* h, int phase, int grid, int sb, int se,
                           hipEvent_t e0, hipEvent_t e1) {
  if (phase) {
    auto k = k_sweep<float, U32, PF, MINW>;
    SweepPtrs<float> P = ptrs32(h);
    if (h->accel_on) P.f = h->aa32.f;
    if (e0)
      hipExtLaunchKernelGGL(k, dim3(grid), dim3(256), 0, h->stream, e0, e1, 0, P, h->prm,
                            (const SolveState*)h->st, sb, se, 1);
    else
      hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, h->stream, P, h->prm, h->st, sb, se, 1);
  } else {
    auto k = k_sweep<double, U64, PF, MINW>;
    SweepPtrs<double> P = ptrs64(h, h->part);
    if (h->accel_on) P.f = h->aa64.f;
    if (e0)
      hipExtLaunchKernelGGL(k, dim3(grid), dim3(256), 0, h->stream, e0, e1, 0, P, h->prm,
                            (const SolveState*)h->st, sb, se, 0);
    else
      hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, h->stream, P, h->prm, h->st, sb, se, 0);
  }
}

static void launch_sweep_w4(fc3d_hip_problem* h, int sb, int se, hipEvent_t e0, hipEvent_t e1) {
  auto k = k_sweep_w4<float, 4>;
  SweepPtrs<float> P = ptrs32(h);
  if (h->accel_on) P.f = h->aa32.f;
  const dim3 g((unsigned)(se - sb)), b(256);
  if (e0)
    hipExtLaunchKernelGGL(k, g, b, 0, h->stream, e0, e1, 0, P, h->prm, (const SolveState*)h->st,
                          sb, se, 1);
  else
    hipLaunchKernelGGL(k, g, b, 0, h->stream, P, h->prm, h->st, sb, se, 1);
}

static void launch_sweep(fc3d_hip_problem* h, int phase, int schedule, int grid, int sb, int se,
                         hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (schedule == FC3D_SCHEDULE_W4 && phase) {
    launch_sweep_w4(h, sb, se, e0, e1);
    return;
  }
  switch (schedule) {
    case 2: launch_sweep_t<4, 8, 0, 3>(h, phase, grid, sb, se, e0, e1); break;
    case 3: launch_sweep_t<4, 8, 0, 4>(h, phase, grid, sb, se, e0, e1); break;
    case 4: launch_sweep_t<4, 8, 16, 2>(h, phase, grid, sb, se, e0, e1); break;
    case 5: launch_sweep_t<4, 8, 16, 3>(h, phase, grid, sb, se, e0, e1); break;
    case 6: launch_sweep_t<2, 4, 8, 4>(h, phase, grid, sb, se, e0, e1); break;
    case 8: launch_sweep_t<8, 16, 0, 1>(h, phase, grid, sb, se, e0, e1); break;
    case 9: launch_sweep_t<8, 16, 16, 1>(h, phase, grid, sb, se, e0, e1); break;
    default: launch_sweep_t<4, 8, 0, 2>(h, phase, grid, sb, se, e0, e1); break;  // 1, 7, 10
  }
}

static void launch_sweep_vel(fc3d_hip_problem* h, int phase, int sb, int se, hipEvent_t e0,
                             hipEvent_t e1) {
  const dim3 g((unsigned)((se - sb + 3) / 4)), b(256);
  if (phase) {
    auto k = h->frames ? k_sweep_vel<float, true> : k_sweep_vel<float, false>;
    SweepPtrs<float> P = ptrs32(h);
    if (h->accel_on) P.f = h->aa32.f;
    if (e0)
      hipExtLaunchKernelGGL(k, g, b, 0, h->stream, e0, e1, 0, P, bodies32(h), h->prm,
                            (const SolveState*)h->st, sb, se, 1);
    else
      hipLaunchKernelGGL(k, g, b, 0, h->stream, P, bodies32(h), h->prm, h->st, sb, se, 1);
  } else {
    auto k = h->frames ? k_sweep_vel<double, true> : k_sweep_vel<double, false>;
    SweepPtrs<double> P = ptrs64(h, h->part);
    if (h->accel_on) P.f = h->aa64.f;
    if (e0)
      hipExtLaunchKernelGGL(k, g, b, 0, h->stream, e0, e1, 0, P, bodies64(h), h->prm,
                            (const SolveState*)h->st, sb, se, 0);
    else
      hipLaunchKernelGGL(k, g, b, 0, h->stream, P, bodies64(h), h->prm, h->st, sb, se, 0);
  }
}

// First colour of the velocity-form tail (k_sweep_vel_tail): the run of last
// colours with at most two rounds of the tail block's waves each; ncolours
// if that run is shorter than two colours, or with FC3D_HIP_VEL_TAIL=0 in
// the environment (one launch per colour; the results are the same).
static int vel_tail_begin(const fc3d_hip_problem* h, int phase) {
  const char* env = std::getenv("FC3D_HIP_VEL_TAIL");
  if (env && env[0] == '0') return h->ncolours;
  const int waves = (phase ? vel_tail_threads<float>() : vel_tail_threads<double>()) / 64;
  int c = h->ncolours;
  while (c > 0 && h->colour_slice_ptr[(size_t)c] - h->colour_slice_ptr[(size_t)c - 1] <= 2 * waves)
    c--;
  return h->ncolours - c >= 2 ? c : h->ncolours;
}

static void launch_sweep_vel_tail(fc3d_hip_problem* h, int phase, int c0, hipEvent_t e0,
                                  hipEvent_t e1) {
  const dim3 g(1);
  if (phase) {
    auto k = h->frames ? k_sweep_vel_tail<float, true> : k_sweep_vel_tail<float, false>;
    const dim3 b(vel_tail_threads<float>());
    SweepPtrs<float> P = ptrs32(h);
    if (h->accel_on) P.f = h->aa32.f;
    if (e0)
      hipExtLaunchKernelGGL(k, g, b, 0, h->stream, e0, e1, 0, P, bodies32(h), h->prm, h->st,
                            (const int*)h->d_cps, c0, h->ncolours, 1, 1, h->nslices);
    else
      hipLaunchKernelGGL(k, g, b, 0, h->stream, P, bodies32(h), h->prm, h->st, h->d_cps, c0,
                         h->ncolours, 1, 1, h->nslices);
  } else {
    auto k = h->frames ? k_sweep_vel_tail<double, true> : k_sweep_vel_tail<double, false>;
    const dim3 b(vel_tail_threads<double>());
    SweepPtrs<double> P = ptrs64(h, h->part);
    if (h->accel_on) P.f = h->aa64.f;
    if (e0)
      hipExtLaunchKernelGGL(k, g, b, 0, h->stream, e0, e1, 0, P, bodies64(h), h->prm, h->st,
                            (const int*)h->d_cps, c0, h->ncolours, 0, 1, h->nslices);
    else
      hipLaunchKernelGGL(k, g, b, 0, h->stream, P, bodies64(h), h->prm, h->st, h->d_cps, c0,
                         h->ncolours, 0, 1, h->nslices);
  }
}

// Body velocities from the phase's reaction (st / as as in k_vinit).
static void launch_vinit(fc3d_hip_problem* h, int phase, const SolveState* st,
                         const AccelState* as) {
  const dim3 g((unsigned)((h->nbodies + 255) / 256)), b(256);
  if (phase)
    hipLaunchKernelGGL(h->frames ? (k_vinit<float, true>) : (k_vinit<float, false>), g, b, 0,
                       h->stream, bodies32(h), (const float4*)h->r32,
                       h->nbodies, st, as, 1);
  else
    hipLaunchKernelGGL(h->frames ? (k_vinit<double, true>) : (k_vinit<double, false>), g, b, 0,
                       h->stream, bodies64(h), (const double4*)h->r,
                       h->nbodies, st, as, 0);
}

static DFArgs df_args(fc3d_hip_problem* h) {
  DFArgs D;
  D.dep_ptr = h->dep_ptr;
  D.dep_idx = h->dep_idx;
  D.qlist = h->qlist;
  D.done = h->done;
  D.df = h->df;
  for (int x = 0; x < 9; x++) D.qoff[x] = h->dfh.qoff[x];
  return D;
}

static void launch_sweep_df(fc3d_hip_problem* h, int phase, hipEvent_t e0, hipEvent_t e1) {
  const dim3 g((unsigned)h->df_grid), b(256);
  const DFArgs D = df_args(h);
  if (phase) {
    auto k = k_sweep_df<float, 8, 0, 2>;
    SweepPtrs<float> P = ptrs32(h);
    if (h->accel_on) P.f = h->aa32.f;
    if (e0)
      hipExtLaunchKernelGGL(k, g, b, 0, h->stream, e0, e1, 0, P, h->prm, (const SolveState*)h->st,
                            D, 1);
    else
      hipLaunchKernelGGL(k, g, b, 0, h->stream, P, h->prm, h->st, D, 1);
  } else {
    auto k = k_sweep_df<double, 4, 0, 2>;
    SweepPtrs<double> P = ptrs64(h, h->part);
    if (h->accel_on) P.f = h->aa64.f;
    if (e0)
      hipExtLaunchKernelGGL(k, g, b, 0, h->stream, e0, e1, 0, P, h->prm, (const SolveState*)h->st,
                            D, 0);
    else
      hipLaunchKernelGGL(k, g, b, 0, h->stream, P, h->prm, h->st, D, 0);
  }
  hipLaunchKernelGGL(k_df_advance, dim3(1), dim3(64), 0, h->stream, h->df);
}

// Event pairs one timed sweep uses, and the algorithmic bytes of pair c.
static int vel_tail_begin(const fc3d_hip_problem* h, int phase);
static int timed_launches(const fc3d_hip_problem* h, int phase, int schedule) {
  if (h->vel_on) {
    const int ct = vel_tail_begin(h, phase);
    return ct + (ct < h->ncolours ? 1 : 0);
  }
  return schedule == FC3D_SCHEDULE_DATAFLOW ? 1 : h->ncolours;
}
static double timed_bytes_all(const fc3d_hip_problem* h, int phase, int schedule, int c);
// Without the PLI state, an FP32 pass moves 8 bytes a contact less.
static double timed_bytes(const fc3d_hip_problem* h, int phase, int schedule, int c) {
  double t = timed_bytes_all(h, phase, schedule, c);
  if (!phase || !h->pli32_off) return t;
  const size_t nc = h->colour_nc.size();
  size_t end = (size_t)c + 1;
  if (h->vel_on ? c >= vel_tail_begin(h, phase) : schedule == FC3D_SCHEDULE_DATAFLOW) end = nc;
  for (size_t k = h->vel_on || schedule != FC3D_SCHEDULE_DATAFLOW ? (size_t)c : 0; k < end; k++)
    t -= 8.0 * h->colour_nc[k];
  return t;
}
static double timed_bytes_all(const fc3d_hip_problem* h, int phase, int schedule, int c) {
  if (h->vel_on) {
    const std::vector<double>& b = phase ? h->vel_bytes32 : h->vel_bytes64;
    if (c < vel_tail_begin(h, phase)) return b[(size_t)c];
    double t = 0.0;  // the tail launch: all its colours
    for (size_t k = (size_t)c; k < b.size(); k++) t += b[k];
    return t;
  }
  const std::vector<double>& b = phase ? h->colour_bytes32 : h->colour_bytes64;
  if (schedule != FC3D_SCHEDULE_DATAFLOW) return b[(size_t)c];
  double t = 0.0;
  for (double v : b) t += v;
  return t;
}

static void enqueue_sweep(fc3d_hip_problem* h, int phase, int schedule, hipEvent_t* ev,
                          int nev) {
  if (h->vel_on) {
    const int ct = vel_tail_begin(h, phase);
    for (int c = 0; c < ct; c++) {
      const int sb = h->colour_slice_ptr[(size_t)c], se = h->colour_slice_ptr[(size_t)c + 1];
      const bool t = ev && 2 * c + 1 < nev;
      launch_sweep_vel(h, phase, sb, se, t ? ev[2 * c] : nullptr, t ? ev[2 * c + 1] : nullptr);
    }
    if (ct < h->ncolours) {
      // the tail launch also ends the sweep (k_finalize_sweep's work)
      const bool t = ev && 2 * ct + 1 < nev;
      launch_sweep_vel_tail(h, phase, ct, t ? ev[2 * ct] : nullptr, t ? ev[2 * ct + 1] : nullptr);
    }
  } else if (schedule == FC3D_SCHEDULE_DATAFLOW) {
    launch_sweep_df(h, phase, ev && nev >= 2 ? ev[0] : nullptr, ev && nev >= 2 ? ev[1] : nullptr);
  } else {
    for (int c = 0; c < h->ncolours; c++) {
      const int sb = h->colour_slice_ptr[(size_t)c], se = h->colour_slice_ptr[(size_t)c + 1];
      const int grid = (se - sb + 3) / 4;
      if (ev && 2 * c + 1 < nev)
        launch_sweep(h, phase, schedule, grid, sb, se, ev[2 * c], ev[2 * c + 1]);
      else
        launch_sweep(h, phase, schedule, grid, sb, se);
    }
  }
  if (!h->vel_on || vel_tail_begin(h, phase) >= h->ncolours)
    hipLaunchKernelGGL(k_finalize_sweep, dim3(1), dim3(1024), 0, h->stream, h->prm, h->st,
                       h->part, h->nslices, phase,
                       (const int*)(h->freeze > 0 && !phase ? h->d_fcnt : nullptr));
  if (!phase) {
    hipLaunchKernelGGL(k_residual, dim3((h->nslices + 3) / 4), dim3(256), 0, h->stream,
                       ptrs64(h, h->part2), h->st, h->nslices, 0, nullptr, 0);
    hipLaunchKernelGGL(k_finalize_residual, dim3(1), dim3(1024), 0, h->stream, h->prm, h->st,
                       h->part2, h->nslices, 0);
  }
  if (h->accel_on) {
    const dim3 g((unsigned)h->aa_nblocks), b(256);
    const int nv = 2 * h->nbodies;  // velocity form: two vec4 per body
    const dim3 gv((unsigned)std::max(1, std::min((nv + 255) / 256, 1024)));
    if (phase) {
      hipLaunchKernelGGL(k_aa_update<float>, g, b, 0, h->stream, h->r32, h->aa32, h->st, h->ast,
                         h->prm, h->nslots, 1);
      if (h->vel_on)
        hipLaunchKernelGGL(k_aa_update_v<float>, gv, b, 0, h->stream, h->d_V32, h->aa32, h->st,
                           h->ast, nv, 1);
      hipLaunchKernelGGL(k_aa_finalize, dim3(1), dim3(1024), 0, h->stream, h->aa32.part,
                         h->aa_nblocks, h->st, h->ast, h->prm, 1);
      hipLaunchKernelGGL(k_aa_apply<float>, g, b, 0, h->stream, h->r32, h->aa32, h->st, h->ast,
                         h->prm, h->nslots, 1);
    } else {
      hipLaunchKernelGGL(k_aa_update<double>, g, b, 0, h->stream, h->r, h->aa64, h->st, h->ast,
                         h->prm, h->nslots, 0);
      if (h->vel_on)
        hipLaunchKernelGGL(k_aa_update_v<double>, gv, b, 0, h->stream, h->d_V, h->aa64, h->st,
                           h->ast, nv, 0);
      hipLaunchKernelGGL(k_aa_finalize, dim3(1), dim3(1024), 0, h->stream, h->aa64.part,
                         h->aa_nblocks, h->st, h->ast, h->prm, 0);
      hipLaunchKernelGGL(k_aa_apply<double>, g, b, 0, h->stream, h->r, h->aa64, h->st, h->ast,
                         h->prm, h->nslots, 0);
    }
    // an Anderson step moved r: the body velocities take the same step
    if (h->vel_on) {
      if (phase) {
        hipLaunchKernelGGL(k_aa_apply_v<float>, gv, b, 0, h->stream, h->d_V32, h->aa32, h->st,
                           h->ast, h->prm, nv, 1);
      } else {
        hipLaunchKernelGGL(k_aa_apply_v<double>, gv, b, 0, h->stream, h->d_V, h->aa64, h->st,
                           h->ast, h->prm, nv, 0);
      }
    }
  }
}

// Anderson buffers of one phase, allocated once (f zeroed: the sweep never
// writes padding slots).
template <typename T>
static int alloc_accel(fc3d_hip_problem* h, AccelBufs<T>* A) {
  if (A->f) return 0;
  const size_t ns = (size_t)h->nslots;
  using V = typename vec4<T>::type;
  int rc;
  if ((rc = dev_alloc(&A->f, ns)) || (rc = dev_alloc(&A->fprev, ns)) ||
      (rc = dev_alloc(&A->gprev, ns)) || (rc = dev_alloc(&A->dF, ns * AA_MAX)) ||
      (rc = dev_alloc(&A->dG, ns * AA_MAX)) ||
      (rc = dev_alloc(&A->part, (size_t)h->aa_nblocks * 2 * AA_MAX)))
    return rc;
  FC3D_HIP_TRY(hipMemset(A->f, 0, sizeof(V) * ns));
  return 0;
}

// Body-velocity history of one phase (velocity form), sized for the attached
// bodies; dropped with the factors.
template <typename T>
static int alloc_accel_bodies(fc3d_hip_problem* h, AccelBufs<T>* A) {
  if (A->vprev) return 0;
  const size_t nv = 2 * (size_t)h->nbodies;
  int rc;
  if ((rc = dev_alloc(&A->vprev, nv)) || (rc = dev_alloc(&A->dV, nv * AA_MAX))) return rc;
  return 0;
}

static int ensure_accel(fc3d_hip_problem* h, int mixed) {
  h->aa_nblocks = std::min((h->nslots + 255) / 256, 1024);  // grid-stride update
  int rc;
  if (!h->ast && (rc = dev_alloc(&h->ast, 1))) return rc;
  if ((rc = alloc_accel(h, &h->aa64))) return rc;
  if (mixed && (rc = alloc_accel(h, &h->aa32))) return rc;
  return 0;
}

// Isotropic bodies: (m_t, m_r) of each body's diagonal M^-1, in FP32.
static __global__ void k_minv_iso(const double* __restrict__ minv, float* __restrict__ m2,
                                  int nbodies) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbodies) return;
  m2[2 * (size_t)k] = float(minv[6 * (size_t)k]);
  m2[2 * (size_t)k + 1] = float(minv[6 * (size_t)k + 3]);
}

// FP32 copies of the velocity-form factors.
static int ensure_vel32(fc3d_hip_problem* h) {
  if (h->d_H32 || h->d_F32) return 0;
  const size_t nh = (size_t)h->nslices * (h->frames ? 15 : 36) * 64;
  const size_t nm = (size_t)h->nbodies * (h->minv_iso ? 2 : h->minv_diag ? 6 : 36);
  int rc;
  float** dst = h->frames ? &h->d_F32 : &h->d_H32;
  if ((rc = dev_alloc(dst, nh)) || (rc = dev_alloc(&h->d_minv32, nm)) ||
      (rc = dev_alloc(&h->d_V32, 2 * (size_t)h->nbodies)))
    return rc;
  hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((nh + 255) / 256)), dim3(256), 0,
                     h->stream, h->frames ? h->d_F : h->d_H, *dst, nh);
  if (h->minv_iso)
    hipLaunchKernelGGL(k_minv_iso, dim3((unsigned)((h->nbodies + 255) / 256)), dim3(256), 0,
                       h->stream, h->d_minv, h->d_minv32, h->nbodies);
  else
    hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((nm + 255) / 256)), dim3(256),
                       0, h->stream, h->d_minv, h->d_minv32, nm);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

static void free_accel_bodies(fc3d_hip_problem* h) {
  void* ptrs[] = {h->aa64.vprev, h->aa64.dV, h->aa32.vprev, h->aa32.dV};
  for (void* p : ptrs)
    if (p) hipFree(p);
  h->aa64.vprev = h->aa64.dV = nullptr;
  h->aa32.vprev = h->aa32.dV = nullptr;
}

// Frees the attached factors of W and everything sized by the bodies.
static void drop_bodies(fc3d_hip_problem* h) {
  void* old[] = {h->d_ba, h->d_bb, h->d_inc_ptr, h->d_inc, h->d_H, h->d_minv, h->d_vfree, h->d_V,
                 h->d_H32, h->d_minv32, h->d_V32, h->d_F, h->d_F32};
  for (void* q : old)
    if (q) hipFree(q);
  h->d_ba = h->d_bb = h->d_inc_ptr = h->d_inc = nullptr;
  h->d_H = h->d_minv = h->d_vfree = nullptr;
  h->d_V = nullptr;
  h->d_H32 = h->d_minv32 = nullptr;
  h->d_V32 = nullptr;
  h->d_F = nullptr;
  h->d_F32 = nullptr;
  h->frames = 0;
  h->minv_iso = 0;
  h->nbodies = 0;
  free_accel_bodies(h);
  // the focus rounds' per-body arrays are sized by the old body count
  if (h->d_bmark) hipFree(h->d_bmark);
  if (h->d_bkey) hipFree(h->d_bkey);
  h->d_bmark = nullptr;
  h->d_bkey = nullptr;
  h->stage32.bkey = nullptr;
  h->stage64.bkey = nullptr;
  // captured velocity-form sweeps hold the freed pointers
  for (int ph = 0; ph < 2; ph++)
    if (h->graph[ph] && h->graph_vel[ph]) {
      hipGraphExecDestroy(h->graph[ph]);
      h->graph[ph] = nullptr;
    }
}

static void free_accel(fc3d_hip_problem* h) {
  void* ptrs[] = {h->aa64.f,  h->aa64.fprev, h->aa64.gprev, h->aa64.dF, h->aa64.dG,
                  h->aa64.part, h->aa32.f,   h->aa32.fprev, h->aa32.gprev, h->aa32.dF,
                  h->aa32.dG, h->aa32.part, h->ast};
  for (void* p : ptrs)
    if (p) hipFree(p);
  free_accel_bodies(h);
}

static int accel_stats(fc3d_hip_problem* h, fc3d_hip_stats* s) {
  AccelState a;
  FC3D_HIP_TRY(hipMemcpyAsync(&a, h->ast, sizeof a, hipMemcpyDeviceToHost, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  s->accel_steps += a.steps;
  s->accel_restarts += a.resets;
  s->accel_rejects += a.rejects;
  return 0;
}

// Uploads the dependency lists and queues of the single-launch sweep.
static int ensure_dataflow(fc3d_hip_problem* h) {
  if (h->df_ready) return 0;
  int rc;
  int cus = 0;
  FC3D_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
  h->df_grid = std::max(8, (cus * 2) / 8 * 8);  // two 4-wave blocks per CU, a multiple of 8
  if ((rc = dev_upload(&h->dep_ptr, h->dfh.dep_ptr)) ||
      (rc = dev_upload(&h->dep_idx, h->dfh.dep_idx)) ||
      (rc = dev_upload(&h->qlist, h->dfh.qlist)) || (rc = dev_alloc(&h->done, (size_t)h->nslices)) ||
      (rc = dev_alloc(&h->df, 1)))
    return rc;
  FC3D_HIP_TRY(hipMemset(h->done, 0, sizeof(int) * (size_t)h->nslices));
  DataflowState d;
  std::memset(&d, 0, sizeof d);
  d.epoch = 1;
  FC3D_HIP_TRY(hipMemcpy(h->df, &d, sizeof d, hipMemcpyHostToDevice));
  h->df_ready = 1;
  return 0;
}

static int dataflow_error(fc3d_hip_problem* h) {
  if (!h->df_ready) return 0;
  DataflowState d;
  FC3D_HIP_TRY(hipMemcpyAsync(&d, h->df, sizeof d, hipMemcpyDeviceToHost, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  if (d.error) {
    set_error("single-launch sweep: a dependency wait exceeded its bound");
    return NUMERICS_ERR_DEVICE;
  }
  return 0;
}

static int reset_accel(fc3d_hip_problem* h) {
  AccelState a;
  std::memset(&a, 0, sizeof a);
  a.nf_prev = INFINITY;
  FC3D_HIP_TRY(hipMemcpyAsync(h->ast, &a, sizeof a, hipMemcpyHostToDevice, h->stream));
  return 0;
}

static int build_graph(fc3d_hip_problem* h, int phase, int sweeps, int schedule) {
  if (h->graph[phase] && h->graph_sweeps[phase] == sweeps && h->graph_schedule[phase] == schedule &&
      h->graph_accel[phase] == h->accel_on && h->graph_vel[phase] == h->vel_on &&
      h->graph_tail[phase] == vel_tail_begin(h, phase) && h->graph_pli32_off[phase] == h->pli32_off &&
      h->graph_freeze[phase] == h->freeze && h->graph_focus[phase] == h->focus_passes &&
      h->graph_focus_lds[phase] == h->focus_lds)
    return 0;
  if (h->graph[phase]) {
    hipGraphExecDestroy(h->graph[phase]);
    h->graph[phase] = nullptr;
  }
  hipGraph_t g = nullptr;
  FC3D_HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < sweeps; k++) enqueue_sweep(h, phase, schedule, nullptr, 0);
  if (h->focus_passes > 0) enqueue_focus(h, phase);
  FC3D_HIP_TRY(hipStreamEndCapture(h->stream, &g));
  FC3D_HIP_TRY(hipGraphInstantiate(&h->graph[phase], g, nullptr, nullptr, 0));
  hipGraphDestroy(g);
  h->graph_sweeps[phase] = sweeps;
  h->graph_schedule[phase] = schedule;
  h->graph_accel[phase] = h->accel_on;
  h->graph_vel[phase] = h->vel_on;
  h->graph_tail[phase] = vel_tail_begin(h, phase);
  h->graph_pli32_off[phase] = h->pli32_off;
  h->graph_freeze[phase] = h->freeze;
  h->graph_focus[phase] = h->focus_passes;
  h->graph_focus_lds[phase] = h->focus_lds;
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

// Learned warm start with its fallback: one sweep of `phase` from the
// fallback start (zero, or with `previous` the resident reaction, i.e. the
// previous time step's solution) and one from the network's start, each
// followed by the FP64 residual; the solve continues from the state with the
// smaller residual (the fallback on a tie or a non-finite residual).  The kept
// sweep counts as the solve's first; the other one is extra work inside the
// solve's time.
// The learned warm start (warm_start 2, 3): the network's start and the
// fallback (zero, or the resident reaction) each get `trial` full sweeps of
// the solve's own kind (rows of W or velocity form, this phase's
// precision), and the one with the lower FP64 residual after them is kept,
// with its solve state, so its sweeps count; the other leg's sweeps are
// reported apart.  The zero start wins ties and non-finite residuals.
static int learned_start(fc3d_hip_problem* h, int phase, int schedule, int trial_sweeps,
                         int previous, SolveState* s0, fc3d_hip_stats* out) {
  const size_t ns = (size_t)h->nslots;
  const unsigned nb = (unsigned)((ns + 255) / 256);
  int rc;
  if (!h->feat) {
    if ((rc = dev_alloc(&h->feat, ns * MLP_NF)) || (rc = dev_alloc(&h->ypred, ns * 3)) ||
        (rc = dev_alloc(&h->featX, (ns + 63) / 64 * 64 * MLP_K1)) ||
        (rc = dev_alloc(&h->rbak, ns)) || (rc = dev_alloc(&h->plibak, ns)))
      return rc;
  }
  if (previous)  // keep the resident reaction for the fallback trial
    FC3D_HIP_TRY(hipMemcpyAsync(h->x, h->r, sizeof(double4) * ns, hipMemcpyDeviceToDevice,
                                h->stream));
  hipLaunchKernelGGL(k_features, dim3(nb), dim3(256), 0, h->stream, ptrs64(h, h->part), h->nslots,
                     h->feat, (float*)nullptr);
  if ((rc = mlp_forward_device(h->model, h->nslots, h->feat, h->featX, h->ypred, h->stream)))
    return rc;
  const int m = trial_sweeps > 0 ? trial_sweeps : 1;
  double res[2];
  SolveState after[2];
  for (int trial = 0; trial < 2; trial++) {
    if (trial == 0 && previous)
      FC3D_HIP_TRY(hipMemcpyAsync(h->r, h->x, sizeof(double4) * ns, hipMemcpyDeviceToDevice,
                                  h->stream));
    else if (trial == 0)
      hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((4 * ns + 255) / 256)), dim3(256), 0,
                         h->stream, reinterpret_cast<double*>(h->r), 0.0, 4 * ns);
    else
      hipLaunchKernelGGL(k_apply_start, dim3(nb), dim3(256), 0, h->stream, ptrs64(h, h->part),
                         h->nslots, (const float*)h->ypred);
    hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, h->stream, h->pli, 1.0, ns);
    FC3D_HIP_TRY(hipMemcpyAsync(h->st, s0, sizeof *s0, hipMemcpyHostToDevice, h->stream));
    if (phase)
      hipLaunchKernelGGL(k_r64_to_r32, dim3(nb), dim3(256), 0, h->stream, h->r, h->r32, h->pli,
                         h->pli32, ns);
    if (h->vel_on) launch_vinit(h, phase, nullptr, nullptr);
    for (int k = 0; k < m; k++) enqueue_sweep(h, phase, schedule, nullptr, 0);
    if (phase)
      hipLaunchKernelGGL(k_r32_to_r64, dim3(nb), dim3(256), 0, h->stream, h->r32, h->r,
                         h->pli32, h->pli, ns);
    if ((rc = final_residual(h))) return rc;
    FC3D_HIP_TRY(hipMemcpyAsync(&after[trial], h->st, sizeof(SolveState), hipMemcpyDeviceToHost,
                                h->stream));
    FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
    res[trial] = after[trial].full;
    if (trial == 0) {
      FC3D_HIP_TRY(hipMemcpyAsync(h->rbak, h->r, sizeof(double4) * ns, hipMemcpyDeviceToDevice,
                                  h->stream));
      FC3D_HIP_TRY(hipMemcpyAsync(h->plibak, h->pli, sizeof(double) * ns,
                                  hipMemcpyDeviceToDevice, h->stream));
    }
  }
  const int keep = res[1] < res[0];
  if (!keep) {
    FC3D_HIP_TRY(hipMemcpyAsync(h->r, h->rbak, sizeof(double4) * ns, hipMemcpyDeviceToDevice,
                                h->stream));
    FC3D_HIP_TRY(hipMemcpyAsync(h->pli, h->plibak, sizeof(double) * ns, hipMemcpyDeviceToDevice,
                                h->stream));
  }
  out->warm_start_used = keep ? 2 : (previous ? 0 : 1);
  out->start_residual_fallback = res[0];
  out->start_residual_learned = res[1];
  out->start_trial_sweeps = after[keep ? 0 : 1].iter;  // the discarded leg's sweeps
  *s0 = after[keep];
  // the kept leg's state: its sweeps count; its residual field held the
  // trial's final residual, which the convergence logic recomputes
  s0->full = INFINITY;
  FC3D_HIP_TRY(hipMemcpyAsync(h->st, s0, sizeof *s0, hipMemcpyHostToDevice, h->stream));
  FC3D_HIP_TRY(hipGetLastError());
  return 0;
}

// Autotuner: times every schedule on two sweeps of this very problem (from
// the resident reaction, restored afterwards) and keeps the fastest.  The
// choice is cached on the handle and, per size class, for the process.
struct TuneKey {
  int log2_slices, colours, phase;
};
static std::vector<std::pair<TuneKey, int>> g_tune_cache;

static int tune_schedule(fc3d_hip_problem* h, int phase, double norm_q, double local_tol,
                         int* best) {
  int lg = 0;
  while ((1 << (lg + 1)) <= h->nslices) lg++;
  for (auto& e : g_tune_cache)
    if (e.first.log2_slices == lg && e.first.colours == h->ncolours && e.first.phase == phase) {
      *best = e.second;
      return 0;
    }
  const size_t ns = (size_t)h->nslots;
  void* saved = phase ? (void*)h->u : (void*)h->x;  // scratch of at least ns * 16 / 32 bytes
  void* live = phase ? (void*)h->r32 : (void*)h->r;
  const size_t bytes = ns * (phase ? sizeof(float4) : sizeof(double4));
  FC3D_HIP_TRY(hipMemcpyAsync(saved, live, bytes, hipMemcpyDeviceToDevice, h->stream));
  std::vector<hipEvent_t> ev(2 * (size_t)h->ncolours);
  for (auto& e : ev) FC3D_HIP_TRY(hipEventCreate(&e));
  SolveState t;
  std::memset(&t, 0, sizeof t);
  t.iter = -1000000;  // never reaches max_iter while tuning
  t.notconv = 1;
  t.phase = phase;
  t.incr = INFINITY;
  t.full = INFINITY;
  t.working_tol = 0.0;  // never converges while tuning
  t.local_tol = local_tol;
  t.norm_q = norm_q;
  double best_ms = 1e300;
  *best = 1;
  for (int sc = 1; sc <= FC3D_NUM_SCHEDULES; sc++) {
    if (sc == FC3D_SCHEDULE_DATAFLOW) {
      int rc = ensure_dataflow(h);
      if (rc) return rc;
    }
    FC3D_HIP_TRY(hipMemcpyAsync(h->st, &t, sizeof t, hipMemcpyHostToDevice, h->stream));
    FC3D_HIP_TRY(hipMemcpyAsync(live, saved, bytes, hipMemcpyDeviceToDevice, h->stream));
    double ms_total = 0.0;
    for (int rep = 0; rep < 3; rep++) {
      enqueue_sweep(h, phase, sc, ev.data(), (int)ev.size());
      FC3D_HIP_TRY(hipGetLastError());
      FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
      if (rep == 0) continue;  // first sweep warms caches and code
      for (int c = 0; c < timed_launches(h, phase, sc); c++) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, ev[2 * (size_t)c], ev[2 * (size_t)c + 1]);
        ms_total += ms;
      }
    }
    if (ms_total < best_ms) {
      best_ms = ms_total;
      *best = sc;
    }
  }
  for (auto& e : ev) hipEventDestroy(e);
  FC3D_HIP_TRY(hipMemcpyAsync(live, saved, bytes, hipMemcpyDeviceToDevice, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  {
    int rc = dataflow_error(h);
    if (rc) return rc;
  }
  g_tune_cache.push_back({{lg, h->ncolours, phase}, *best});
  return 0;
}

// Frames layout detection: every contact's blocks must be H_a = [T | T S(l_a)]
// and H_b = -[T | T S(l_b)] for one 3x3 T and lever arms l_a, l_b (a side
// without a body is free).  l is read off the skew matrix sg T^-1 Rot, and the
// blocks rebuilt the way the kernels rebuild them (frames_to_h) must equal
// the given ones to 1e-13 of the block's largest entry.  On success Fs holds
// [slice][15][64]: T row-major, l_a, l_b.
static bool rigid_frames(int nc, const int* body_a, const int* body_b, const double* H,
                         const std::vector<int>& orig2slot, int nslices, std::vector<double>* Fs) {
  Fs->assign((size_t)nslices * 15 * 64, 0.0);
  for (int c = 0; c < nc; c++) {
    const double* hc = H + 36 * (size_t)c;
    const bool has[2] = {body_a[c] >= 0, body_b[c] >= 0};
    double f[15] = {0};
    // T from side a, or from side b's negated translational block
    const int ref = has[0] ? 0 : has[1] ? 1 : -1;
    if (ref >= 0) {
      const double sg = ref ? -1.0 : 1.0;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) f[3 * i + j] = sg * hc[18 * ref + i + 3 * j];
    }
    double big = 0.0;
    for (int e = 0; e < 36; e++) big = std::max(big, std::fabs(hc[e]));
    // T^-1 by cofactors
    const double* T = f;
    const double det = T[0] * (T[4] * T[8] - T[5] * T[7]) - T[1] * (T[3] * T[8] - T[5] * T[6]) +
                       T[2] * (T[3] * T[7] - T[4] * T[6]);
    if (ref >= 0 && !(std::fabs(det) > 1e-12 * big * big * big)) return false;
    double Ti[9] = {0};
    if (ref >= 0) {
      Ti[0] = (T[4] * T[8] - T[5] * T[7]) / det;
      Ti[1] = (T[2] * T[7] - T[1] * T[8]) / det;
      Ti[2] = (T[1] * T[5] - T[2] * T[4]) / det;
      Ti[3] = (T[5] * T[6] - T[3] * T[8]) / det;
      Ti[4] = (T[0] * T[8] - T[2] * T[6]) / det;
      Ti[5] = (T[2] * T[3] - T[0] * T[5]) / det;
      Ti[6] = (T[3] * T[7] - T[4] * T[6]) / det;
      Ti[7] = (T[1] * T[6] - T[0] * T[7]) / det;
      Ti[8] = (T[0] * T[4] - T[1] * T[3]) / det;
    }
    for (int side = 0; side < 2; side++) {
      if (!has[side]) continue;
      const double sg = side ? -1.0 : 1.0;
      double X[9];  // sg T^-1 Rot, row-major
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double v = 0.0;
          for (int k = 0; k < 3; k++) v += Ti[3 * i + k] * hc[18 * side + k + 3 * (3 + j)];
          X[3 * i + j] = sg * v;
        }
      double* l = f + 9 + 3 * side;
      l[0] = 0.5 * (X[7] - X[5]);
      l[1] = 0.5 * (X[2] - X[6]);
      l[2] = 0.5 * (X[3] - X[1]);
    }
    // rebuild as the kernels do and compare
    double hr[36];
    for (int side = 0; side < 2; side++) {
      const double* l = f + 9 + 3 * side;
      const double S[9] = {0.0, -l[2], l[1], l[2], 0.0, -l[0], -l[1], l[0], 0.0};
      const double sg = side ? -1.0 : 1.0;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          hr[18 * side + i + 3 * j] = sg * f[3 * i + j];
          const double rot = f[3 * i] * S[j] + f[3 * i + 1] * S[3 + j] + f[3 * i + 2] * S[6 + j];
          hr[18 * side + i + 3 * (3 + j)] = sg * rot;
        }
    }
    for (int side = 0; side < 2; side++) {
      if (!has[side]) continue;
      for (int e = 0; e < 18; e++)
        if (std::fabs(hr[18 * side + e] - hc[18 * side + e]) > 1e-13 * big) return false;
    }
    const int slot = orig2slot[(size_t)c], sl = slot / 64, lane = slot % 64;
    for (int e = 0; e < 15; e++) (*Fs)[((size_t)sl * 15 + e) * 64 + (size_t)lane] = f[e];
  }
  return true;
}

static DevParams make_params(const fc3d_hip_params* p) {
  DevParams d;
  std::memset(&d, 0, sizeof d);
  d.accel_depth = p->accel_depth;
  d.accel_every = p->accel_every;
  d.accel_reg = p->accel_reg;
  d.lp.solver = p->local_solver;
  d.lp.newton_max_iter = p->local_max_iter;
  d.lp.formulation = p->local_formulation;
  d.lp.ls_max_iter = p->ls_max_iter;
  d.lp.hybrid_strategy = p->hybrid_strategy;
  d.lp.hybrid_max_loop = p->hybrid_max_loop;
  d.lp.hybrid_max_iter = p->hybrid_max_iter;
  d.lp32 = d.lp;
  if (p->fp32_local_solver) d.lp32.solver = p->fp32_local_solver;
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
  RowInfo info((size_t)nc);
  std::string err;
  if (int e = validate_rows(0, nc, row_ptr, col_idx, &info, &err)) {
    set_error("fc3d_hip_problem_create: %s", err.c_str());
    return e;
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

  HostLayout L;
  int ncol = 0;
  append_problem(L, 0, nc, row_ptr, col_idx, blocks, q, mu, h->colour.data(), info,
                 default_sort_window(), &ncol, &h->block_dst);
  if (L.nplanes * 64 > 0x7fffffffLL) {
    set_error("fc3d_hip_problem_create: sliced layout exceeds int32 indexing");
    return fail(NUMERICS_ERR_INVALID_ARGUMENT);
  }
  h->nslices = L.nslices;
  h->nslots = L.nslots;
  h->nplanes = L.nplanes;
  h->colour_slice_ptr = L.colour_slice_ptr;
  if ((rc = dev_upload(&h->d_cps, h->colour_slice_ptr))) return fail(rc);
  h->slot2orig = L.slot2orig;
  h->orig2slot = L.orig2slot;
  h->colour_bytes64 = L.colour_bytes64;
  h->colour_bytes32 = L.colour_bytes32;
  h->colour_nc.assign(h->colour_slice_ptr.size() - 1, 0.0);
  for (int c : h->colour) h->colour_nc[(size_t)c] += 1.0;
  build_dataflow(L, h->colour.data(), &h->dfh);
  if ((rc = dev_upload(&h->len, L.len)) || (rc = dev_upload(&h->dpos, L.dpos)) ||
      (rc = dev_upload(&h->koff, L.koff)) || (rc = dev_upload(&h->kmax, L.kmax)) ||
      (rc = dev_upload(&h->col, L.col)) || (rc = dev_upload(&h->vals, L.vals)) ||
      (rc = dev_upload(&h->diag, L.diag)) || (rc = dev_upload(&h->q, L.q)) ||
      (rc = dev_upload(&h->mu, L.mu)) || (rc = dev_alloc(&h->r, (size_t)h->nslots)) ||
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
                  h->prm,  h->st,    h->feat,   h->ypred, h->featX, h->rbak, h->plibak};
  for (void* p : ptrs)
    if (p) hipFree(p);
  free_accel(h);
  free_stage_bufs(h);
  void* dfp[] = {h->dep_ptr, h->dep_idx, h->qlist, h->done,    h->df,    h->d_block_dst,
                 h->d_o2s,    h->d_cps,    h->d_ba,   h->d_bb,   h->d_inc_ptr, h->d_inc, h->d_H,
                 h->d_minv,   h->d_vfree,  h->d_V,    h->d_H32,     h->d_minv32, h->d_V32,
                 h->d_F,      h->d_F32,    h->d_frz,  h->d_le2,  h->d_fcnt,
                 h->d_fs,     h->d_bmark,  h->d_slice_colour, h->d_focus_cnt,
                 h->d_focus_list, h->d_fdiff32, h->d_fdiff64};
  for (void* p : dfp)
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

int fc3d_hip_problem_update(fc3d_hip_problem* h, const double* blocks, const double* q,
                            const double* mu) {
  if (!h) return NUMERICS_ERR_NULL_POINTER;
  int rc = use_device(h->device);
  if (rc) return rc;
  if (blocks || q) {
    // attached factors of W (velocity form) no longer describe the problem
    FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
    drop_bodies(h);
  }
  const size_t ns = (size_t)h->nslots;
  if (blocks && h->nnzb > 0) {
    if (!h->d_block_dst && (rc = dev_upload(&h->d_block_dst, h->block_dst))) return rc;
    double* d = nullptr;
    if ((rc = dev_alloc(&d, (size_t)h->nnzb * 9))) return rc;
    hipError_t e = hipMemcpyAsync(d, blocks, sizeof(double) * 9 * (size_t)h->nnzb,
                                  hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
      const int64_t n9 = 9 * h->nnzb;
      hipLaunchKernelGGL(k_scatter_blocks, dim3((unsigned)((n9 + 255) / 256)), dim3(256), 0,
                         h->stream, n9, (const double*)d, (const int64_t*)h->d_block_dst, h->vals,
                         h->diag);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d);
    if (e != hipSuccess) {
      set_error("fc3d_hip_problem_update: %s", hipGetErrorString(e));
      return NUMERICS_ERR_DEVICE;
    }
  }
  if ((q || mu) && h->nc > 0) {
    if (!h->d_o2s && (rc = dev_upload(&h->d_o2s, h->orig2slot))) return rc;
    double *dq = nullptr, *dmu = nullptr;
    if ((q && (rc = dev_alloc(&dq, 3 * (size_t)h->nc))) ||
        (mu && (rc = dev_alloc(&dmu, (size_t)h->nc)))) {
      hipFree(dq);
      return rc;
    }
    hipError_t e = hipSuccess;
    if (q) e = hipMemcpyAsync(dq, q, sizeof(double) * 3 * (size_t)h->nc, hipMemcpyHostToDevice,
                              h->stream);
    if (e == hipSuccess && mu)
      e = hipMemcpyAsync(dmu, mu, sizeof(double) * (size_t)h->nc, hipMemcpyHostToDevice,
                         h->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_scatter_contacts, dim3((unsigned)((h->nc + 255) / 256)), dim3(256), 0,
                         h->stream, h->nc, (const double*)dq, (const double*)dmu,
                         (const int*)h->d_o2s, h->q, h->mu);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(dq);
    hipFree(dmu);
    if (e != hipSuccess) {
      set_error("fc3d_hip_problem_update: %s", hipGetErrorString(e));
      return NUMERICS_ERR_DEVICE;
    }
    if (q) {
      // |q| in the caller's order, as dnrm2 does (fc3d_nsgs.c:571)
      double s2 = 0.0;
      for (int64_t i = 0; i < 3 * (int64_t)h->nc; i++) s2 += q[i] * q[i];
      h->norm_q = std::sqrt(s2);
    }
  }
  // FP32 copies follow the FP64 values
  if (h->vals32) {
    auto conv = [&](const double* a, float* b, size_t n) {
      if (n) hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((n + 255) / 256)),
                                dim3(256), 0, h->stream, a, b, n);
    };
    if (blocks) {
      conv(h->vals, h->vals32, (size_t)h->nplanes * 9 * 64);
      conv(h->diag, h->diag32, (size_t)h->nslices * 9 * 64);
    }
    if (q) conv(h->q, h->q32, (size_t)h->nslices * 3 * 64);
    if (mu) conv(h->mu, h->mu32, ns);
    FC3D_HIP_TRY(hipGetLastError());
    FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  }
  return 0;
}

int fc3d_hip_problem_set_bodies(fc3d_hip_problem* h, int nbodies, const int* body_a,
                                const int* body_b, const double* H, const double* Minv,
                                int minv_diagonal, const double* vfree) {
  if (!h || !body_a || !body_b || !H || !Minv || !vfree) return NUMERICS_ERR_NULL_POINTER;
  if (nbodies <= 0) {
    set_error("fc3d_hip_problem_set_bodies: nbodies must be positive");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  const int nc = h->nc;
  for (int c = 0; c < nc; c++)
    if (body_a[c] < -1 || body_a[c] >= nbodies || body_b[c] < -1 || body_b[c] >= nbodies) {
      set_error("fc3d_hip_problem_set_bodies: contact %d has a body outside [-1, %d)", c, nbodies);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
  int rc = use_device(h->device);
  if (rc) return rc;
  // incidence (slot, side) per body, ascending slot
  std::vector<int> ptr((size_t)nbodies + 1, 0), inc;
  for (int c = 0; c < nc; c++) {
    if (body_a[c] >= 0) ptr[(size_t)body_a[c] + 1]++;
    if (body_b[c] >= 0) ptr[(size_t)body_b[c] + 1]++;
  }
  for (int k = 0; k < nbodies; k++) ptr[(size_t)k + 1] += ptr[(size_t)k];
  inc.resize((size_t)ptr[(size_t)nbodies]);
  {
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int slot = 0; slot < h->nslots; slot++) {
      const int c = h->slot2orig[(size_t)slot];
      if (c < 0) continue;
      if (body_a[c] >= 0) inc[(size_t)fill[(size_t)body_a[c]]++] = 2 * slot;
      if (body_b[c] >= 0) inc[(size_t)fill[(size_t)body_b[c]]++] = 2 * slot + 1;
    }
  }
  // the velocity update of a colour is race-free only if a body's contacts
  // all have different colours (the pattern of W must hold every such pair)
  {
    std::vector<int> mark((size_t)h->ncolours, -1);
    for (int k = 0; k < nbodies; k++)
      for (int t = ptr[(size_t)k]; t < ptr[(size_t)k + 1]; t++) {
        const int c = h->slot2orig[(size_t)(inc[(size_t)t] >> 1)];
        const int col = h->colour[(size_t)c];
        if (mark[(size_t)col] == k && !(t > ptr[(size_t)k] && (inc[(size_t)t] >> 1) ==
                                                                   (inc[(size_t)t - 1] >> 1))) {
          set_error("fc3d_hip_problem_set_bodies: two contacts of body %d share colour %d (W "
                    "lacks the block between them)", k, col);
          return NUMERICS_ERR_INVALID_ARGUMENT;
        }
        mark[(size_t)col] = k;
      }
  }
  // the factors must reproduce the resident diagonal blocks of W
  {
    std::vector<double> diag((size_t)h->nslices * 9 * 64);
    FC3D_HIP_TRY(hipMemcpy(diag.data(), h->diag, sizeof(double) * diag.size(),
                           hipMemcpyDeviceToHost));
    for (int c = 0; c < nc; c++) {
      const int slot = h->orig2slot[(size_t)c], sl = slot / 64, lane = slot % 64;
      const int kb[2] = {body_a[c], body_b[c]};
      double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, big = 0.0, err = 0.0;
      for (int s1 = 0; s1 < 2; s1++)
        for (int s2 = 0; s2 < 2; s2++) {
          if (kb[s1] < 0 || kb[s1] != kb[s2]) continue;
          const double* h1 = H + 36 * (size_t)c + 18 * s1;
          const double* h2 = H + 36 * (size_t)c + 18 * s2;
          const int k = kb[s1];
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
              double v = 0.0;
              for (int l = 0; l < 6; l++)
                for (int m = 0; m < 6; m++) {
                  const double mi = minv_diagonal ? (l == m ? Minv[6 * (size_t)k + l] : 0.0)
                                                  : Minv[36 * (size_t)k + l + 6 * m];
                  v += h1[i + 3 * l] * mi * h2[j + 3 * m];
                }
              w[i + 3 * j] += v;
            }
        }
      for (int e = 0; e < 9; e++) {
        const double d = diag[((size_t)sl * 9 + e) * 64 + (size_t)lane];
        big = std::max(big, std::fabs(d));
        err = std::max(err, std::fabs(d - w[e]));
      }
      if (err > 1e-10 * big + 1e-300) {
        set_error("fc3d_hip_problem_set_bodies: H M^-1 H^T differs from W on the diagonal block "
                  "of contact %d (%g)", c, err);
        return NUMERICS_ERR_INVALID_ARGUMENT;
      }
    }
  }
  // ... and the resident q: the velocity form computes a contact's local
  // velocity as H v with v = v_free + M^-1 H^T r, which is q + W r only when
  // q = H v_free (a q with other terms would be solved as if they were not
  // there)
  {
    std::vector<double> qs((size_t)h->nslices * 3 * 64);
    FC3D_HIP_TRY(hipMemcpy(qs.data(), h->q, sizeof(double) * qs.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < nc; c++) {
      const int slot = h->orig2slot[(size_t)c], sl = slot / 64, lane = slot % 64;
      const int kb[2] = {body_a[c], body_b[c]};
      for (int i = 0; i < 3; i++) {
        double hv = 0.0, mag = 0.0;
        for (int side = 0; side < 2; side++) {
          if (kb[side] < 0) continue;
          for (int l = 0; l < 6; l++) {
            const double t = H[36 * (size_t)c + 18 * side + i + 3 * l] * vfree[6 * (size_t)kb[side] + l];
            hv += t;
            mag += std::fabs(t);
          }
        }
        const double qi = qs[((size_t)sl * 3 + i) * 64 + (size_t)lane];
        const double err = std::fabs(qi - hv);
        if (err > 1e-10 * std::max(mag, std::fabs(qi)) + 1e-300) {
          set_error("fc3d_hip_problem_set_bodies: H v_free differs from q at contact %d, "
                    "component %d (%g)", c, i, err);
          return NUMERICS_ERR_INVALID_ARGUMENT;
        }
      }
    }
  }
  // per-slot bodies and H planes
  std::vector<int> ba((size_t)h->nslots, -1), bb((size_t)h->nslots, -1);
  std::vector<double> Hs((size_t)h->nslices * 36 * 64, 0.0);
  for (int c = 0; c < nc; c++) {
    const int slot = h->orig2slot[(size_t)c], sl = slot / 64, lane = slot % 64;
    ba[(size_t)slot] = body_a[c];
    bb[(size_t)slot] = body_b[c];
    for (int e = 0; e < 36; e++) Hs[((size_t)sl * 36 + e) * 64 + (size_t)lane] = H[36 * (size_t)c + e];
  }
  // frames layout when every contact has rigid-body blocks: T and the lever
  // arms (15 values a contact) instead of H (36), beside the resident W_cc
  // (9); the sweep rebuilds H in registers.  FC3D_HIP_FRAMES=0 in the
  // environment keeps the general layout.
  std::vector<double> Fs;
  const char* fenv = std::getenv("FC3D_HIP_FRAMES");
  const bool frames = !(fenv && fenv[0] == '0') &&
                      rigid_frames(nc, body_a, body_b, H, h->orig2slot, h->nslices, &Fs);
  // drop earlier factors, upload the new ones
  drop_bodies(h);
  std::vector<double> mv(Minv, Minv + (size_t)nbodies * (minv_diagonal ? 6 : 36));
  std::vector<double> vf(vfree, vfree + 6 * (size_t)nbodies);
  if ((rc = dev_upload(&h->d_ba, ba)) || (rc = dev_upload(&h->d_bb, bb)) ||
      (rc = dev_upload(&h->d_inc_ptr, ptr)) || (rc = dev_upload(&h->d_inc, inc)) ||
      (rc = frames ? dev_upload(&h->d_F, Fs) : dev_upload(&h->d_H, Hs)) ||
      (rc = dev_upload(&h->d_minv, mv)) || (rc = dev_upload(&h->d_vfree, vf)) ||
      (rc = dev_alloc(&h->d_V, 2 * (size_t)nbodies)))
    return rc;
  h->frames = frames ? 1 : 0;
  h->nbodies = nbodies;
  h->minv_diag = minv_diagonal ? 1 : 0;
  // FC3D_HIP_MINV_ISO=0 in the environment keeps 6 values a body (the same
  // numbers; for comparison)
  const char* ienv = std::getenv("FC3D_HIP_MINV_ISO");
  bool iso = minv_diagonal != 0 && !(ienv && ienv[0] == '0');
  for (int k = 0; iso && k < nbodies; k++) {
    const double* m = Minv + 6 * (size_t)k;
    iso = m[1] == m[0] && m[2] == m[0] && m[4] == m[3] && m[5] == m[3];
  }
  h->minv_iso = iso ? 1 : 0;
  // algorithmic bytes per contact of one velocity-form colour pass, each
  // unique value counted once (the padding of the 4-wide r and v vectors is
  // not), in values of the sweep precision: H (36), or in the frames layout
  // T and the lever arms (15); W_cc (9), mu (1), r read and written (3 + 3),
  // PLI step read and written (2); per body side present: its velocity read
  // and written (6 + 6) and M^-1 (6 or 36; 2 in the FP32 copy of isotropic
  // bodies); plus len and the two body ids (12 bytes).  A two-body FP32
  // contact without the PLI state moves 364 bytes, 280 in the frames layout,
  // 248 with isotropic bodies as well.
  const double mb = minv_diagonal ? 6.0 : 36.0;
  const double mb32 = h->minv_iso ? 2.0 : mb;  // the FP32 copy of an isotropic M^-1
  h->vel_bytes64.assign((size_t)h->ncolours, 0.0);
  h->vel_bytes32.assign((size_t)h->ncolours, 0.0);
  for (int c = 0; c < nc; c++) {
    const int col = h->colour[(size_t)c];
    const double sides = (body_a[c] >= 0 ? 1.0 : 0.0) + (body_b[c] >= 0 ? 1.0 : 0.0);
    const double vals = (frames ? 15 : 36) + 9 + 1 + 6 + 2 + sides * 12;
    h->vel_bytes64[(size_t)col] += 8.0 * (vals + sides * mb) + 12.0;
    h->vel_bytes32[(size_t)col] += 4.0 * (vals + sides * mb32) + 12.0;
  }
  return 0;
}

int fc3d_hip_problem_body_layout(const fc3d_hip_problem* h) {
  if (!h) return NUMERICS_ERR_NULL_POINTER;
  if (!h->nbodies) return 0;
  return h->frames ? 2 : 1;
}

int fc3d_hip_problem_features(fc3d_hip_problem* h, float* feat, float* target) {
  if (!h || !feat) return NUMERICS_ERR_NULL_POINTER;
  int rc = use_device(h->device);
  if (rc) return rc;
  const size_t ns = (size_t)h->nslots;
  float *df = nullptr, *dt = nullptr;
  if ((rc = dev_alloc(&df, ns * MLP_NF)) || (rc = dev_alloc(&dt, ns * 3))) {
    hipFree(df);
    return rc;
  }
  hipLaunchKernelGGL(k_features, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, h->stream,
                     ptrs64(h, h->part), h->nslots, df, target ? dt : (float*)nullptr);
  std::vector<float> hf(ns * MLP_NF), ht(target ? ns * 3 : 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(hf.data(), df, sizeof(float) * hf.size(), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && target)
    e = hipMemcpyAsync(ht.data(), dt, sizeof(float) * ht.size(), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(df);
  hipFree(dt);
  if (e != hipSuccess) {
    set_error("fc3d_hip_problem_features: %s", hipGetErrorString(e));
    return NUMERICS_ERR_DEVICE;
  }
  for (int i = 0; i < h->nc; i++) {
    const size_t slot = (size_t)h->orig2slot[(size_t)i];
    std::memcpy(feat + (size_t)i * MLP_NF, hf.data() + slot * MLP_NF, sizeof(float) * MLP_NF);
    if (target) std::memcpy(target + (size_t)i * 3, ht.data() + slot * 3, sizeof(float) * 3);
  }
  return 0;
}

int fc3d_hip_problem_set_model(fc3d_hip_problem* h, fc3d_hip_mlp* m) {
  if (!h) return NUMERICS_ERR_NULL_POINTER;
  if (m && mlp_device(m) != h->device) {
    set_error("fc3d_hip_problem_set_model: model on device %d, problem on device %d",
              mlp_device(m), h->device);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  h->model = m;
  return 0;
}

int fc3d_hip_solve(fc3d_hip_problem* h, const fc3d_hip_params* p, fc3d_hip_stats* s) {
  return fc3d_hip_solve_observed(h, p, s, nullptr, nullptr);
}

int fc3d_hip_solve_observed(fc3d_hip_problem* h, const fc3d_hip_params* p, fc3d_hip_stats* s,
                            fc3d_hip_sweep_observer observer, void* env) {
  if (!h || !p || !s) return NUMERICS_ERR_NULL_POINTER;
  const int ls = p->local_solver;
  if (!(ls == 565 || ls == 551 || ls == 550 || ls == 553 || ls == 552)) {
    set_error("fc3d_hip_solve: local solver %d not supported", ls);
    return NUMERICS_ERR_INVALID_SOLVER;
  }
  const int ls32 = p->fp32_local_solver;
  if (!(ls32 == 0 || ls32 == 565 || ls32 == 551 || ls32 == 550 || ls32 == 553 || ls32 == 552)) {
    set_error("fc3d_hip_solve: FP32 local solver %d not supported", ls32);
    return NUMERICS_ERR_INVALID_SOLVER;
  }
  if (p->error_evaluation < 0 || p->error_evaluation > 3) {
    set_error("fc3d_hip_solve: unknown error evaluation %d", p->error_evaluation);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if ((p->rho_strategy < 0 || p->rho_strategy > 3) &&
      (ls == 565 || ls == 551 || ls == 550 || ls32 == 565 || ls32 == 551 || ls32 == 550)) {
    set_error("fc3d_hip_solve: rho strategy %d not supported", p->rho_strategy);
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  if (p->warm_start < 0 || p->warm_start > 3) {
    set_error("fc3d_hip_solve: warm_start %d is not 0, 1, 2 or 3", p->warm_start);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (p->accel_depth < 0 || p->accel_depth > AA_MAX || p->accel_every < 1 ||
      !(p->accel_reg >= 0.0)) {
    set_error("fc3d_hip_solve: acceleration depth %d (0..%d), every %d (>= 1), reg %g (>= 0)",
              p->accel_depth, AA_MAX, p->accel_every, p->accel_reg);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (p->velocity_form && !h->d_H && !h->d_F) {
    set_error("fc3d_hip_solve: velocity_form needs the factors of W "
              "(fc3d_hip_problem_set_bodies)");
    return NUMERICS_ERR_INVALID_OPTION;
  }
  if (p->focus_passes < 0 || (p->focus_passes > 0 &&
                               (!p->velocity_form || p->accel_depth > 0 ||
                                p->focus_every < 1 || p->focus_count < 1 || p->focus_after < 0 ||
                                !(p->focus_share >= 0.0) || !(p->focus_omega > 0.0) ||
                                !(p->focus_omega < 2.0)))) {
    set_error("fc3d_hip_solve: focused passes need the velocity form, no acceleration, "
              "focus_every >= 1, focus_count >= 1, focus_after >= 0, focus_share >= 0 and "
              "0 < focus_omega < 2");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (p->freeze_contacts < 0 ||
      (p->freeze_contacts > 0 && (p->precision != 0 || p->velocity_form || p->accel_depth > 0 ||
                                  p->warm_start >= 2))) {
    set_error("fc3d_hip_solve: contact freezing (%d) needs FP64 sweeps on the rows of W, without "
              "acceleration or a learned start", p->freeze_contacts);
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  if (p->warm_start >= 2 && !h->model) {
    set_error("fc3d_hip_solve: warm_start %d needs a model (fc3d_hip_problem_set_model)",
              p->warm_start);
    return NUMERICS_ERR_INVALID_OPTION;
  }
  int rc = use_device(h->device);
  if (rc) return rc;
  std::memset(s, 0, sizeof *s);
  s->num_colours = h->ncolours;
  s->warm_start_used = p->warm_start == 1 ? 1 : 0;

  h->freeze = 0;  // the autotuner's trial sweeps do not freeze
  h->freeze_tmp = (double)h->nc * (double)h->nc / 1000.;
  DevParams dp = make_params(p);
  dp.freeze = p->freeze_contacts;
  dp.nc = h->nc;
  dp.freeze_tmp = h->freeze_tmp;
  dp.focus_every = p->focus_every;
  dp.focus_after = p->focus_after;
  dp.focus_count = p->focus_count;
  dp.focus_share = p->focus_share;
  dp.focus_omega = p->focus_omega;
  FC3D_HIP_TRY(hipMemcpyAsync(h->prm, &dp, sizeof dp, hipMemcpyHostToDevice, h->stream));
  if (p->warm_start == 1)
    FC3D_HIP_TRY(hipMemsetAsync(h->r, 0, sizeof(double4) * (size_t)h->nslots, h->stream));
  const size_t ns = (size_t)h->nslots;
  hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, h->stream,
                     h->pli, 1.0, ns);
  const int mixed = p->precision == 1;
  if (mixed && (rc = ensure_fp32(h))) return rc;

  {
    const int eff32 = ls32 ? ls32 : ls;
    h->pli32_off = eff32 == 550 || eff32 == 551 || eff32 == 552;
  }
  h->accel_on = 0;  // the autotuner times plain matrix-form sweeps
  h->vel_on = 0;
  int sched[2] = {p->schedule, p->schedule};
  if (p->schedule < 0 || p->schedule > FC3D_NUM_SCHEDULES) {
    set_error("fc3d_hip_solve: schedule %d out of range [0, %d]", p->schedule, FC3D_NUM_SCHEDULES);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (p->schedule == 0 && p->velocity_form) {
    sched[0] = sched[1] = 1;  // the velocity form has one kernel of its own
  } else if (p->schedule == 0) {
    for (int ph = 0; ph <= mixed; ph++) {
      if (!h->tuned_schedule[ph]) {
        if (ph == 1)
          hipLaunchKernelGGL(k_r64_to_r32, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                             h->stream, h->r, h->r32, h->pli, h->pli32, ns);
        if ((rc = tune_schedule(h, ph, h->norm_q, p->local_tol, &h->tuned_schedule[ph])))
          return rc;
        hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                           h->stream, h->pli, 1.0, ns);
      }
      sched[ph] = h->tuned_schedule[ph];
    }
  }
  s->schedule = sched[mixed];
  if ((sched[0] == FC3D_SCHEDULE_DATAFLOW || (mixed && sched[1] == FC3D_SCHEDULE_DATAFLOW)) &&
      (rc = ensure_dataflow(h)))
    return rc;
  if (p->accel_depth > 0) {
    if ((rc = ensure_accel(h, mixed))) return rc;
    h->accel_on = 1;
  }
  h->focus_passes = 0;
  if (p->velocity_form && p->focus_passes > 0) {
    if ((rc = ensure_focus(h, mixed))) return rc;
    hipLaunchKernelGGL(k_focus_reset, dim3(1), dim3(64), 0, h->stream, h->d_fs);
    h->focus_passes = p->focus_passes;
    // FC3D_HIP_FOCUS_HBM=1: the passes keep the body velocities in HBM (the
    // same bits, slower; for comparison)
    const char* fh = std::getenv("FC3D_HIP_FOCUS_HBM");
    h->focus_lds = !(fh && fh[0] == '1');
  }
  if (p->velocity_form) {
    if (mixed && (rc = ensure_vel32(h))) return rc;
    if (p->accel_depth > 0 && ((rc = alloc_accel_bodies(h, &h->aa64)) ||
                               (mixed && (rc = alloc_accel_bodies(h, &h->aa32)))))
      return rc;
    h->vel_on = 1;
  }

  if (p->freeze_contacts > 0) {
    // fc3d_nsgs.c:302-313, 712: counters and light-error terms start at 0
    if (!h->d_frz && ((rc = dev_alloc(&h->d_frz, (size_t)h->nslots)) ||
                      (rc = dev_alloc(&h->d_le2, (size_t)h->nslots)) ||
                      (rc = dev_alloc(&h->d_fcnt, (size_t)h->nslices))))
      return rc;
    FC3D_HIP_TRY(hipMemsetAsync(h->d_frz, 0, sizeof(int) * (size_t)h->nslots, h->stream));
    FC3D_HIP_TRY(hipMemsetAsync(h->d_le2, 0, sizeof(double) * (size_t)h->nslots, h->stream));
    FC3D_HIP_TRY(hipMemsetAsync(h->d_fcnt, 0, sizeof(int) * (size_t)h->nslices, h->stream));
    h->freeze = p->freeze_contacts;
  }

  SolveState s0;
  std::memset(&s0, 0, sizeof s0);
  s0.norm_r = 1e24;  // fc3d_nsgs.c:575
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
  if (p->warm_start >= 2 &&
      (rc = learned_start(h, mixed, sched[mixed], p->start_trial_sweeps, p->warm_start == 3, &s0,
                          s)))
    return rc;
  const int chunk = p->sweeps_per_launch > 0 ? p->sweeps_per_launch : 1;
  // FC3D_HIP_NO_GRAPH=1 enqueues the same kernels directly instead of
  // replaying a captured graph (for tools that do not follow graph launches)
  const char* ng = getenv("FC3D_HIP_NO_GRAPH");
  const int no_graph = ng && ng[0] == '1';

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
    if (h->accel_on) {
      if (phase == 0 && mixed && (rc = accel_stats(h, s))) return rc;
      if ((rc = reset_accel(h))) return rc;
    }
    if (h->vel_on) launch_vinit(h, phase, nullptr, nullptr);  // v from the phase's reaction
    // The FP64 sweeps that end a mixed solve are few (2-4 at an FP32 local
    // tolerance of 1e-7), so they replay 4 sweeps at a time: the sweeps of a
    // replay that follow the stop still launch, and each costs about a
    // working sweep's loads (its waves load their slots before they test the
    // stop flag; ~140 us on 1M contacts).
    const int ck = phase == 0 && mixed ? std::min(chunk, 4) : chunk;
    // every replay advances the sweep counter or stops; the bound only
    // guards against a launch that silently did not run
    const long max_replays = (long)p->max_iter / (p->instrument ? 1 : ck) + 4;
    for (long replay = 0;; replay++) {
      SolveState cur;
      if (replay > max_replays) {
        set_error("fc3d_hip_solve: sweep loop did not stop after %ld replays", replay);
        return NUMERICS_ERR_UNKNOWN;
      }
      if (p->instrument) {
        enqueue_sweep(h, phase, sched[phase], kev.data(), (int)kev.size());
        if (h->focus_passes > 0) enqueue_focus(h, phase);
        FC3D_HIP_TRY(hipGetLastError());
        FC3D_HIP_TRY(hipMemcpyAsync(&cur, h->st, sizeof cur, hipMemcpyDeviceToHost, h->stream));
        FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
        for (int c = 0; c < timed_launches(h, phase, sched[phase]); c++) {
          float ms = 0.f;
          hipEventElapsedTime(&ms, kev[2 * (size_t)c], kev[2 * (size_t)c + 1]);
          const double by = timed_bytes(h, phase, sched[phase], c);
          s->sweep_kernel_ms += ms;
          s->sweep_kernel_launches += 1;
          s->sweep_kernel_bytes += by;
          if (phase) {
            s->sweep_kernel_ms_fp32 += ms;
            s->sweep_kernel_launches_fp32 += 1;
            s->sweep_kernel_bytes_fp32 += by;
            if (h->vel_on && c >= vel_tail_begin(h, phase)) {
              s->sweep_tail_ms_fp32 += ms;
              s->sweep_tail_launches_fp32 += 1;
              s->sweep_tail_bytes_fp32 += by;
            }
          }
        }
      } else if (no_graph) {
        for (int k = 0; k < ck; k++) enqueue_sweep(h, phase, sched[phase], nullptr, 0);
        if (h->focus_passes > 0) enqueue_focus(h, phase);
        FC3D_HIP_TRY(hipGetLastError());
        FC3D_HIP_TRY(hipMemcpyAsync(&cur, h->st, sizeof cur, hipMemcpyDeviceToHost, h->stream));
        FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
      } else {
        if ((rc = build_graph(h, phase, ck, sched[phase]))) return rc;
        FC3D_HIP_TRY(hipGraphLaunch(h->graph[phase], h->stream));
        FC3D_HIP_TRY(hipMemcpyAsync(&cur, h->st, sizeof cur, hipMemcpyDeviceToHost, h->stream));
        FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
      }
      if (observer) {
        // the iterate is readable in FP64 while the observer runs
        if (phase == 1) {
          hipLaunchKernelGGL(k_r32_to_r64, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                             h->stream, h->r32, h->r, h->pli32, h->pli, ns);
          FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
        }
        observer(env, cur.iter, cur.incr);
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
  if (h->accel_on && (rc = accel_stats(h, s))) return rc;
  if ((rc = dataflow_error(h))) return rc;
  float ms = 0.f;
  hipEventElapsedTime(&ms, t0, t1);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  for (auto& e : kev) hipEventDestroy(e);
  if (h->focus_passes > 0) {
    FocusState fst;
    FC3D_HIP_TRY(hipMemcpy(&fst, h->d_fs, sizeof fst, hipMemcpyDeviceToHost));
    s->focus_rounds = fst.rounds;
    s->focus_listed = fst.listed_sum;
    s->focus_due = fst.due;
    s->focus_last_share = fst.last_share;
    s->focus_lds_rounds = fst.lds_rounds;
  }
  s->solve_ms = ms;
  s->iterations = fin.iter;
  s->incr_error = incr;
  s->residual = full;
  s->info = notconv;
  if (fin.iter == p->max_iter && notconv > 0) s->info = NUMERICS_ERR_MAX_ITER;
  return 0;
}

}  // extern "C"
