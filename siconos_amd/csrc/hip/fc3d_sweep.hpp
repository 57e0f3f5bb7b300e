// fc3d_sweep.hpp — device-side pieces shared by the single-problem solver
// (fc3d_problem.hip) and the batched solver (fc3d_batch.hip): parameter and
// state blocks, the sliced-ELL access, and the per-slot sweep and residual.
#pragma once
#include "fc3d_device.hpp"

namespace fc3d {

template <typename T> struct vec4;
template <> struct vec4<double> { using type = double4; };
template <> struct vec4<float> { using type = float4; };

struct DevParams {
  LocalParams lp;
  LocalParams lp32;     // local solver of the FP32 sweeps
  double tol;           // user tolerance
  double local_tol;     // local tolerance at solve start (FP64 phase)
  double fp32_local_tol;
  double fp32_tol;
  double omega;
  double rho_constant;
  double accel_reg;     // Anderson: Tikhonov weight relative to the Gram trace
  int accel_depth;      // Anderson history columns (0: off)
  int accel_every;      // Anderson step on every accel_every-th sweep
  int mode;             // error evaluation (NonSmoothGaussSeidel_options.h:121-140)
  int freq;             // error evaluation frequency
  int max_iter;
  int filter;
  int relaxation;
  int rho_strategy;
  int focus_every;      // focused passes: a round after every focus_every-th sweep
  int focus_after;      // ... from this sweep on
  int focus_count;      // contacts of largest |r_new - r_old|^2 that seed a round
  double focus_share;   // least share of sum |r_new - r_old|^2 they must carry
  double focus_omega;   // relaxation of the focus passes (1: none)
  int freeze;           // contact freezing: sweeps a frozen contact is skipped (0: off)
  int nc;               // contacts of the problem (freezing's reset rule)
  double freeze_tmp;    // nc * nc / 1000 (fc3d_nsgs.c:722)
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
  double nf;      // sqrt of the last sweep's sum |r_new - r_old|^2
  double norm_r;  // |r| after the last light-error evaluation (1e24 before)
  int unfreeze;   // contact freezing: every counter reads 0 in this sweep
  int focus_last; // focused passes: sweep count at the last round
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
  T* pli;           // per-slot PLI step (persistent local state); null when the
                    // local solver keeps none (550, 551, 552): not loaded, not stored
  double* part;     // per-slice partial sums
  typename vec4<T>::type* f = nullptr;  // if set: r_new - r_old of the sweep
  // contact freezing (FP64 sweeps; null when off): per-slot counters of the
  // sweeps still to skip, the slot's last light-error term, per-slice counts
  // of frozen slots after the sweep, and the solve state the criteria read
  int* frz = nullptr;
  double* le2 = nullptr;
  int* fcnt = nullptr;
  const SolveState* fst = nullptr;
  int freeze = 0;
  double freeze_tmp = 1.0;
};

// Adds the value of the lane that DPP control CTRL selects (all lanes active).
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)u, CTRL, 0xF, 0xF, false);
  const unsigned hi = __builtin_amdgcn_update_dpp(0u, (unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return v + __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Adds the value of the lane ROWS lanes away (lane ^ ROWS, ROWS = 16 or 32)
// with gfx950's v_permlane{16,32}_swap.  Swapping v with itself leaves each
// lane holding its own value in one result and its partner's in the other
// (which one depends on the lane), and IEEE addition is commutative, so the
// sum of the two is v + v[lane ^ ROWS] bit for bit.
template <int ROWS>
__device__ __forceinline__ double swap_add(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned ul = (unsigned)u, uh = (unsigned)(u >> 32);
  const auto lo = ROWS == 16 ? __builtin_amdgcn_permlane16_swap(ul, ul, false, false)
                             : __builtin_amdgcn_permlane32_swap(ul, ul, false, false);
  const auto hi = ROWS == 16 ? __builtin_amdgcn_permlane16_swap(uh, uh, false, false)
                             : __builtin_amdgcn_permlane32_swap(uh, uh, false, false);
  const double a = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
  const double b = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
  return a + b;
}

// Sum over the wave's 64 lanes as the xor butterfly (partners lane ^ 1, ^ 2,
// ^ 4, ^ 8, ^ 16, ^ 32) computes it, every lane ending with the same bits,
// with no LDS round trip: quad permutations for ^ 1 and ^ 2; after those
// steps each quad holds one value, so the half-row mirror (lane 7 - i of each
// 8) reads what lane ^ 4 holds, and likewise the row mirror for ^ 8; the
// permlane swaps for ^ 16 and ^ 32.  Call with every lane active.
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141>(v);  // row_half_mirror
  v = dpp_add<0x140>(v);  // row_mirror
  v = swap_add<16>(v);
  v = swap_add<32>(v);
  return v;
}

// Fixed-order sums of NV interleaved partials (part[i * NV + j], j < NV) as
// 1024 threads add them: thread t adds entries t, t + 1024, ... in that order
// (sixteen loads in flight at a time, so a sweep's ~16k slice partials cost
// one memory round trip), each wave of 64 threads adds its values with
// wave_sum, then the 16 wave sums are added in order.  A workgroup of R
// threads (R divides 1024, R >= 64) computes the same bits: thread r stands
// for threads r, r + R, ..., and its wave for the matching waves.  sh holds
// NV * 16 doubles; the results are valid on thread 0.
template <int NV, int R = 1024>
__device__ void block_sums(const double* __restrict__ part, int n, double* sh, double* out) {
  static_assert(R >= 64 && R <= 1024 && 1024 % R == 0, "R must divide 1024");
  constexpr int M = 1024 / R;          // threads each real thread stands for
  constexpr int MB = M < 2 ? M : 2;    // of those, loaded together
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();  // sh may hold an earlier call's sums
#pragma unroll
  for (int m0 = 0; m0 < M; m0 += MB) {
    double v[MB][NV];
#pragma unroll
    for (int mm = 0; mm < MB; mm++)
#pragma unroll
      for (int j = 0; j < NV; j++) v[mm][j] = 0.0;
    for (int base = 0; base < n; base += 16 * 1024) {
      double t[MB][16][NV];
#pragma unroll
      for (int mm = 0; mm < MB; mm++)
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const int k = base + (int)threadIdx.x + (m0 + mm) * R + u * 1024;
#pragma unroll
          for (int j = 0; j < NV; j++) t[mm][u][j] = k < n ? part[(size_t)k * NV + j] : 0.0;
        }
#pragma unroll
      for (int mm = 0; mm < MB; mm++)
#pragma unroll
        for (int u = 0; u < 16; u++)
#pragma unroll
          for (int j = 0; j < NV; j++)
            if (base + (int)threadIdx.x + (m0 + mm) * R + u * 1024 < n) v[mm][j] += t[mm][u][j];
    }
#pragma unroll
    for (int mm = 0; mm < MB; mm++)
#pragma unroll
      for (int j = 0; j < NV; j++) {
        const double w = wave_sum(v[mm][j]);
        if (lane == 0) sh[j * 16 + wave + (m0 + mm) * (R / 64)] = w;
      }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; j++) {
      double t = 0.0;
      for (int k = 0; k < 16; k++) t += sh[j * 16 + k];
      out[j] = t;
    }
  }
}

// State update at the end of a sweep from its light-error sums (one thread):
// fc3d_nsgs.c:90-96 and the convergence test of the chosen evaluation mode
// (fc3d_nsgs.c:826-863).  The state and the parameters it reads are loaded
// at once into registers and the state stored back whole, so the update
// waits on one memory round trip instead of one per field.
__device__ inline void end_sweep_state(const DevParams* prm, SolveState* st, double light,
                                       double nr2, int want_phase) {
  SolveState x = *st;
  const double fp32_tol = prm->fp32_tol;
  const int max_iter = prm->max_iter, mode = prm->mode;
  x.iter += 1;
  x.nf = sqrt(light);
  double incr = sqrt(light);
  const double nr = sqrt(nr2);
  if (fabs(nr) > DBL_EPSILON) incr /= nr;
  x.incr = incr;
  if (want_phase == 0 && (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT ||
                          mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL))
    x.norm_r = nr;  // calculateLightError keeps |r| (fc3d_nsgs.c:90-96)
  if (want_phase == 1) {
    // FP32 phase: hand over to FP64 sweeps once the FP32 light error is small
    x.fp32_iter += 1;
    x.need_full = 0;
    if (incr < fp32_tol || x.iter >= max_iter) x.stop = 1;
  } else {
    if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT) {
      x.notconv = incr < x.working_tol ? 0 : 1;
      x.need_full = 0;
    } else if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
      if (incr >= x.working_tol) {
        x.notconv = 1;
        x.need_full = 0;
      } else {
        x.need_full = 1;
      }
    } else {
      const int f = x.freq;
      x.need_full = f > 0 ? (x.iter % f == 0) : 1;
      if (!x.need_full) {
        x.full = 1e24;
        x.incr = 1e24;
        x.notconv = 1;
      }
    }
    if (!x.need_full && (x.notconv == 0 || x.iter >= max_iter)) x.stop = 1;
  }
  *st = x;
}

// y += sum_k W_k r_{col(k)} over the off-diagonal blocks of the lane's row, in
// the reference row order (SBM_row_prod_no_diag_3x3, SparseBlockMatrix.c:
// 1431-1490).  The slice index is wave-uniform, so its metadata are scalar
// loads; planes are read U at a time (padded entries hold a zero block and a
// valid column) so that U blocks and U neighbour reactions are in flight
// together; only blocks k < len enter the sum.
template <typename T, int U, int PF>
__device__ __forceinline__ void gather_row(const SweepPtrs<T>& P, int s, int lane, int len, T* y,
                                           int koff, int kmax) {
  const int* __restrict__ colp = P.col + (size_t)koff * 64 + lane;
  const T* __restrict__ valp = P.vals + (size_t)koff * 9 * 64 + lane;
  int k0 = 0;
  if (PF > 0) {
    // the first PF column indices are loaded at once, so that only one
    // memory latency separates them from the neighbour reactions
    int jj[PF > 0 ? PF : 1];
#pragma unroll
    for (int u = 0; u < PF; u++)
      if (u < kmax) jj[u] = colp[(size_t)u * 64];
#pragma unroll
    for (int b = 0; b < PF; b += U) {
      if (b >= kmax) break;
      T w[U][9];
      typename vec4<T>::type rj[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (b + u < kmax) {
          rj[u] = P.r[jj[b + u]];
#pragma unroll
          for (int e = 0; e < 9; e++) w[u][e] = valp[((size_t)(b + u) * 9 + e) * 64];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (b + u < len) {
          const T v[3] = {rj[u].x, rj[u].y, rj[u].z};
          mvp3x3<T>(w[u], v, y);
        }
      }
    }
    k0 = PF;
  }
  for (; k0 < kmax; k0 += U) {
    int j[U];
    T w[U][9];
    typename vec4<T>::type rj[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u < kmax) {
        j[u] = colp[(size_t)(k0 + u) * 64];
#pragma unroll
        for (int e = 0; e < 9; e++) w[u][e] = valp[((size_t)(k0 + u) * 9 + e) * 64];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u < kmax) rj[u] = P.r[j[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u < len) {
        const T v[3] = {rj[u].x, rj[u].y, rj[u].z};
        mvp3x3<T>(w[u], v, y);
      }
    }
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD
// b % 8).  This maps the blocks of each XCD onto one contiguous range of
// slices, so that the neighbour gathers of an XCD stay within its own L2.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, nb = gridDim.x;
  const int x = b & 7, q = nb >> 3, rem = nb & 7;
  return x * q + (x < rem ? x : rem) + (b >> 3);
}

// Inputs of one slot's local solve: diagonal block, mu, the old reaction and
// the persistent PLI step; q_loc starts at the slot's q.
template <typename T>
struct SlotInputs {
  LocalProblem<T> lp;
  typename vec4<T>::type ro;
  T pli;
  T ql[3];
};

// A load that marks its line non-temporal when NT: streamed planes then leave
// the body data that later colours reread in L2.
template <bool NT, typename T>
__device__ __forceinline__ T ld(const T* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <typename T, bool NT = false>
__device__ __forceinline__ void load_slot(const SweepPtrs<T>& P, int s, int lane, SlotInputs<T>* in) {
  const int slot = s * 64 + lane;
#pragma unroll
  for (int e = 0; e < 9; e++) in->lp.W[e] = ld<NT>(&P.diag[(size_t)(s * 9 + e) * 64 + lane]);
  in->lp.mu = P.mu[slot];
  in->ro = P.r[slot];
  in->pli = P.pli ? P.pli[slot] : T(0);
#pragma unroll
  for (int e = 0; e < 3; e++) in->ql[e] = P.q[(size_t)(s * 3 + e) * 64 + lane];
}

// The constants of the per-contact step, read from DevParams once per wave
// (all wave-uniform, so they sit in scalar registers).
struct StepConsts {
  LocalParams lp;  // the local solver of this precision's sweeps
  int rho_strategy;
  int relaxation;
  int filter;
  double rho_constant;
  double omega;
};

template <typename T>
__device__ __forceinline__ StepConsts step_consts(const DevParams* __restrict__ prm) {
  StepConsts c;
  c.lp = sizeof(T) == 8 ? prm->lp : prm->lp32;
  c.rho_strategy = prm->rho_strategy;
  c.relaxation = prm->relaxation;
  c.filter = prm->filter;
  c.rho_constant = prm->rho_constant;
  c.omega = prm->omega;
  return c;
}

// The step constants of the focused passes: the sweep's, with the passes' own
// relaxation when focus_omega is not 1.
template <typename T>
__device__ __forceinline__ StepConsts focus_consts(const DevParams* __restrict__ prm) {
  StepConsts c = step_consts<T>(prm);
  if (prm->focus_omega != 1.0) {
    c.relaxation = 1;
    c.omega = prm->focus_omega;
  }
  return c;
}

// Makes every field of c live here, so that its scalar loads are issued
// together before this point instead of one by one where each is first used.
__device__ __forceinline__ void step_consts_ready(const StepConsts& c) {
  asm volatile("" ::"s"(c.lp.solver), "s"(c.lp.newton_max_iter), "s"(c.lp.formulation),
               "s"(c.lp.ls_max_iter), "s"(c.lp.hybrid_strategy), "s"(c.lp.hybrid_max_loop),
               "s"(c.lp.hybrid_max_iter), "s"(c.rho_strategy), "s"(c.relaxation), "s"(c.filter),
               "s"(c.rho_constant), "s"(c.omega));
}

// Local solve, relaxation, light-error term, optional filter and write-back
// for a slot whose local q (q + sum_j W_ij r_j) is in in->ql.
template <typename T>
__device__ __forceinline__ void finish_slot(const SweepPtrs<T>& P, const StepConsts& c, int s,
                                            int lane, T tol, SlotInputs<T>& in, double* light,
                                            double* nr2, typename vec4<T>::type* out_r = nullptr) {
  const int slot = s * 64 + lane;
  LocalProblem<T>& lp = in.lp;
  const typename vec4<T>::type ro = in.ro;
  if (P.frz) {
    // fc3d_nsgs.c:749-756: a frozen contact keeps its reaction and adds its
    // last light-error term again
    const int f0 = P.fst->unfreeze ? 0 : P.frz[slot];
    if (f0 > 0) {
      P.frz[slot] = f0 - 1;
      *light = P.le2[slot];
      *nr2 = double(ro.x * ro.x + ro.y * ro.y + ro.z * ro.z);
      if (out_r) *out_r = ro;
      return;
    }
  }
  T pli = in.pli;
  lp.q[0] = in.ql[0]; lp.q[1] = in.ql[1]; lp.q[2] = in.ql[2];
  const T rold[3] = {ro.x, ro.y, ro.z};
  T lr[3] = {ro.x, ro.y, ro.z};
  T rho[3];
  compute_rho<T>(c.rho_strategy, T(c.rho_constant), lp.W, rho);
  T res = T(0);
  local_solve<T>(lp, lr, c.lp, tol, &pli, rho, &res);
  if (c.relaxation) {
    const T om = T(c.omega);
#pragma unroll
    for (int e = 0; e < 3; e++) lr[e] = om * lr[e] + (T(1) - om) * rold[e];
  }
  const T d0 = rold[0] - lr[0], d1 = rold[1] - lr[1], d2 = rold[2] - lr[2];
  *light = double(d0 * d0 + d1 * d1 + d2 * d2);
  if (P.frz) {
    // fc3d_nsgs.c:723-724, 776-800: freeze the contact for P.freeze sweeps
    // when its step is small against its reaction, or its reaction small
    // against |r|, from the 10th sweep on
    const SolveState* fs = P.fst;
    const double wt = fs->working_tol, nr = fs->norm_r;
    const double crit1 = wt * wt / P.freeze_tmp, crit2 = nr * nr / P.freeze_tmp;
    const double l2 = *light;
    const double sq = double(lr[0] * lr[0] + lr[1] * lr[1] + lr[2] * lr[2]);
    P.le2[slot] = l2;
    const bool relative = l2 <= crit1 * sq, small = sq <= crit2;
    P.frz[slot] = ((relative || small) && fs->iter + 1 >= 10) ? P.freeze : 0;
  }
  const bool accept = !(c.filter && (isnan(res) || isinf(res) || res > T(1)));
  typename vec4<T>::type out = ro;
  if (accept) { out.x = lr[0]; out.y = lr[1]; out.z = lr[2]; }
  P.r[slot] = out;
  if (P.pli) P.pli[slot] = pli;
  if (out_r) *out_r = out;
  if (P.f) {
    typename vec4<T>::type d = out;
    d.x -= ro.x; d.y -= ro.y; d.z -= ro.z; d.w = T(0);
    P.f[slot] = d;
  }
  *nr2 = sizeof(T) == 8 ? double(out.x * out.x + out.y * out.y + out.z * out.z)
                        : double(out.x) * double(out.x) + double(out.y) * double(out.y) +
                              double(out.z) * double(out.z);
}

template <typename T>
__device__ __forceinline__ void finish_slot(const SweepPtrs<T>& P, const DevParams* __restrict__ prm,
                                            int s, int lane, T tol, SlotInputs<T>& in,
                                            double* light, double* nr2,
                                            typename vec4<T>::type* out_r = nullptr) {
  finish_slot<T>(P, step_consts<T>(prm), s, lane, tol, in, light, nr2, out_r);
}

// The per-contact step of a sweep (fc3d_nsgs.c:737-822) for the lane's slot of
// slice s: local problem (fc3d_nsgs_update), local solve, relaxation, light
// error term, optional filter, write-back.  Returns |r_old - r_new|^2 and the
// squared norm of the accepted reaction (0 for padding slots).
// The slice's plane range (wave-uniform scalar loads).
template <typename T>
__device__ __forceinline__ void slice_planes(const SweepPtrs<T>& P, int s, int* koff, int* kmax) {
  *koff = __builtin_amdgcn_readfirstlane(P.koff[s]);
  *kmax = __builtin_amdgcn_readfirstlane(P.kmax[s]);
}

// sweep_slot with the slot's inputs, the slice's plane range and the step
// constants already loaded by the caller.
template <typename T, int U, int PF>
__device__ __forceinline__ void sweep_slot_loaded(const SweepPtrs<T>& P, const StepConsts& c,
                                                  int s, int lane, T tol, int len,
                                                  SlotInputs<T>& in, int koff, int kmax,
                                                  double* light, double* nr2) {
  *light = 0.0;
  *nr2 = 0.0;
  if (len >= 0) {
    gather_row<T, U, PF>(P, s, lane, len, in.ql, koff, kmax);
    finish_slot<T>(P, c, s, lane, tol, in, light, nr2);
  }
}

template <typename T, int U, int PF>
__device__ __forceinline__ void sweep_slot(const SweepPtrs<T>& P, const DevParams* __restrict__ prm,
                                           int s, int lane, T tol, double* light, double* nr2) {
  *light = 0.0;
  *nr2 = 0.0;
  const int len = P.len[s * 64 + lane];
  if (len >= 0) {
    // local-solve inputs first: their loads overlap the gather
    SlotInputs<T> in;
    load_slot<T>(P, s, lane, &in);
    int koff, kmax;
    slice_planes<T>(P, s, &koff, &kmax);
    gather_row<T, U, PF>(P, s, lane, len, in.ql, koff, kmax);
    finish_slot<T>(P, prm, s, lane, tol, in, light, nr2);
  }
}

// Partial gather for the four-wave sweep: planes k = first, first + step, ...
// of the slice, blocks k < len only, summed into y (order of the planes taken;
// not the reference row order, so used only where bit-exactness is not kept).
template <typename T, int U>
__device__ __forceinline__ void gather_planes(const SweepPtrs<T>& P, int s, int lane, int len,
                                              int first, int step, T* y) {
  const int koff = __builtin_amdgcn_readfirstlane(P.koff[s]);
  const int kmax = __builtin_amdgcn_readfirstlane(P.kmax[s]);
  const int* __restrict__ colp = P.col + (size_t)koff * 64 + lane;
  const T* __restrict__ valp = P.vals + (size_t)koff * 9 * 64 + lane;
  for (int k0 = first; k0 < kmax; k0 += U * step) {
    int j[U];
    T w[U][9];
    typename vec4<T>::type rj[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = k0 + u * step;
      if (k < kmax) {
        j[u] = colp[(size_t)k * 64];
#pragma unroll
        for (int e = 0; e < 9; e++) w[u][e] = valp[((size_t)k * 9 + e) * 64];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u * step < kmax) rj[u] = P.r[j[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u * step < len) {
        const T v[3] = {rj[u].x, rj[u].y, rj[u].z};
        mvp3x3<T>(w[u], v, y);
      }
    }
  }
}

// fc3d_compute_error.c:70-104 for the lane's slot: u = q + W r (blocks in the
// reference row order, the diagonal at its own position), u stored, and the
// local error term with the squared norms of r and u returned.
__device__ __forceinline__ void residual_slot(const SweepPtrs<double>& P, int s, int lane,
                                              double* e2, double* r2, double* u2) {
  const int slot = s * 64 + lane;
  const int len = P.len[slot];
  *e2 = 0.0;
  *r2 = 0.0;
  *u2 = 0.0;
  if (len < 0) return;
  double u[3];
#pragma unroll
  for (int e = 0; e < 3; e++) u[e] = P.q[(size_t)(s * 3 + e) * 64 + lane];
  const double4 rs = P.r[slot];
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
      const double4 rj = P.r[j];
      const double v[3] = {rj.x, rj.y, rj.z};
      mvp3x3<double>(w, v, u);
    }
  }
  if (dpos >= kmax) mvp3x3<double>(dg, rv, u);
  P.u[slot] = make_double4(u[0], u[1], u[2], 0.0);
  *e2 = unitary_error<double>(rv, u, P.mu[slot]);
  *r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
  *u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
}

}  // namespace fc3d
