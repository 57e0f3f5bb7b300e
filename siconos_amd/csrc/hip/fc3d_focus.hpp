// fc3d_focus.hpp — focused passes of the velocity-form sweep (no reference
// counterpart; enabled by fc3d_hip_params.focus_passes).
//
// On granular piles the coloured Gauss-Seidel sweep reaches a residual near
// the tolerance after a few hundred sweeps and then spends most of its sweeps
// on a handful of small hyperstatic clusters (a few bodies carrying many
// contacts), whose reactions drift along near-null directions of W: on a
// 1M-contact pile, after 300 sweeps the top 0.1% of contacts carry 90% of
// sum |r_new - r_old|^2.  A focus round gives those clusters extra
// Gauss-Seidel passes of their own:
//   1. the sweep that ends a focus interval stores r_new - r_old per contact
//      (BodyPtrs::fdiff);
//   2. k_focus_hist bins |r_new - r_old|^2 by the top bits of its FP32
//      representation, and k_focus_threshold picks the bin at which the
//      largest DevParams::focus_count contacts are reached (every contact
//      that moved at all, when fewer did); the round is skipped unless those
//      carry focus_share of the sum (a spread-out error is the sweep's own
//      work);
//   3. k_focus_mark marks the bodies of the selected contacts and
//      k_focus_collect lists, per colour, every contact touching a marked
//      body (the clusters and their first ring);
//   4. k_focus_passes runs `passes` Gauss-Seidel passes over the list in one
//      workgroup, colour after colour, with the sweep's own per-contact step
//      and the passes' own relaxation (DevParams::focus_omega).
// k_focus_clear, k_focus_hist and k_focus_threshold test whether a round is
// due; the threshold kernel records the round in SolveState::focus_last and
// sets FocusState::active, which the later kernels follow.
// Within a colour the list order comes from atomics, but a colour's contacts
// share no body, so the passes' results do not depend on it.  A round whose
// lists overflow their capacity is skipped whole, so the selection is
// deterministic too.  The result is still judged on the FP64 residual.
#pragma once
#include "fc3d_velocity.hpp"

namespace fc3d {

constexpr int FOCUS_BINS = 2048;  // top 11 bits of a non-negative float (sign bit 0)
constexpr int FOCUS_FINITE = 0x7f800000 >> 20;  // first key of +inf and NaN

struct FocusState {
  unsigned hist[FOCUS_BINS];
  int threshold;  // key of the selection: contacts with key >= threshold
  int round;      // mark of this round (grows across solves; bodies hold the last mark)
  int active;     // 1: this round selects, lists and runs its passes
  int rounds;     // rounds with passes in this solve
  int listed_sum; // contacts listed over those rounds
  int due;        // rounds due in this solve
  int nlocal;     // local body ids handed out by this round's k_focus_stage
  int lds_rounds; // rounds with passes whose body velocities were in LDS
  int pad;
  double last_share;  // the last due round: share of sum |dr|^2 its selection carried
  double last_total;  // ... and that sum
};

// At solve start: the per-solve statistics (the round mark keeps growing).
static __global__ void k_focus_reset(FocusState* fs) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    fs->active = 0;
    fs->rounds = 0;
    fs->listed_sum = 0;
    fs->due = 0;
    fs->lds_rounds = 0;
    fs->last_share = 0.0;
    fs->last_total = 0.0;
  }
}

__device__ __forceinline__ int focus_key(float d2) { return (int)(__float_as_uint(d2) >> 20); }

// A round is due once focus_every sweeps have passed since the last one
// (SolveState::focus_last), from sweep focus_after on, unless the solve
// stopped with the last sweep.  Rounds are enqueued at the end of each graph
// replay; k_focus_threshold decides and records, the later kernels follow
// FocusState::active.
__device__ __forceinline__ bool focus_round_due(const DevParams* prm, const SolveState* st,
                                                int want_phase) {
  return !st->stop && st->phase == want_phase && st->iter >= prm->focus_after &&
         st->iter - st->focus_last >= prm->focus_every;
}

static __global__ void k_focus_clear(const DevParams* prm, const SolveState* st, FocusState* fs,
                              int* cnt, int ncolours, int want_phase) {
  // (runs before k_focus_threshold, so it still sees the round as due)
  if (!focus_round_due(prm, st, want_phase)) return;
  for (int b = threadIdx.x; b < FOCUS_BINS; b += blockDim.x) fs->hist[b] = 0u;
  for (int c = threadIdx.x; c < ncolours; c += blockDim.x) cnt[c] = 0;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_focus_hist(const DevParams* prm, const SolveState* st,
                                                     FocusState* fs, const int* __restrict__ len,
                                                     const typename vec4<T>::type* __restrict__ d,
                                                     int nslots, int want_phase) {
  __shared__ unsigned h[FOCUS_BINS];
  if (!focus_round_due(prm, st, want_phase)) return;
  for (int b = threadIdx.x; b < FOCUS_BINS; b += blockDim.x) h[b] = 0u;
  __syncthreads();
  for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < nslots;
       slot += gridDim.x * blockDim.x) {
    if (len[slot] < 0) continue;
    const typename vec4<T>::type v = d[slot];
    const float d2 = float(v.x * v.x + v.y * v.y + v.z * v.z);
    atomicAdd(&h[focus_key(d2)], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < FOCUS_BINS; b += blockDim.x)
    if (h[b]) atomicAdd(&fs->hist[b], h[b]);
}

// The threshold bin and whether the round runs, in one workgroup of 1024
// threads: the bins' counts and masses (count times the bin's smallest value)
// go to LDS, suffix sums are taken by a block scan, and the highest bin at
// which the suffix count reaches focus_count is the threshold.  Keys from
// FOCUS_FINITE on hold infinities and NaNs, which no selection takes.
constexpr int FOCUS_SCAN = 1024;
static __global__ __launch_bounds__(FOCUS_SCAN) void k_focus_threshold(const DevParams* prm,
                                                                SolveState* st, FocusState* fs,
                                                                int want_phase) {
  __shared__ long cnt[FOCUS_SCAN];
  __shared__ double mass[FOCUS_SCAN];
  __shared__ int thr_sh;
  const int t = threadIdx.x;
  const bool due = focus_round_due(prm, st, want_phase);
  __syncthreads();  // every thread has read the state before thread 0 updates it
  if (!due) {
    if (t == 0) fs->active = 0;
    return;
  }
  // thread t owns bins 2t and 2t + 1; the scan runs from the top bin down,
  // so it is taken over the reversed thread order
  const int u = FOCUS_SCAN - 1 - t;
  long c = 0;
  double m = 0.0;
  for (int b = 2 * u; b < 2 * u + 2; b++) {
    if (b > 0 && b < FOCUS_FINITE) {
      const unsigned h = fs->hist[b];
      c += h;
      m += (double)h * (double)__uint_as_float((unsigned)b << 20);
    }
  }
  cnt[t] = c;
  mass[t] = m;
  if (t == 0) thr_sh = -1;
  __syncthreads();
  for (int off = 1; off < FOCUS_SCAN; off <<= 1) {  // inclusive scan over t
    const long ca = t >= off ? cnt[t - off] : 0;
    const double ma = t >= off ? mass[t - off] : 0.0;
    __syncthreads();
    cnt[t] += ca;
    mass[t] += ma;
    __syncthreads();
  }
  // bins above thread t's pair: the scan value of thread t - 1
  const long above = t > 0 ? cnt[t - 1] : 0;
  const long hi = fs->hist[2 * u + 1] * (long)(2 * u + 1 < FOCUS_FINITE);
  const int K = prm->focus_count;
  if (above < K && above + hi >= K) thr_sh = 2 * u + 1;
  else if (above + hi < K && cnt[t] >= K && 2 * u > 0) thr_sh = 2 * u;
  __syncthreads();
  if (t != 0) return;
  st->focus_last = st->iter;
  // fewer than focus_count contacts moved (a local solver that stops at
  // once on a contact already within its tolerance leaves r_new == r_old):
  // every one that moved is taken, from the lowest nonzero bin up
  const int thr = thr_sh < 0 && cnt[FOCUS_SCAN - 1] > 0 ? 1 : thr_sh;
  const double total = mass[FOCUS_SCAN - 1];
  double top = 0.0;
  if (thr > 0) {
    // mass of the bins >= thr: the scan through the pair holding thr, less
    // the pair's lower bin when thr is its upper one
    const int tt = FOCUS_SCAN - 1 - thr / 2;
    top = mass[tt];
    if (thr & 1) {
      const int b = thr - 1;
      if (b > 0) top -= (double)fs->hist[b] * (double)__uint_as_float((unsigned)b << 20);
    }
  }
  fs->round += 1;
  fs->due += 1;
  fs->nlocal = 0;
  fs->last_share = total > 0.0 ? top / total : 0.0;
  fs->last_total = total;
  fs->threshold = thr;
  fs->active = thr > 0 && total > 0.0 && top >= prm->focus_share * total;
}

static __global__ void k_focus_mark(const FocusState* fs,
                             const int* __restrict__ len, const int* __restrict__ ba,
                             const int* __restrict__ bb,
                             const float4* __restrict__ d32, const double4* __restrict__ d64,
                             int* __restrict__ bmark, int nslots) {
  if (!fs->active) return;
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots || len[slot] < 0) return;
  float d2;
  if (d32) {
    const float4 v = d32[slot];
    d2 = v.x * v.x + v.y * v.y + v.z * v.z;
  } else {
    const double4 v = d64[slot];
    d2 = float(v.x * v.x + v.y * v.y + v.z * v.z);
  }
  const int key = focus_key(d2);
  if (key < fs->threshold || key >= FOCUS_FINITE) return;
  const int r = fs->round;
  if (ba[slot] >= 0) bmark[ba[slot]] = r;
  if (bb[slot] >= 0) bmark[bb[slot]] = r;
}

static __global__ void k_focus_collect(const FocusState* fs, const int* __restrict__ len, const int* __restrict__ ba,
                                const int* __restrict__ bb, const int* __restrict__ bmark,
                                const int* __restrict__ slice_colour, int* cnt, int* list,
                                int capc, int nslots) {
  if (!fs->active) return;
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots || len[slot] < 0) return;
  const int r = fs->round;
  const bool hot = (ba[slot] >= 0 && bmark[ba[slot]] == r) || (bb[slot] >= 0 && bmark[bb[slot]] == r);
  if (!hot) return;
  const int c = slice_colour[slot >> 6];
  const int pos = atomicAdd(&cnt[c], 1);
  if (pos < capc) list[(size_t)c * capc + pos] = slot;
}

// Staged passes (diagonal M^-1): the listed contacts' data are copied once
// per round into a compact buffer, entry i of colour c at off[c] + j, each
// field a plane of N entries, so that a pass reads coalesced lines that stay
// in L2; the reactions live there during the passes and go back at the end.
// The bodies the entries touch get local ids 0 .. nlocal-1 (blist maps them
// back, -1 for an id lost to a race), so that the passes can keep their
// velocities in LDS.
template <typename T>
struct FocusStage {
  int* slot;   // [N] the entry's slot
  int* ba;     // [N]
  int* bb;     // [N]
  int* la;     // [N] local id of body a (-1: none)
  int* lb;     // [N]
  int* blist;  // [capl] body of each local id (-1: unused)
  unsigned long long* bkey;  // [nbodies] (round << 32) | local id of the body's last round
  int capl;    // 2 N: at most two ids are handed out per entry
  T* H;        // [36][N]; with frames, planes 0-14 hold the frame (row-major)
               // and both lever arms
  int frames = 0;  // 1: the entries are in the frames layout (FP32 stage)
  T* W;        // [9][N]
  T* mu;       // [N]
  T* minv;     // [12][N]: side a 0-5, side b 6-11 (diagonal M^-1)
  typename vec4<T>::type* r;  // [N]
  T* pli;      // [N] (null when the local solver keeps no PLI state)
  int N;
};

// off[c]: first entry of colour c (the list counts scanned); -1 in off[0]
// marks an overflow (the round is then skipped).
__device__ __forceinline__ void focus_offsets(const int* cnt, int ncolours, int capc, int* off) {
  int acc = 0;
  for (int c = 0; c < ncolours; c++) {
    off[c] = acc;
    if (cnt[c] > capc) {
      off[0] = -1;
      return;
    }
    acc += cnt[c];
  }
  off[ncolours] = acc;
}

constexpr int FOCUS_MAX_COLOURS = 64;

// The body's local id in this round (stamp = round << 32): the first entry
// that claims the body with a compare-and-swap gives it a fresh id.  The ids
// depend on the order of the claims, the passes' results do not (an id only
// places the body's velocity in LDS).
__device__ __forceinline__ int focus_local_body(unsigned long long* bkey, int body,
                                                unsigned long long stamp, int* nlocal, int* blist,
                                                int capl) {
  const unsigned long long old = bkey[body];
  if ((old >> 32) == (stamp >> 32)) return (int)(unsigned)old;
  const int id = atomicAdd(nlocal, 1);
  const unsigned long long prev = atomicCAS(&bkey[body], old, stamp | (unsigned)id);
  // prev is old (claimed here) or another claim of this round
  const bool won = prev == old;
  if (id < capl) blist[id] = won ? body : -1;
  return won ? id : (int)(unsigned)prev;
}

template <typename T>
__global__ void k_focus_stage(FocusState* fs, SweepPtrs<T> P, BodyPtrs<T> B,
                              const int* __restrict__ cnt, const int* __restrict__ list, int capc,
                              int ncolours, FocusStage<T> S) {
  if (!fs->active) return;
  __shared__ int off[FOCUS_MAX_COLOURS + 1];
  if (threadIdx.x == 0) focus_offsets(cnt, ncolours, capc, off);
  __syncthreads();
  if (off[0] < 0) return;
  const int total = off[ncolours];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    int c = 0;
    while (off[c + 1] <= i) c++;
    const int slot = list[(size_t)c * capc + (i - off[c])];
    const int s = slot >> 6, lane = slot & 63;
    S.slot[i] = slot;
    const int a = B.ba[slot], b = B.bb[slot];
    S.ba[i] = a;
    S.bb[i] = b;
    const unsigned long long stamp = (unsigned long long)(unsigned)fs->round << 32;
    S.la[i] = a >= 0 ? focus_local_body(S.bkey, a, stamp, &fs->nlocal, S.blist, S.capl) : -1;
    S.lb[i] = b >= 0 ? focus_local_body(S.bkey, b, stamp, &fs->nlocal, S.blist, S.capl) : -1;
    if (S.frames) {  // the frame and the arms, as the FP32 sweep reads them
#pragma unroll
      for (int e = 0; e < 15; e++) S.H[(size_t)e * S.N + i] = B.F[((size_t)s * 15 + e) * 64 + lane];
    } else if (B.F) {  // frames layout: the blocks rebuilt as the sweep rebuilds them
      T f[15], hh[36];
#pragma unroll
      for (int e = 0; e < 15; e++) f[e] = B.F[((size_t)s * 15 + e) * 64 + lane];
      frames_to_h<T>(f, hh);
#pragma unroll
      for (int e = 0; e < 36; e++) S.H[(size_t)e * S.N + i] = hh[e];
    } else {
#pragma unroll 4
      for (int e = 0; e < 36; e++) S.H[(size_t)e * S.N + i] = B.H[((size_t)s * 36 + e) * 64 + lane];
    }
#pragma unroll
    for (int e = 0; e < 9; e++) S.W[(size_t)e * S.N + i] = P.diag[((size_t)s * 9 + e) * 64 + lane];
    S.mu[i] = P.mu[slot];
    T ma[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, mb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (a >= 0) minv_diagonal<T>(B, (size_t)a, ma);
    if (b >= 0) minv_diagonal<T>(B, (size_t)b, mb);
#pragma unroll
    for (int l = 0; l < 6; l++) {
      S.minv[(size_t)l * S.N + i] = ma[l];
      S.minv[(size_t)(6 + l) * S.N + i] = mb[l];
    }
    S.r[i] = P.r[slot];
    if (S.pli) S.pli[i] = P.pli[slot];
  }
}

template <typename T>
__global__ void k_focus_unstage(const FocusState* fs, SweepPtrs<T> P, const int* __restrict__ cnt,
                                int capc, int ncolours, FocusStage<T> S) {
  if (!fs->active) return;
  __shared__ int off[FOCUS_MAX_COLOURS + 1];
  if (threadIdx.x == 0) focus_offsets(cnt, ncolours, capc, off);
  __syncthreads();
  if (off[0] < 0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < off[ncolours];
       i += gridDim.x * blockDim.x) {
    const int slot = S.slot[i];
    P.r[slot] = S.r[i];
    if (S.pli) P.pli[slot] = S.pli[i];
  }
}

// The velocity-form step of staged entry i: the arithmetic of vel_slot with
// the entry's data read from the stage (the reaction is updated there).
template <typename T>
__device__ __forceinline__ void focus_step(const FocusStage<T>& S, const BodyPtrs<T>& B,
                                           const StepConsts& sc, T tol, int i) {
  const size_t N = (size_t)S.N;
  SlotInputs<T> in;
#pragma unroll
  for (int e = 0; e < 9; e++) in.lp.W[e] = S.W[e * N + i];
  in.lp.mu = S.mu[i];
  in.ro = S.r[i];
  in.pli = S.pli ? S.pli[i] : T(0);
  const int kb[2] = {S.ba[i], S.bb[i]};
  T h[36];
#pragma unroll
  for (int e = 0; e < 15; e++) h[e] = S.H[e * N + i];
  if (!S.frames) {
#pragma unroll
    for (int e = 15; e < 36; e++) h[e] = S.H[e * N + i];
  }
  typename vec4<T>::type va[2][2];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const size_t k = kb[side] >= 0 ? (size_t)kb[side] : 0;
    va[side][0] = B.V[2 * k];
    va[side][1] = B.V[2 * k + 1];
  }
  T u[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (kb[side] < 0) continue;
    const T v[6] = {va[side][0].x, va[side][0].y, va[side][0].z,
                    va[side][1].x, va[side][1].y, va[side][1].z};
    if (S.frames) {
      frames_hv<T>(h, side, v, u);
    } else {
#pragma unroll
      for (int l = 0; l < 6; l++)
#pragma unroll
        for (int e = 0; e < 3; e++) u[e] += h[18 * side + e + 3 * l] * v[l];
    }
  }
  const T rold[3] = {in.ro.x, in.ro.y, in.ro.z};
  T wr[3] = {T(0), T(0), T(0)};
  mvp3x3<T>(in.lp.W, rold, wr);
#pragma unroll
  for (int e = 0; e < 3; e++) in.ql[e] = u[e] - wr[e];
  // finish_slot on the stage: its slot index is the entry's
  SweepPtrs<T> Pc;
  Pc.r = S.r;
  Pc.pli = S.pli;
  Pc.f = nullptr;
  typename vec4<T>::type out;
  double light, nr2;
  finish_slot<T>(Pc, sc, i >> 6, i & 63, tol, in, &light, &nr2, &out);
  const T dr[3] = {out.x - rold[0], out.y - rold[1], out.z - rold[2]};
  T w[3] = {T(0), T(0), T(0)};
  if (S.frames) frames_tdr<T>(h, dr, w);
  typename vec4<T>::type vn[2][2];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (kb[side] < 0) continue;
    T g[6], dv[6];
    if (S.frames)
      ht_dr<T, true>(h, side, dr, w, g);
    else
      ht_dr<T, false>(h, side, dr, w, g);
#pragma unroll
    for (int l = 0; l < 6; l++) dv[l] = S.minv[(6 * side + l) * N + i] * g[l];
    const bool again = side == 1 && kb[1] == kb[0];
    typename vec4<T>::type v0 = again ? vn[0][0] : va[side][0];
    typename vec4<T>::type v1 = again ? vn[0][1] : va[side][1];
    v0.x += dv[0]; v0.y += dv[1]; v0.z += dv[2];
    v1.x += dv[3]; v1.y += dv[4]; v1.z += dv[5];
    vn[side][0] = v0;
    vn[side][1] = v1;
    B.V[2 * (size_t)kb[side]] = v0;
    B.V[2 * (size_t)kb[side] + 1] = v1;
  }
}

// The passes with the bodies' velocities in LDS: one workgroup of
// FOCUS_THREADS threads, thread t taking entries off[c] + t, off[c] + t + nt,
// ... of each colour c.  A colour step is bound by the dependent arithmetic
// of one wave's local solves, not by memory: on the 1M-contact pile 256
// threads solve in 103 ms where 512 take 127 ms (with the velocities in HBM,
// 113 ms), and loading the next colour's entry into registers during the
// step (107 ms) costs more than the latency it hides.
constexpr int FOCUS_THREADS = 256;
constexpr int FOCUS_LDS_BYTES = 144 * 1024;

// One staged entry's data, loaded before its step.
template <typename T>
struct FocusItem {
  int i;  // the entry
  int la, lb;
  T W[9];
  T mu;
  typename vec4<T>::type r;
  T pli;
  T h[36];
  T minv[12];
};

template <typename T>
__device__ __forceinline__ void focus_fetch(const FocusStage<T>& S, int i, FocusItem<T>& it) {
  it.i = i;
  const size_t N = (size_t)S.N;
#pragma unroll
  for (int e = 0; e < 9; e++) it.W[e] = S.W[e * N + i];
  it.mu = S.mu[i];
  it.r = S.r[i];
  it.pli = S.pli ? S.pli[i] : T(0);
  it.la = S.la[i];
  it.lb = S.lb[i];
#pragma unroll
  for (int e = 0; e < 15; e++) it.h[e] = S.H[e * N + i];
  if (!S.frames) {
#pragma unroll
    for (int e = 15; e < 36; e++) it.h[e] = S.H[e * N + i];
  }
#pragma unroll
  for (int l = 0; l < 12; l++) it.minv[l] = S.minv[l * N + i];
}

// focus_step with the velocities in LDS (vl: six values per local body):
// the same arithmetic in the same order, so both give the same bits.
template <typename T>
__device__ __forceinline__ void focus_step_lds(const FocusStage<T>& S, T* vl,
                                               const StepConsts& sc, T tol,
                                               const FocusItem<T>& it) {
  SlotInputs<T> in;
#pragma unroll
  for (int e = 0; e < 9; e++) in.lp.W[e] = it.W[e];
  in.lp.mu = it.mu;
  in.ro = it.r;
  in.pli = it.pli;
  const int kb[2] = {it.la, it.lb};
  T va[2][6];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int k = kb[side] >= 0 ? kb[side] : 0;
#pragma unroll
    for (int l = 0; l < 6; l++) va[side][l] = vl[6 * k + l];
  }
  T u[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (kb[side] < 0) continue;
    if (S.frames) {
      frames_hv<T>(it.h, side, va[side], u);
    } else {
#pragma unroll
      for (int l = 0; l < 6; l++)
#pragma unroll
        for (int e = 0; e < 3; e++) u[e] += it.h[18 * side + e + 3 * l] * va[side][l];
    }
  }
  const T rold[3] = {in.ro.x, in.ro.y, in.ro.z};
  T wr[3] = {T(0), T(0), T(0)};
  mvp3x3<T>(in.lp.W, rold, wr);
#pragma unroll
  for (int e = 0; e < 3; e++) in.ql[e] = u[e] - wr[e];
  SweepPtrs<T> Pc;
  Pc.r = S.r;
  Pc.pli = S.pli;
  Pc.f = nullptr;
  typename vec4<T>::type out;
  double light, nr2;
  finish_slot<T>(Pc, sc, it.i >> 6, it.i & 63, tol, in, &light, &nr2, &out);
  const T dr[3] = {out.x - rold[0], out.y - rold[1], out.z - rold[2]};
  T w[3] = {T(0), T(0), T(0)};
  if (S.frames) frames_tdr<T>(it.h, dr, w);
  T vn0[6];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (kb[side] < 0) continue;
    const bool again = side == 1 && kb[1] == kb[0];
    T g[6], vn[6];
    if (S.frames)
      ht_dr<T, true>(it.h, side, dr, w, g);
    else
      ht_dr<T, false>(it.h, side, dr, w, g);
#pragma unroll
    for (int l = 0; l < 6; l++) {
      const T dv = it.minv[6 * side + l] * g[l];
      vn[l] = (again ? vn0[l] : va[side][l]) + dv;
    }
#pragma unroll
    for (int l = 0; l < 6; l++) {
      if (side == 0) vn0[l] = vn[l];
      vl[6 * kb[side] + l] = vn[l];
    }
  }
}

// The staged passes: with the velocities in LDS when the round's bodies fit
// (use_lds and nlocal <= the LDS capacity), else in HBM (focus_step).
template <typename T>
__global__ __launch_bounds__(FOCUS_THREADS) void k_focus_passes_staged(
    FocusStage<T> S, BodyPtrs<T> B, const DevParams* __restrict__ prm, const SolveState* st,
    FocusState* fs, const int* __restrict__ cnt, int capc, int ncolours, int passes,
    int want_phase, int use_lds) {
  constexpr int nt = FOCUS_THREADS;
  constexpr int capb = FOCUS_LDS_BYTES / (6 * (int)sizeof(T));
  __shared__ int off[FOCUS_MAX_COLOURS + 1];
  __shared__ int ne[FOCUS_MAX_COLOURS];  // the colours with entries, in order
  __shared__ int nne_sh;
  __shared__ __align__(16) T vl[6 * capb];
  if (!fs->active) return;
  if (threadIdx.x == 0) {
    focus_offsets(cnt, ncolours, capc, off);
    int n = 0;
    if (off[0] >= 0)
      for (int c = 0; c < ncolours; c++)
        if (off[c + 1] > off[c]) ne[n++] = c;
    nne_sh = n;
  }
  __syncthreads();
  if (off[0] < 0 || off[ncolours] == 0) return;
  const StepConsts sc = focus_consts<T>(prm);
  const T tol = T(want_phase ? prm->fp32_local_tol : st->local_tol);
  const int nne = nne_sh;
  const int nloc = fs->nlocal;
  if (use_lds && nloc <= capb) {
    if (threadIdx.x == 0) fs->lds_rounds += 1;
    for (int id = threadIdx.x; id < nloc; id += nt) {
      const int body = S.blist[id];
      if (body < 0) continue;
      const typename vec4<T>::type v0 = B.V[2 * (size_t)body], v1 = B.V[2 * (size_t)body + 1];
      T* p = vl + 6 * id;
      p[0] = v0.x; p[1] = v0.y; p[2] = v0.z;
      p[3] = v1.x; p[4] = v1.y; p[5] = v1.z;
    }
    __syncthreads();
    for (int pass = 0; pass < passes; pass++) {
      for (int q = 0; q < nne; q++) {
        const int c = ne[q];
        for (int i = off[c] + (int)threadIdx.x; i < off[c + 1]; i += nt) {
          FocusItem<T> x;
          focus_fetch<T>(S, i, x);
          focus_step_lds<T>(S, vl, sc, tol, x);
        }
        __syncthreads();
      }
    }
    for (int id = threadIdx.x; id < nloc; id += nt) {
      const int body = S.blist[id];
      if (body < 0) continue;
      const T* p = vl + 6 * id;
      typename vec4<T>::type v0 = B.V[2 * (size_t)body], v1 = B.V[2 * (size_t)body + 1];
      v0.x = p[0]; v0.y = p[1]; v0.z = p[2];
      v1.x = p[3]; v1.y = p[4]; v1.z = p[5];
      B.V[2 * (size_t)body] = v0;
      B.V[2 * (size_t)body + 1] = v1;
    }
  } else {
    for (int pass = 0; pass < passes; pass++) {
      for (int q = 0; q < nne; q++) {
        const int c = ne[q];
        for (int i = off[c] + (int)threadIdx.x; i < off[c + 1]; i += nt)
          focus_step<T>(S, B, sc, tol, i);
        __syncthreads();
      }
    }
  }
  if (threadIdx.x == 0) {
    fs->rounds += 1;
    fs->listed_sum += off[ncolours];
  }
}

// The passes, in one workgroup: colour after colour, a workgroup barrier
// between colours (the next colour reads this colour's body velocities).
// The workgroup is as large as k_sweep_vel's registers allow on one CU.
// FRAMES: the contacts' blocks are in the frames layout.
template <typename T, bool FRAMES>
__global__ __launch_bounds__(vel_tail_threads<T>()) void k_focus_passes(
    SweepPtrs<T> P, BodyPtrs<T> B, const DevParams* __restrict__ prm, const SolveState* st,
    FocusState* fs, const int* __restrict__ cnt, const int* __restrict__ list, int capc,
    int ncolours, int passes, int want_phase) {
  constexpr int nt = vel_tail_threads<T>();
  if (!fs->active) return;
  int listed = 0;
  for (int c = 0; c < ncolours; c++) {
    if (cnt[c] > capc) return;  // overflow: the round is skipped whole
    listed += cnt[c];
  }
  if (listed == 0) return;
  for (int pass = 0; pass < passes; pass++) {
    for (int c = 0; c < ncolours; c++) {
      const int n = cnt[c];
      if (n == 0) continue;  // uniform: every thread skips the empty colour
      const int* lc = list + (size_t)c * capc;
      for (int i = threadIdx.x; i < n; i += nt) {
        const int slot = lc[i];
        double light, nr2;
        vel_slot<T, false, FRAMES, true>(P, B, prm, st, slot >> 6, slot & 63, want_phase, light,
                                         nr2);
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    fs->rounds += 1;
    fs->listed_sum += listed;
  }
}

}  // namespace fc3d
