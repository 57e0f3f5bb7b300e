// fc3d_accel.hpp — Anderson acceleration of the NSGS sweep (no reference
// counterpart; off unless fc3d_hip_params.accel_depth > 0).
//
// The fixed-point map is one coloured sweep, x -> G(x).  With f_k = G(x_k) - x_k
// and the last m differences dF_i = f_{i+1} - f_i, dG_i = G(x_{i+1}) - G(x_i),
// an Anderson step replaces x_{k+1} = G(x_k) by
//     x_{k+1} = G(x_k) - dG gamma,   (dF^T dF + reg tr(dF^T dF) I) gamma = dF^T f_k.
// Three rules keep it stable on contact problems, whose map is only piecewise
// smooth:
//   * the step is taken on every `every`-th sweep only (plain sweeps between
//     Anderson steps restore the contact statuses an extrapolation upsets);
//   * the history restarts whenever |f_k| grows over the previous sweep;
//   * safeguard: a step is judged by the sweep that follows it.  If |f| there
//     exceeds |f_k| of the sweep the step extrapolated, the step is rejected:
//     x returns to G(x_k) (kept in gprev), the history restarts and no step is
//     taken for `backoff` sweeps (2, doubling per rejection up to 64, halved
//     per accepted step).  The rejected sweep counts as a sweep.
// Everything runs on the device from device state, so the kernels sit in the
// captured sweep graphs: the sweep writes f (SweepPtrs::f), then
//   k_aa_update    new history column and 2m fixed-order dot products, or
//                  the restore of a rejected step,
//   k_aa_finalize  Gram matrix update, restart and safeguard rules, small
//                  solve for gamma,
//   k_aa_apply     x -= dG gamma.
#pragma once
#include "fc3d_sweep.hpp"

namespace fc3d {

constexpr int AA_MAX = 8;

struct AccelState {
  int count;      // valid history columns
  int head;       // ring position of the next column
  int have_prev;  // fprev / gprev hold the previous sweep
  int apply;      // 1: k_aa_apply uses gamma this sweep
  int sweeps;     // sweeps seen in this phase
  int resets;
  int steps;      // Anderson steps taken
  int pending;    // 1: the last sweep started from an Anderson step
  int rejects;    // steps undone by the safeguard
  int backoff;    // sweeps without steps after a rejection
  int wait_until; // no step before this sweep count
  int pad;
  double nf_prev;
  double nf_ref;  // |f| of the sweep the pending step extrapolated
  double gram[AA_MAX][AA_MAX];
  double gamma[AA_MAX];
};

template <typename T>
struct AccelBufs {
  using V = typename vec4<T>::type;
  V* f;       // written by the sweep
  V* fprev;
  V* gprev;
  V* dF;      // [AA_MAX][nslots] ring
  V* dG;      // [AA_MAX][nslots] ring
  double* part;  // [blocks][2 AA_MAX]
  // velocity form: body velocities follow r linearly, so they keep their own
  // difference history and take the same step (2 entries per body)
  V* vprev;   // [2 nbodies]
  V* dV;      // [AA_MAX][2 nbodies] ring
};

template <typename V>
__device__ __forceinline__ double dot3(const V& a, const V& b) {
  return double(a.x) * double(b.x) + double(a.y) * double(b.y) + double(a.z) * double(b.z);
}

// Safeguard / restart / new-column decision, identical in every thread of
// every kernel that runs before k_aa_finalize changes the state.
__device__ __forceinline__ void aa_decide(const SolveState* st, const AccelState* as, int* reset,
                                          int* form, int* reject) {
  *reject = (as->pending && st->nf > as->nf_ref) ? 1 : 0;
  *reset = (!*reject && as->count > 0 && st->nf > as->nf_prev) ? 1 : 0;
  *form = (as->have_prev && !*reset && !*reject) ? 1 : 0;
}

__device__ __forceinline__ bool aa_valid(int i, int head_after, int count_after, int m) {
  const int age = (head_after - 1 - i + 2 * m) % m;  // 0: newest
  return age < count_after;
}

template <typename T>
__global__ __launch_bounds__(256) void k_aa_update(typename vec4<T>::type* __restrict__ r,
                                                   AccelBufs<T> A, const SolveState* st,
                                                   const AccelState* as, const DevParams* prm,
                                                   int nslots, int want_phase) {
  __shared__ double sh[4][2 * AA_MAX];
  if (st->stop || st->phase != want_phase) return;
  const int m = prm->accel_depth;
  int reset, form, reject;
  aa_decide(st, as, &reset, &form, &reject);
  if (reject) {
    // back to G(x_k), the point the rejected step extrapolated
    for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < nslots;
         slot += gridDim.x * blockDim.x)
      r[slot] = A.gprev[slot];
    return;
  }
  const int head = as->head;
  const int count_after = form ? min(as->count + 1, m) : 0;
  const int head_after = form ? (head + 1) % m : head;
  double acc[2 * AA_MAX];
#pragma unroll
  for (int j = 0; j < 2 * AA_MAX; j++) acc[j] = 0.0;
  // grid-stride over slots: each thread adds its slots in a fixed order
  for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < nslots;
       slot += gridDim.x * blockDim.x) {
    using V = typename vec4<T>::type;
    const V f = A.f[slot];
    const V g = r[slot];
    if (form) {
      const V fp = A.fprev[slot], gp = A.gprev[slot];
      V dF = f, dG = g;
      dF.x -= fp.x; dF.y -= fp.y; dF.z -= fp.z; dF.w = T(0);
      dG.x -= gp.x; dG.y -= gp.y; dG.z -= gp.z; dG.w = T(0);
      A.dF[(size_t)head * nslots + slot] = dF;
      A.dG[(size_t)head * nslots + slot] = dG;
#pragma unroll
      for (int i = 0; i < AA_MAX; i++) {
        if (i >= m || !aa_valid(i, head_after, count_after, m)) continue;
        const V c = i == head ? dF : A.dF[(size_t)i * nslots + slot];
        acc[i] += dot3(c, dF);
        acc[AA_MAX + i] += dot3(c, f);
      }
    }
    A.fprev[slot] = f;
    A.gprev[slot] = g;
  }
  if (!form) return;  // uniform over the grid
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 2 * AA_MAX; j++) {
    const double v = wave_sum(acc[j]);
    if (lane == 0) sh[wave][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < 2 * AA_MAX) {
    const int j = threadIdx.x;
    A.part[(size_t)blockIdx.x * 2 * AA_MAX + j] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}

// Gaussian elimination with partial pivoting on n <= AA_MAX unknowns; false
// if a pivot vanishes or the result is not finite.
__device__ inline bool aa_solve(double (*a)[AA_MAX], double* b, int n) {
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int i = c + 1; i < n; i++)
      if (fabs(a[i][c]) > fabs(a[p][c])) p = i;
    if (!(fabs(a[p][c]) > 0.0)) return false;
    if (p != c) {
      for (int j = 0; j < n; j++) {
        const double t = a[c][j];
        a[c][j] = a[p][j];
        a[p][j] = t;
      }
      const double t = b[c];
      b[c] = b[p];
      b[p] = t;
    }
    for (int i = c + 1; i < n; i++) {
      const double l = a[i][c] / a[c][c];
      for (int j = c; j < n; j++) a[i][j] -= l * a[c][j];
      b[i] -= l * b[c];
    }
  }
  for (int c = n - 1; c >= 0; c--) {
    double v = b[c];
    for (int j = c + 1; j < n; j++) v -= a[c][j] * b[j];
    b[c] = v / a[c][c];
    if (!isfinite(b[c])) return false;
  }
  return true;
}

static __global__ __launch_bounds__(1024) void k_aa_finalize(const double* __restrict__ part,
                                                      int nblocks, const SolveState* st,
                                                      AccelState* as, const DevParams* prm,
                                                      int want_phase) {
  __shared__ double sh[16][2 * AA_MAX];
  if (st->stop || st->phase != want_phase) return;
  const int m = prm->accel_depth;
  int reset, form, reject;
  aa_decide(st, as, &reset, &form, &reject);
  double S[2 * AA_MAX];
#pragma unroll
  for (int j = 0; j < 2 * AA_MAX; j++) S[j] = 0.0;
  if (form) {
    const double2* p2 = reinterpret_cast<const double2*>(part);
    for (int b = threadIdx.x; b < nblocks; b += 1024) {
      double2 t[AA_MAX];
#pragma unroll
      for (int j = 0; j < AA_MAX; j++) t[j] = p2[(size_t)b * AA_MAX + j];
#pragma unroll
      for (int j = 0; j < AA_MAX; j++) {
        S[2 * j] += t[j].x;
        S[2 * j + 1] += t[j].y;
      }
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 2 * AA_MAX; j++) {
    const double v = wave_sum(S[j]);
    if (lane == 0) sh[wave][j] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  as->sweeps += 1;
  as->apply = 0;
  if (reject) {
    as->pending = 0;
    as->rejects += 1;
    as->count = 0;
    as->have_prev = 0;  // the restored point has no sweep difference yet
    as->backoff = min(max(2, 2 * as->backoff), 64);
    as->wait_until = as->sweeps + as->backoff;
    return;
  }
  if (as->pending) {
    as->pending = 0;
    as->backoff /= 2;
  }
  if (reset) {
    as->count = 0;
    as->resets += 1;
  } else if (form) {
    double tot[2 * AA_MAX];
    for (int j = 0; j < 2 * AA_MAX; j++) {
      double v = 0.0;
      for (int w = 0; w < 16; w++) v += sh[w][j];
      tot[j] = v;
    }
    const int head = as->head;
    const int count_after = min(as->count + 1, m);
    const int head_after = (head + 1) % m;
    for (int i = 0; i < m; i++)
      if (aa_valid(i, head_after, count_after, m)) {
        as->gram[head][i] = tot[i];
        as->gram[i][head] = tot[i];
      }
    as->head = head_after;
    as->count = count_after;
    for (int i = 0; i < AA_MAX; i++) as->gamma[i] = 0.0;
    if (as->sweeps % prm->accel_every == 0 && as->sweeps >= as->wait_until) {
      // compact the valid ring columns, oldest first
      int idx[AA_MAX], n = 0;
      for (int age = count_after - 1; age >= 0; age--) idx[n++] = (head_after - 1 - age + 2 * m) % m;
      double a[AA_MAX][AA_MAX], b[AA_MAX];
      double tr = 0.0;
      for (int i = 0; i < n; i++) tr += as->gram[idx[i]][idx[i]];
      for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) a[i][j] = as->gram[idx[i]][idx[j]];
        a[i][i] += prm->accel_reg * tr;
        b[i] = tot[AA_MAX + idx[i]];
      }
      if (tr > 0.0 && aa_solve(a, b, n)) {
        for (int i = 0; i < n; i++) as->gamma[idx[i]] = b[i];
        as->apply = 1;
        as->steps += 1;
        as->pending = 1;
        as->nf_ref = st->nf;
      }
    }
  }
  as->have_prev = 1;
  as->nf_prev = st->nf;
}

template <typename T>
__global__ __launch_bounds__(256) void k_aa_apply(typename vec4<T>::type* __restrict__ r,
                                                  AccelBufs<T> A, const SolveState* st,
                                                  const AccelState* as, const DevParams* prm,
                                                  int nslots, int want_phase) {
  if (st->stop || st->phase != want_phase || !as->apply) return;
  const int m = prm->accel_depth;
  for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < nslots;
       slot += gridDim.x * blockDim.x) {
    typename vec4<T>::type x = r[slot];
    double x0 = x.x, x1 = x.y, x2 = x.z;
    for (int i = 0; i < m; i++) {
      const double gmm = as->gamma[i];
      if (gmm == 0.0) continue;
      const typename vec4<T>::type d = A.dG[(size_t)i * nslots + slot];
      x0 -= gmm * double(d.x);
      x1 -= gmm * double(d.y);
      x2 -= gmm * double(d.z);
    }
    x.x = T(x0);
    x.y = T(x1);
    x.z = T(x2);
    r[slot] = x;
  }
}

// Velocity form: body-velocity history column (same decision as k_aa_update).
template <typename T>
__global__ __launch_bounds__(256) void k_aa_update_v(typename vec4<T>::type* __restrict__ Vb,
                                                     AccelBufs<T> A, const SolveState* st,
                                                     const AccelState* as, int nv,
                                                     int want_phase) {
  if (st->stop || st->phase != want_phase) return;
  int reset, form, reject;
  aa_decide(st, as, &reset, &form, &reject);
  if (reject) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x)
      Vb[i] = A.vprev[i];
    return;
  }
  const int head = as->head;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    const typename vec4<T>::type v = Vb[i];
    if (form) {
      typename vec4<T>::type d = v;
      const typename vec4<T>::type vp = A.vprev[i];
      d.x -= vp.x; d.y -= vp.y; d.z -= vp.z; d.w = T(0);
      A.dV[(size_t)head * nv + i] = d;
    }
    A.vprev[i] = v;
  }
}

// Velocity form: V -= dV gamma, the Anderson step of r mapped to the bodies.
template <typename T>
__global__ __launch_bounds__(256) void k_aa_apply_v(typename vec4<T>::type* __restrict__ Vb,
                                                    AccelBufs<T> A, const SolveState* st,
                                                    const AccelState* as, const DevParams* prm,
                                                    int nv, int want_phase) {
  if (st->stop || st->phase != want_phase || !as->apply) return;
  const int m = prm->accel_depth;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    typename vec4<T>::type x = Vb[i];
    double x0 = x.x, x1 = x.y, x2 = x.z;
    for (int j = 0; j < m; j++) {
      const double gmm = as->gamma[j];
      if (gmm == 0.0) continue;
      const typename vec4<T>::type d = A.dV[(size_t)j * nv + i];
      x0 -= gmm * double(d.x);
      x1 -= gmm * double(d.y);
      x2 -= gmm * double(d.z);
    }
    x.x = T(x0);
    x.y = T(x1);
    x.z = T(x2);
    Vb[i] = x;
  }
}

}  // namespace fc3d
