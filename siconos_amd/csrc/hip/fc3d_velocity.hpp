// fc3d_velocity.hpp — the velocity form of the coloured sweep (no reference
// counterpart; enabled by fc3d_hip_params.velocity_form once the factors of
// W = H M^-1 H^T are attached with fc3d_hip_problem_set_bodies).
//
// With body velocities v = v_free + M^-1 H^T r kept up to date, the local
// velocity of contact c is u_c = H_{c,a} v_a + H_{c,b} v_b, so
//     q_loc = q_c + sum_{j != c} W_cj r_j = u_c - W_cc r_c,
// the same local problem as the matrix form in exact arithmetic.  After the
// local solve, v_a += M_a^-1 H_{c,a}^T (r_new - r_old) and likewise for b.
// Contacts of one colour share no body, so a colour's updates never collide.
// Per contact this reads the 3x6 blocks of H and two body velocities instead
// of the ~k neighbour blocks of W and k neighbour reactions.
//
// Frames layout: when every contact's blocks have the rigid-body form
//     H_{c,a} = [ T | T S(l_a) ],   H_{c,b} = -[ T | T S(l_b) ]
// (S(l) w = l x w; T the contact frame, l the lever arms, as the reference's
// NewtonEuler3DR Jacobians are built), only T and the two lever arms are
// stored (15 values instead of 36) and the blocks are rebuilt in registers;
// W_cc is read from W's diagonal blocks in both layouts.
#pragma once
#include "fc3d_accel.hpp"

namespace fc3d {

template <typename T>
struct BodyPtrs {
  const int* ba;     // per slot: body of side a, -1 for none or padding
  const int* bb;     // per slot: body of side b
  const T* H;        // [slice][36][64]: side a entries 0-17, side b 18-35 (3x6 column-major)
  const T* F;        // frames layout: [slice][15][64]: T row-major (0-8), l_a (9-11), l_b (12-14)
  const T* minv;     // per body: 6 diagonal entries, or 36 (6x6 column-major),
                     // or with minv_iso 2: (m_t, m_r), M^-1 = diag(m_t I3, m_r I3)
  int minv_diag;
  int minv_iso = 0;
  typename vec4<T>::type* V;  // per body: (v0, v1, v2, _), (v3, v4, v5, _); in FP32 with
                              // minv_iso the fourth components hold m_t and m_r, so the
                              // velocity gather brings M^-1 along
  const double* vfree;        // per body: 6
  const int* inc_ptr;         // per body: range in inc
  const int* inc;             // 2 slot + side, ascending slot
  // focused passes: r_new - r_old of the sweep that ends a focus interval
  typename vec4<T>::type* fdiff = nullptr;
};

// The diagonal of body k's M^-1 (minv_diag), from 6 entries or, for bodies
// with isotropic translational and rotational parts, from 2: the same
// values either way.
template <typename T>
__device__ __forceinline__ void minv_diagonal(const BodyPtrs<T>& B, size_t k, T* m) {
  if (B.minv_iso) {
    const T mt = B.minv[2 * k], mr = B.minv[2 * k + 1];
    m[0] = m[1] = m[2] = mt;
    m[3] = m[4] = m[5] = mr;
  } else {
#pragma unroll
    for (int l = 0; l < 6; l++) m[l] = B.minv[6 * k + l];
  }
}

template <typename T>
__device__ __forceinline__ void minv_apply(const BodyPtrs<T>& B, int k, const T* g, T* dv) {
  if (B.minv_diag) {
    T m[6];
    minv_diagonal<T>(B, (size_t)k, m);
#pragma unroll
    for (int l = 0; l < 6; l++) dv[l] = m[l] * g[l];
  } else {
    const T* m = B.minv + 36 * (size_t)k;
#pragma unroll
    for (int l = 0; l < 6; l++) {
      T s = T(0);
#pragma unroll
      for (int t = 0; t < 6; t++) s += m[l + 6 * t] * g[t];
      dv[l] = s;
    }
  }
}

// The two 3x6 blocks of a contact (column-major, side a then side b) from its
// frame and lever arms: (T S(l))_{ij} = sum_k T_ik S(l)_kj.
template <typename T>
__device__ __forceinline__ void frames_to_h(const T* f, T* h) {
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const T* l = f + 9 + 3 * side;
    const T S[9] = {T(0), -l[2], l[1], l[2], T(0), -l[0], -l[1], l[0], T(0)};
    const T sg = side ? T(-1) : T(1);
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
      for (int j = 0; j < 3; j++) {
        h[18 * side + i + 3 * j] = sg * f[3 * i + j];
        const T rot = f[3 * i] * S[j] + f[3 * i + 1] * S[3 + j] + f[3 * i + 2] * S[6 + j];
        h[18 * side + i + 3 * (3 + j)] = sg * rot;
      }
    }
  }
}

// Loads a slot's 3x6 blocks from either layout (non-temporal with NT).
template <typename T, bool FRAMES, bool NT = false>
__device__ __forceinline__ void load_h(const BodyPtrs<T>& B, int s, int lane, T* h) {
  if (FRAMES) {
    T f[15];
#pragma unroll
    for (int e = 0; e < 15; e++) f[e] = ld<NT>(&B.F[((size_t)s * 15 + e) * 64 + lane]);
    frames_to_h<T>(f, h);
  } else {
#pragma unroll
    for (int e = 0; e < 36; e++) h[e] = ld<NT>(&B.H[((size_t)s * 36 + e) * 64 + lane]);
  }
}

// v_k = v_free,k + M_k^-1 sum_{c at k} H_{c,k}^T r_c for every body (in a fixed
// order of its contacts).  With `as`, only after an Anderson step changed r.
template <typename T, bool FRAMES>
__global__ void k_vinit(BodyPtrs<T> B, const typename vec4<T>::type* __restrict__ r, int nbodies,
                        const SolveState* st, const AccelState* as, int want_phase) {
  if (st && (st->stop || st->phase != want_phase)) return;
  if (as && !as->apply) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbodies) return;
  T g[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int t = B.inc_ptr[k]; t < B.inc_ptr[k + 1]; t++) {
    const int e = B.inc[t], slot = e >> 1, side = e & 1;
    const int s = slot >> 6, lane = slot & 63;
    const typename vec4<T>::type rc = r[slot];
    T hh[36];
    const T* h;
    if (FRAMES) {
      load_h<T, FRAMES>(B, s, lane, hh);
      h = hh + 18 * side;
    } else {
      // only this side's 18 entries (the other side's are not needed)
#pragma unroll
      for (int e = 0; e < 18; e++) hh[e] = B.H[((size_t)s * 36 + 18 * side + e) * 64 + lane];
      h = hh;
    }
#pragma unroll
    for (int l = 0; l < 6; l++)
      g[l] += h[3 * l] * rc.x + h[3 * l + 1] * rc.y + h[3 * l + 2] * rc.z;
  }
  T dv[6];
  minv_apply<T>(B, k, g, dv);
  typename vec4<T>::type v0, v1;
  v0.x = T(B.vfree[6 * (size_t)k]) + dv[0];
  v0.y = T(B.vfree[6 * (size_t)k + 1]) + dv[1];
  v0.z = T(B.vfree[6 * (size_t)k + 2]) + dv[2];
  v0.w = T(0);
  v1.x = T(B.vfree[6 * (size_t)k + 3]) + dv[3];
  v1.y = T(B.vfree[6 * (size_t)k + 4]) + dv[4];
  v1.z = T(B.vfree[6 * (size_t)k + 5]) + dv[5];
  v1.w = T(0);
  if (sizeof(T) == 4 && B.minv_iso) {
    // (m_t, m_r) beside the velocity: every later writer of V keeps the
    // fourth components as they are
    v0.w = B.minv[2 * (size_t)k];
    v1.w = B.minv[2 * (size_t)k + 1];
  }
  B.V[2 * (size_t)k] = v0;
  B.V[2 * (size_t)k + 1] = v1;
}

// u += H_k v for side k in the frames layout (f: the frame T row-major and
// both lever arms; v = (v_lin, v_ang)): sg T (v_lin + l_k x v_ang), sg = -1
// on side b.
template <typename T>
__device__ __forceinline__ void frames_hv(const T* f, int side, const T* v, T* u) {
  const T* l = f + 9 + 3 * side;
  const T a[3] = {v[0] + (l[1] * v[5] - l[2] * v[4]), v[1] + (l[2] * v[3] - l[0] * v[5]),
                  v[2] + (l[0] * v[4] - l[1] * v[3])};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const T t = f[3 * i] * a[0] + f[3 * i + 1] * a[1] + f[3 * i + 2] * a[2];
    u[i] += side ? -t : t;
  }
}

// w = T^T dr in the frames layout.
template <typename T>
__device__ __forceinline__ void frames_tdr(const T* f, const T* dr, T* w) {
#pragma unroll
  for (int j = 0; j < 3; j++) w[j] = f[j] * dr[0] + f[3 + j] * dr[1] + f[6 + j] * dr[2];
}

// g = H_k^T dr for side k of a contact: from the 3x6 blocks h, or (DIRECT:
// the FP32 frames layout, h holding the frame T row-major and both lever
// arms) as sg (w, w x l_k) with w = T^T dr, sg = -1 on side b.
template <typename T, bool DIRECT>
__device__ __forceinline__ void ht_dr(const T* h, int side, const T* dr, const T* w, T* g) {
  if constexpr (DIRECT) {
    const T* l = h + 9 + 3 * side;
    const T sg = side ? T(-1) : T(1);
    g[0] = sg * w[0];
    g[1] = sg * w[1];
    g[2] = sg * w[2];
    g[3] = sg * (w[1] * l[2] - w[2] * l[1]);
    g[4] = sg * (w[2] * l[0] - w[0] * l[2]);
    g[5] = sg * (w[0] * l[1] - w[1] * l[0]);
  } else {
#pragma unroll
    for (int l = 0; l < 6; l++)
      g[l] = h[18 * side + 3 * l] * dr[0] + h[18 * side + 3 * l + 1] * dr[1] +
             h[18 * side + 3 * l + 2] * dr[2];
  }
}

// The velocity-form step of one contact, slot s * 64 + lane (s need not be
// wave-uniform): returns false, with nothing written, when CHECK finds the
// stop flag set; light and nr2 receive the slot's light-error term and
// squared reaction (0 for padding).  With B.fdiff set, the sweep that ends a
// focus interval also stores r_new - r_old.  FOCUS: a focused pass's step
// (focus_consts).
template <typename T, bool CHECK, bool FRAMES, bool FOCUS = false>
__device__ __forceinline__ bool vel_slot(const SweepPtrs<T>& P, const BodyPtrs<T>& B,
                                         const DevParams* __restrict__ prm,
                                         const SolveState* __restrict__ st, int s, int lane,
                                         int want_phase, double& light, double& nr2) {
  const int slot = s * 64 + lane;
  light = 0.0;
  nr2 = 0.0;
  // Every load of the slot is issued before the stop flag is tested, so the
  // flag's latency overlaps theirs (all addresses are in bounds: padding
  // slots read body 0).  The empty asm makes the loaded values live before
  // the test, which keeps the compiler from sinking the loads below it.
  const int stop = CHECK ? st->stop : 0, phase = CHECK ? st->phase : want_phase;
  const bool record = B.fdiff && st->iter + 1 - st->focus_last >= prm->focus_every;
  const StepConsts sc = FOCUS ? focus_consts<T>(prm) : step_consts<T>(prm);
  const T tol = T(want_phase ? prm->fp32_local_tol : st->local_tol);
  const int len = P.len[slot];
  SlotInputs<T> in;
  // FP64: the streamed planes are loaded non-temporal, which keeps V and M^-1
  // in L2 (5% a sweep); in FP32 it costs 9%
  constexpr bool NT = sizeof(T) == 8;
  // W_cc is read in both layouts (the frames layout stores it beside T and
  // the lever arms: rebuilding it needs M^-1 before the solve, one more
  // dependent load)
  load_slot<T, NT>(P, s, lane, &in);
  const int kb[2] = {B.ba[slot], B.bb[slot]};
  // FP32 frames layout: H is never formed.  H_k v = sg T (v_lin + l_k x
  // v_ang) and H_k^T dr = sg (w, w x l_k), w = T^T dr: fewer operations on
  // the dependent chain and 21 registers fewer than the 36 entries of H.
  constexpr bool DIRECT = FRAMES && sizeof(T) == 4;
  T h[DIRECT ? 15 : 36];
  if constexpr (DIRECT) {
#pragma unroll
    for (int e = 0; e < 15; e++) h[e] = B.F[((size_t)s * 15 + e) * 64 + lane];
  } else {
    load_h<T, FRAMES, NT>(B, s, lane, h);
  }
  typename vec4<T>::type va[2][2];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const size_t k = kb[side] >= 0 ? (size_t)kb[side] : 0;
    va[side][0] = B.V[2 * k];
    va[side][1] = B.V[2 * k + 1];
  }
  step_consts_ready(sc);
  asm volatile("" ::"s"(tol));
  if (CHECK) {
    asm volatile("" ::"v"(len), "v"(in.lp.mu), "v"(in.ro.x), "v"(in.pli), "v"(h[0]),
                 "v"(h[DIRECT ? 14 : 35]), "v"(va[0][0].x), "v"(va[1][1].x), "s"(stop),
                 "s"(phase));
    if (stop || phase != want_phase) return false;
  }
  if (len >= 0) {
    // u = H_a v_a + H_b v_b, then q_loc = u - W_cc r_old
    T u[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int side = 0; side < 2; side++) {
      if (kb[side] < 0) continue;
      if constexpr (DIRECT) {
        const T v[6] = {va[side][0].x, va[side][0].y, va[side][0].z,
                        va[side][1].x, va[side][1].y, va[side][1].z};
        frames_hv<T>(h, side, v, u);
      } else {
        const T v[6] = {va[side][0].x, va[side][0].y, va[side][0].z,
                        va[side][1].x, va[side][1].y, va[side][1].z};
#pragma unroll
        for (int l = 0; l < 6; l++)
#pragma unroll
          for (int i = 0; i < 3; i++) u[i] += h[18 * side + i + 3 * l] * v[l];
      }
    }
    const T rold[3] = {in.ro.x, in.ro.y, in.ro.z};
    T wr[3] = {T(0), T(0), T(0)};
    mvp3x3<T>(in.lp.W, rold, wr);
#pragma unroll
    for (int i = 0; i < 3; i++) in.ql[i] = u[i] - wr[i];
    typename vec4<T>::type out;
    finish_slot<T>(P, sc, s, lane, tol, in, &light, &nr2, &out);
    const T dr[3] = {out.x - rold[0], out.y - rold[1], out.z - rold[2]};
    if (record) {
      typename vec4<T>::type d;
      d.x = dr[0]; d.y = dr[1]; d.z = dr[2]; d.w = T(0);
      B.fdiff[slot] = d;
    }
    // dv = M^-1 H^T dr for both sides.  With diagonal M^-1 both sides'
    // entries are loaded at once, so the update waits on one round trip
    // (none in FP32 with isotropic bodies: M^-1 rides in the velocities).
    // FP32 keeps the body velocities loaded before the solve; FP64 reads
    // them back with M^-1 (keeping them live through its solve spills more).
    T dv[2][6];
    typename vec4<T>::type vb[2][2];
    T w[3] = {T(0), T(0), T(0)};  // DIRECT: T^T dr, shared by both sides
    if constexpr (DIRECT) frames_tdr<T>(h, dr, w);
    if (B.minv_diag) {
      T m[2][6];
#pragma unroll
      for (int side = 0; side < 2; side++) {
        const size_t k = kb[side] >= 0 ? (size_t)kb[side] : 0;
        if (sizeof(T) == 4 && B.minv_iso) {
          // FP32, isotropic bodies: M^-1 came with the velocities (their
          // fourth components), so no load follows the solve
          m[side][0] = m[side][1] = m[side][2] = va[side][0].w;
          m[side][3] = m[side][4] = m[side][5] = va[side][1].w;
        } else {
          minv_diagonal<T>(B, k, m[side]);
        }
        if constexpr (sizeof(T) == 8) {
          vb[side][0] = B.V[2 * k];
          vb[side][1] = B.V[2 * k + 1];
        } else {
          vb[side][0] = va[side][0];
          vb[side][1] = va[side][1];
        }
      }
#pragma unroll
      for (int side = 0; side < 2; side++) {
        if (kb[side] < 0) continue;
        T g[6];
        ht_dr<T, DIRECT>(h, side, dr, w, g);
#pragma unroll
        for (int l = 0; l < 6; l++) dv[side][l] = m[side][l] * g[l];
      }
    } else {
#pragma unroll
      for (int side = 0; side < 2; side++) {
        if (kb[side] < 0) continue;
        T g[6];
        ht_dr<T, DIRECT>(h, side, dr, w, g);
        minv_apply<T>(B, kb[side], g, dv[side]);
        if constexpr (sizeof(T) == 8) {
          vb[side][0] = B.V[2 * (size_t)kb[side]];
          vb[side][1] = B.V[2 * (size_t)kb[side] + 1];
        } else {
          vb[side][0] = va[side][0];
          vb[side][1] = va[side][1];
        }
      }
    }
    // a colour's contacts share no body, so vb holds the bodies' current
    // velocities; when both sides name the same body, side b adds to side
    // a's sum
    typename vec4<T>::type vn[2][2];
#pragma unroll
    for (int side = 0; side < 2; side++) {
      if (kb[side] < 0) continue;
      const bool again = side == 1 && kb[1] == kb[0];
      typename vec4<T>::type v0 = again ? vn[0][0] : vb[side][0];
      typename vec4<T>::type v1 = again ? vn[0][1] : vb[side][1];
      v0.x += dv[side][0]; v0.y += dv[side][1]; v0.z += dv[side][2];
      v1.x += dv[side][3]; v1.y += dv[side][4]; v1.z += dv[side][5];
      vn[side][0] = v0;
      vn[side][1] = v1;
      B.V[2 * (size_t)kb[side]] = v0;
      B.V[2 * (size_t)kb[side] + 1] = v1;
    }
  }
  return true;
}

// One slice (64 contacts) of one colour in the velocity form, for the
// calling wave; writes the slice's partial sums.  With CHECK, the stop flag
// is tested after the slot's loads are issued (the caller tests it
// otherwise) and the function returns without writing if it is set.
template <typename T, bool CHECK, bool FRAMES>
__device__ __forceinline__ void vel_slice(const SweepPtrs<T>& P, const BodyPtrs<T>& B,
                                          const DevParams* __restrict__ prm,
                                          const SolveState* __restrict__ st, int s, int lane,
                                          int want_phase) {
  double light, nr2;
  if (!vel_slot<T, CHECK, FRAMES>(P, B, prm, st, s, lane, want_phase, light, nr2)) return;
  light = wave_sum(light);
  nr2 = wave_sum(nr2);
  if (lane == 0) {
    P.part[2 * (size_t)s] = light;
    P.part[2 * (size_t)s + 1] = nr2;
  }
}

// One colour of one sweep in the velocity form: one wave per slice.  A wave
// waits on a chain of dependent loads (slot state, body ids, V), so waves in
// flight matter more than registers: the FP32 kernel is held to 168 VGPRs
// (3 waves/SIMD, no spills), 7% faster a sweep than at 2 waves/SIMD; the
// FP64 one to 256 (2 waves/SIMD, ~200 B a lane spilled), 9% faster than at
// 1 wave/SIMD with no spills.
template <typename T, bool FRAMES>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 2 : 3) void k_sweep_vel(SweepPtrs<T> P,
                                                                          BodyPtrs<T> B,
                                                      const DevParams* __restrict__ prm,
                                                      const SolveState* __restrict__ st,
                                                      int s_begin, int s_end, int want_phase) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = __builtin_amdgcn_readfirstlane(s_begin + xcd_block() * 4 + wave);
  if (s >= s_end) return;
  vel_slice<T, true, FRAMES>(P, B, prm, st, s, lane, want_phase);
}

// The last colours of a sweep, each of at most two rounds of the block's
// waves, in one workgroup: colour after colour, separated by workgroup
// barriers, in place of one launch per colour.  The same per-slice
// computation and partial sums as k_sweep_vel, so the result is identical.
template <typename T>
constexpr int vel_tail_threads() {
  return sizeof(T) == 8 ? 256 : 512;  // the waves k_sweep_vel's registers allow per CU
}

template <typename T, bool FRAMES>
//
// With finalize, the same workgroup then ends the sweep as k_finalize_sweep
// does (the sums of all nslices slice partials, with the same bits, and the
// state update), which saves that launch.
__global__ __launch_bounds__(vel_tail_threads<T>()) void k_sweep_vel_tail(
    SweepPtrs<T> P, BodyPtrs<T> B, const DevParams* __restrict__ prm, SolveState* st,
    const int* __restrict__ colour_slices, int c_begin, int c_end, int want_phase,
    int finalize, int nslices) {
  __shared__ double sh[2 * 16];
  if (st->stop || st->phase != want_phase) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int waves = vel_tail_threads<T>() / 64;
  for (int c = c_begin; c < c_end; c++) {
    const int sb = colour_slices[c], se = colour_slices[c + 1];
    for (int s = sb + wave; s < se; s += waves)
      vel_slice<T, false, FRAMES>(P, B, prm, st, s, lane, want_phase);
    __syncthreads();  // the next colour reads this colour's body velocities
  }
  if (!finalize) return;
  double sums[2];
  block_sums<2, vel_tail_threads<T>()>(P.part, nslices, sh, sums);
  if (threadIdx.x == 0) end_sweep_state(prm, st, sums[0], sums[1], want_phase);
}

}  // namespace fc3d
