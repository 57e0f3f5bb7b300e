// fc2d_problem.hip — 2-D friction-contact NSGS on gfx950 (fc2d_hip.h).
//
// A solve of fc2d_nsgs (fc2d_nsgs.c:297-538) sweeps the colours of the
// contact graph one after the other, in one of two schedules:
//   one workgroup: one launch of 1024 threads runs the whole solve, colours
//     separated by workgroup barriers, each thread taking the colour's
//     contacts in strides; the light error, the LIGHT_WITH_FULL_FINAL
//     tolerance logic and fc2d_compute_error (fc2d_compute_error.c:104-148)
//     run in the same kernel, so a solve has no host round trip;
//   per colour: one thread per contact over the whole GPU, one launch per
//     colour, then a one-workgroup launch for the light error and the
//     tolerance logic; the sweeps are captured in a graph and replayed
//     until a device-side stop flag is set.
// Both run the same per-contact step and the same reductions, so they give
// the same iterates and decisions bit for bit.
//
// Data (colour-major slot order, one slot per contact):
//   rows: 2 scalar rows per slot in CSR form, entries in the reference's
//         summation order; column = slot-space scalar index 2*slot(d)+comp;
//   diag: 2x2 diagonal block (column-major), det its determinant;
//   q, mu, r, u: per slot.
// Reductions (light error, norms, residual) sum each thread's slots in
// order, then the threads in a fixed order: deterministic run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "fc2d_hip.h"
#include "fc3d_internal.hpp"

namespace fc3d {
int colour_on_device(int nc, const int* d_rp, const int* d_ci, int* d_colour, hipStream_t s,
                     int* ncolours_out, std::vector<int>* host_colour);
}

namespace {

using fc3d::set_error;

constexpr int FC2D_THREADS = 1024;

struct Fc2dArgs {
  int nc;
  int ncolours;
  const int* cptr;     // colour -> slot range
  const int* rp;       // 2 nc + 1
  const int* rc;       // entry -> slot-space scalar column
  const double* rv;    // entry values
  const double* diag;  // 4 per slot
  const double* det;
  const double* q;     // 2 per slot
  const double* mu;
  double* r;
  double* u;
  double* light;       // per slot, scratch
  int row_mode;
  int own_zero;
  double norm_q;
};

struct Fc2dOut {
  int iterations;
  int notconv;
  int velocity_set;
  int pad;
  double error;
};

// fc2d_nsgs.c:247-284
__device__ __forceinline__ void local_solve(const double* W, double D, const double* q, double mu,
                                            double* P) {
  const double Wnn = W[0], Wtn = W[1], Wnt = W[2], Wtt = W[3];
  if (q[0] > 0) {
    P[0] = 0;
    P[1] = 0;
  } else {
    P[0] = -(Wtt * q[0] - Wnt * q[1]) / D;
    P[1] = -(-Wtn * q[0] + Wnn * q[1]) / D;
    const double muPn = mu * P[0];
    if (fabs(P[1]) > muPn) {
      if (P[1] + muPn < 0) {
        P[0] = -q[0] / (Wnn - mu * Wnt);
        P[1] = -mu * P[0];
      } else {
        P[0] = -q[0] / (Wnn + mu * Wnt);
        P[1] = mu * P[0];
      }
    }
  }
}

// fc2d_compute_error.c:37-56
__device__ __forceinline__ double unitary_error(const double* r, const double* u, double mu) {
  double w0 = r[0] - u[0] - mu * fabs(u[1]);
  double w1 = r[1] - u[1];
  const double normT = fabs(w1);
  if (mu * normT <= -w0) {
    w0 = 0.;
    w1 = 0.;
  } else if (normT > mu * w0) {
    w0 = (mu * normT + w0) / (mu * mu + 1.0);
    w1 = mu * w0 * (w1 < 0 ? -1 : w1 > 0 ? 1 : 0);
  }
  w0 = r[0] - w0;
  w1 = r[1] - w1;
  return w0 * w0 + w1 * w1;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over the workgroup: waves by shuffles, then the 16 wave sums in order.
__device__ double block_sum(double v, double* sh) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < FC2D_THREADS / 64; w++) t += sh[w];
  return t;
}

// acc += the row's terms in order; own-contact entries (column in the pair
// of slot s) add w * 0 when own_zero, else nothing.  x: the reactions read
// (r for the sweep's local problem; own entries never read r there).
// Entries are loaded CHUNK at a time (columns and values, then the reactions
// they select) before being added in order, so the loads of a chunk overlap
// instead of forming one dependent chain per entry.
template <bool RESIDUAL, int CHUNK = 8>
__device__ __forceinline__ double row_terms(const Fc2dArgs& A, int s, int beg, int end,
                                            double acc) {
  for (int e0 = beg; e0 < end; e0 += CHUNK) {
    int col[CHUNK];
    double w[CHUNK], x[CHUNK];
#pragma unroll
    for (int k = 0; k < CHUNK; k++)
      if (e0 + k < end) {
        col[k] = A.rc[e0 + k];
        w[k] = A.rv[e0 + k];
      }
#pragma unroll
    for (int k = 0; k < CHUNK; k++)
      if (e0 + k < end) x[k] = (RESIDUAL || (col[k] >> 1) != s) ? A.r[col[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < CHUNK; k++)
      if (e0 + k < end) {
        if (RESIDUAL || (col[k] >> 1) != s)
          acc += w[k] * x[k];
        else if (A.own_zero)
          acc += w[k] * 0.0;
      }
  }
  return acc;
}

// q_loc of slot s (fc2d_nsgs.c:97-113 through NM_row_prod_no_diag2), row by
// row, C entries in flight at a time.  (Reading both rows' chunks together
// measured slower: 83 against 61 ms a solve of 7002 contacts.)
// rp: the slot's row pointers rp[2 s], rp[2 s + 1], rp[2 s + 2].
template <int C>
__device__ __forceinline__ void local_q(const Fc2dArgs& A, int s, const int* rp, double* y) {
  for (int i = 0; i < 2; i++) {
    const int row = 2 * s + i;
    if (A.row_mode == FC2D_ROWS_DOT)
      y[i] = A.q[row] + row_terms<false, C>(A, s, rp[i], rp[i + 1], 0.0);
    else
      y[i] = row_terms<false, C>(A, s, rp[i], rp[i + 1], A.q[row]);
  }
}

// The per-contact step of a sweep (fc2d_nsgs.c:448-471): local problem,
// local solve, light-error term, write-back.
template <int C>
__device__ __forceinline__ void sweep_contact(const Fc2dArgs& A, int s, const int* rp) {
  double ql[2], P[2] = {A.r[2 * s], A.r[2 * s + 1]};
  local_q<C>(A, s, rp, ql);
  local_solve(A.diag + 4 * s, A.det[s], ql, A.mu[s], P);
  const double x0 = A.r[2 * s] - P[0], x1 = A.r[2 * s + 1] - P[1];
  A.light[s] = x0 * x0 + x1 * x1;
  A.r[2 * s] = P[0];
  A.r[2 * s + 1] = P[1];
}

// fc2d_compute_error with the resident r; writes u.  All threads return it.
__device__ double full_error(const Fc2dArgs& A, double* sh) {
  double e = 0.0, r2 = 0.0, u2 = 0.0;
  for (int s = threadIdx.x; s < A.nc; s += FC2D_THREADS) {
    double u[2];
    for (int i = 0; i < 2; i++) {
      const int row = 2 * s + i;
      u[i] = row_terms<true>(A, s, A.rp[row], A.rp[row + 1], A.q[row]);
      A.u[row] = u[i];
    }
    const double rr[2] = {A.r[2 * s], A.r[2 * s + 1]};
    e += unitary_error(rr, u, A.mu[s]);
    r2 += rr[0] * rr[0] + rr[1] * rr[1];
    u2 += u[0] * u[0] + u[1] * u[1];
  }
  e = block_sum(e, sh);
  r2 = block_sum(r2, sh);
  u2 = block_sum(u2, sh);
  double err = sqrt(e);
  const double scale = fmax(sqrt(r2), fmax(sqrt(u2), A.norm_q));
  if (fabs(scale) > DBL_EPSILON) err /= scale;
  return err;
}

__global__ __launch_bounds__(FC2D_THREADS) void k_fc2d_error(Fc2dArgs A, double* out) {
  __shared__ double sh[FC2D_THREADS / 64];
  const double e = full_error(A, sh);
  if (threadIdx.x == 0) *out = e;
}

// The light error and the tolerance logic after a sweep (fc2d_nsgs.c:136-142,
// 191-228), for the light modes.  All threads of the workgroup call it and
// get the same results.
__device__ void end_of_sweep(const Fc2dArgs& A, double* sh, int mode, double tol,
                             double* tolerance, double* error, int* notconv, int* vel) {
  double l = 0.0, r2 = 0.0;
  for (int s = threadIdx.x; s < A.nc; s += FC2D_THREADS) {
    l += A.light[s];
    r2 += A.r[2 * s] * A.r[2 * s] + A.r[2 * s + 1] * A.r[2 * s + 1];
  }
  l = block_sum(l, sh);
  r2 = block_sum(r2, sh);
  double err = sqrt(l);
  const double norm_r = sqrt(r2);
  if (fabs(norm_r) > DBL_EPSILON) err /= norm_r;
  *error = err;
  if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT) {
    *notconv = err < *tolerance ? 0 : 1;
  } else {
    *notconv = 1;
    if (err < *tolerance) {
      *notconv = 0;
      const double abs_err = full_error(A, sh);
      *vel = 1;
      if (abs_err > tol) {
        *tolerance = err / abs_err * tol;
        *notconv = 1;
      }
    }
  }
}

__device__ __forceinline__ bool is_light(int mode) {
  return mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT ||
         mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL;
}

// fc2d_nsgs.c:448-538 (the branch without freezing) in colour order
__global__ __launch_bounds__(FC2D_THREADS) void k_fc2d_nsgs(Fc2dArgs A, int itermax, double tol,
                                                           int mode, Fc2dOut* out) {
  __shared__ double sh[FC2D_THREADS / 64];
  int iter = 0, notconv = 1, vel = 0;
  double error = INFINITY, tolerance = tol;
  const bool light_mode = is_light(mode);
  while (iter < itermax && notconv) {
    ++iter;
    for (int c = 0; c < A.ncolours; c++) {
      for (int s = A.cptr[c] + threadIdx.x; s < A.cptr[c + 1]; s += FC2D_THREADS) {
        const int rp[3] = {A.rp[2 * s], A.rp[2 * s + 1], A.rp[2 * s + 2]};
        sweep_contact<8>(A, s, rp);
      }
      __syncthreads();  // the next colour reads this colour's reactions
    }
    if (light_mode) end_of_sweep(A, sh, mode, tol, &tolerance, &error, &notconv, &vel);
    __syncthreads();
  }
  if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    error = full_error(A, sh);
    vel = 1;
    notconv = error < tol ? 0 : 1;
  }
  if (threadIdx.x == 0) {
    out->iterations = iter;
    out->notconv = notconv;
    out->velocity_set = vel;
    out->error = error;
  }
}

// Solve state of the per-colour schedule, in device memory: the loop
// variables of k_fc2d_nsgs and the solve's parameters, so that one captured
// graph serves every solve of the problem.
struct Fc2dState {
  int iter;
  int notconv;
  int vel;
  int stop;  // no further sweep: converged or iter == itermax
  int itermax;
  int mode;
  double error;
  double tolerance;
  double tol;
};

// One colour of one sweep: one thread per contact of slots [sb, se).
__global__ __launch_bounds__(256) void k_fc2d_colour(Fc2dArgs A, int sb, int se,
                                                     const Fc2dState* st) {
  // the contact's row pointers are loaded before the stop flag is tested, so
  // the two latencies overlap; the empty asm keeps the loads above the test
  const int s = sb + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int t = s < se ? s : sb;
  const int rp[3] = {A.rp[2 * t], A.rp[2 * t + 1], A.rp[2 * t + 2]};
  const int stop = st->stop;
  asm volatile("" ::"v"(rp[0]), "v"(rp[1]), "v"(rp[2]), "s"(stop));
  if (stop) return;
  if (s < se) sweep_contact<8>(A, s, rp);
}

// End of a sweep of the per-colour schedule (the loop test and body tail of
// k_fc2d_nsgs).
__global__ __launch_bounds__(FC2D_THREADS) void k_fc2d_end_sweep(Fc2dArgs A, Fc2dState* st) {
  __shared__ double sh[FC2D_THREADS / 64];
  if (st->stop) return;
  const int iter = st->iter + 1, mode = st->mode;
  int notconv = 1, vel = st->vel;
  double error = st->error, tolerance = st->tolerance;
  if (is_light(mode)) end_of_sweep(A, sh, mode, st->tol, &tolerance, &error, &notconv, &vel);
  __syncthreads();  // every thread has read st
  if (threadIdx.x == 0) {
    st->iter = iter;
    st->notconv = notconv;
    st->vel = vel;
    st->error = error;
    st->tolerance = tolerance;
    st->stop = iter < st->itermax && notconv ? 0 : 1;
  }
}

// After the last sweep of the per-colour schedule: LIGHT_WITH_FULL_FINAL's
// final fc2d_compute_error (fc2d_nsgs.c:525-532), and the results.
__global__ __launch_bounds__(FC2D_THREADS) void k_fc2d_finish(Fc2dArgs A, const Fc2dState* st,
                                                             Fc2dOut* out) {
  __shared__ double sh[FC2D_THREADS / 64];
  int notconv = st->notconv, vel = st->vel;
  double error = st->error;
  if (st->mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
    error = full_error(A, sh);
    vel = 1;
    notconv = error < st->tol ? 0 : 1;
  }
  if (threadIdx.x == 0) {
    out->iterations = st->iter;
    out->notconv = notconv;
    out->velocity_set = vel;
    out->error = error;
  }
}

template <typename T>
int dev_upload(T** p, const std::vector<T>& h) {
  FC3D_HIP_TRY(hipMalloc(p, sizeof(T) * (h.empty() ? 1 : h.size())));
  if (!h.empty()) FC3D_HIP_TRY(hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

struct fc2d_hip_problem {
  int device = 0;
  hipStream_t stream = nullptr;
  int nc = 0;
  int ncolours = 0;
  double norm_q = 0.0;
  int row_mode = 0, own_zero = 0;
  std::vector<int> order;  // slot -> contact
  std::vector<int> slot;   // contact -> slot
  int *cptr = nullptr, *rp = nullptr, *rc = nullptr;
  double *rv = nullptr, *diag = nullptr, *det = nullptr, *q = nullptr, *mu = nullptr;
  double *r = nullptr, *u = nullptr, *light = nullptr, *err = nullptr;
  Fc2dOut* out = nullptr;
  Fc2dState* st = nullptr;        // per-colour schedule
  hipGraphExec_t graph = nullptr;  // FC2D_GRAPH_SWEEPS sweeps of it
  std::vector<int> hcptr;          // colour -> slot range, on the host
  hipEvent_t ev[2] = {nullptr, nullptr};

  Fc2dArgs args() const {
    Fc2dArgs a;
    a.nc = nc;
    a.ncolours = ncolours;
    a.cptr = cptr;
    a.rp = rp;
    a.rc = rc;
    a.rv = rv;
    a.diag = diag;
    a.det = det;
    a.q = q;
    a.mu = mu;
    a.r = r;
    a.u = u;
    a.light = light;
    a.row_mode = row_mode;
    a.own_zero = own_zero;
    a.norm_q = norm_q;
    return a;
  }
};

// Sweeps per graph replay of the per-colour schedule: the host reads the stop
// flag after each replay; sweeps after the stop return at once.
constexpr int FC2D_GRAPH_SWEEPS = 16;

// The schedule a solve uses: the requested one, or automatically one launch
// per colour when some colour holds more contacts than one workgroup sweeps
// in two strides (below that, one workgroup's barriers cost less than the
// launches).  -1 for an unknown request.
static int fc2d_schedule(const fc2d_hip_problem* h, int requested) {
  if (requested == FC2D_SCHEDULE_ONE_WORKGROUP || requested == FC2D_SCHEDULE_PER_COLOUR)
    return requested;
  if (requested != FC2D_SCHEDULE_AUTO) return -1;
  int widest = 0;
  for (size_t c = 0; c + 1 < h->hcptr.size(); c++)
    widest = std::max(widest, h->hcptr[c + 1] - h->hcptr[c]);
  return widest > 2 * FC2D_THREADS ? FC2D_SCHEDULE_PER_COLOUR : FC2D_SCHEDULE_ONE_WORKGROUP;
}

// Captures FC2D_GRAPH_SWEEPS sweeps of the per-colour schedule once per
// problem; the solve's parameters live in the state, not in the graph.
static int build_graph(fc2d_hip_problem* h) {
  if (h->graph) return 0;
  const Fc2dArgs a = h->args();
  hipGraph_t g = nullptr;
  FC3D_HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < FC2D_GRAPH_SWEEPS; k++) {
    for (int c = 0; c < h->ncolours; c++) {
      const int sb = h->hcptr[(size_t)c], se = h->hcptr[(size_t)c + 1];
      if (se <= sb) continue;
      hipLaunchKernelGGL(k_fc2d_colour, dim3((unsigned)((se - sb + 255) / 256)), dim3(256), 0,
                         h->stream, a, sb, se, h->st);
    }
    hipLaunchKernelGGL(k_fc2d_end_sweep, dim3(1), dim3(FC2D_THREADS), 0, h->stream, a, h->st);
  }
  const hipError_t e1 = hipStreamEndCapture(h->stream, &g);
  if (e1 != hipSuccess) {
    set_error("fc2d graph capture: %s", hipGetErrorString(e1));
    return NUMERICS_ERR_DEVICE;
  }
  const hipError_t e2 = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    h->graph = nullptr;
    set_error("fc2d graph instantiation: %s", hipGetErrorString(e2));
    return NUMERICS_ERR_DEVICE;
  }
  return 0;
}

extern "C" {

void fc2d_hip_params_default(fc2d_hip_params* p) {
  if (!p) return;
  p->max_iter = 1000;
  p->tol = 1e-4;
  p->error_evaluation = SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL;
  p->schedule = FC2D_SCHEDULE_AUTO;
}

void fc2d_hip_problem_destroy(fc2d_hip_problem* h) {
  if (!h) return;
  if (fc3d::use_device(h->device) == 0) {
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->graph) hipGraphExecDestroy(h->graph);
    void* ptrs[] = {h->cptr, h->rp, h->rc,    h->rv,  h->diag, h->det, h->q,
                    h->mu,   h->r,  h->u,     h->light, h->err, h->out, h->st};
    for (void* p : ptrs)
      if (p) hipFree(p);
    for (hipEvent_t e : h->ev)
      if (e) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
  }
  delete h;
}

int fc2d_hip_problem_create(int nc, const int64_t* row_ptr, const int64_t* cols,
                            const double* vals, int row_mode, int own_zero, const double* diag,
                            const double* q, const double* mu, int device,
                            fc2d_hip_problem** out) {
  if (!out) return NUMERICS_ERR_NULL_POINTER;
  *out = nullptr;
  if (nc <= 0 || !row_ptr || !diag || !q || !mu ||
      (row_mode != FC2D_ROWS_ACCUMULATE && row_mode != FC2D_ROWS_DOT)) {
    set_error("fc2d_hip_problem_create: invalid arguments (nc=%d)", nc);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  const int64_t n2 = 2 * (int64_t)nc, nnz = row_ptr[n2];
  if (row_ptr[0] != 0 || nnz < 0 || nnz > INT32_MAX || (nnz > 0 && (!cols || !vals))) {
    set_error("fc2d_hip_problem_create: bad row pointers");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  for (int64_t i = 0; i < n2; i++)
    if (row_ptr[i + 1] < row_ptr[i]) {
      set_error("fc2d_hip_problem_create: row pointers decrease at row %lld", (long long)i);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
  for (int64_t e = 0; e < nnz; e++)
    if (cols[e] < 0 || cols[e] >= n2) {
      set_error("fc2d_hip_problem_create: column %lld out of range", (long long)cols[e]);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
  int rc = fc3d::use_device(device);
  if (rc) return rc;
  fc2d_hip_problem* h = new fc2d_hip_problem();
  h->device = device;
  h->nc = nc;
  h->row_mode = row_mode;
  h->own_zero = own_zero ? 1 : 0;
  double nq = 0.0;
  for (int64_t i = 0; i < n2; i++) nq += q[i] * q[i];
  h->norm_q = sqrt(nq);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev[0]) != hipSuccess || hipEventCreate(&h->ev[1]) != hipSuccess) {
    set_error("fc2d_hip_problem_create: stream or event creation failed");
    fc2d_hip_problem_destroy(h);
    return NUMERICS_ERR_DEVICE;
  }
  // conflict graph: contacts referenced by a row, symmetrised, without self
  std::vector<std::vector<int>> nb((size_t)nc);
  for (int c = 0; c < nc; c++)
    for (int i = 0; i < 2; i++)
      for (int64_t e = row_ptr[2 * c + i]; e < row_ptr[2 * c + i + 1]; e++) {
        const int d = (int)(cols[e] >> 1);
        if (d != c) {
          nb[(size_t)c].push_back(d);
          nb[(size_t)d].push_back(c);
        }
      }
  std::vector<int> grp((size_t)nc + 1, 0), gci;
  for (int c = 0; c < nc; c++) {
    auto& v = nb[(size_t)c];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    grp[(size_t)c + 1] = grp[(size_t)c] + (int)v.size();
  }
  gci.reserve((size_t)grp[(size_t)nc]);
  for (int c = 0; c < nc; c++) gci.insert(gci.end(), nb[(size_t)c].begin(), nb[(size_t)c].end());
  nb.clear();
  int *d_rp = nullptr, *d_ci = nullptr, *d_col = nullptr;
  std::vector<int> colour;
  if ((rc = dev_upload(&d_rp, grp)) || (rc = dev_upload(&d_ci, gci)) ||
      hipMalloc(&d_col, sizeof(int) * (size_t)nc) != hipSuccess) {
    if (!rc) set_error("fc2d_hip_problem_create: allocation failed");
    hipFree(d_rp);
    hipFree(d_ci);
    fc2d_hip_problem_destroy(h);
    return rc ? rc : NUMERICS_ERR_DEVICE;
  }
  rc = fc3d::colour_on_device(nc, d_rp, d_ci, d_col, h->stream, &h->ncolours, &colour);
  hipFree(d_rp);
  hipFree(d_ci);
  hipFree(d_col);
  if (rc) {
    fc2d_hip_problem_destroy(h);
    return rc;
  }
  // colour-major order, index order inside a colour
  std::vector<int> cptr((size_t)h->ncolours + 1, 0);
  for (int c = 0; c < nc; c++) cptr[(size_t)colour[(size_t)c] + 1]++;
  for (int k = 0; k < h->ncolours; k++) cptr[(size_t)k + 1] += cptr[(size_t)k];
  h->order.assign((size_t)nc, 0);
  h->slot.assign((size_t)nc, 0);
  {
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    for (int c = 0; c < nc; c++) {
      const int s = fill[(size_t)colour[(size_t)c]]++;
      h->order[(size_t)s] = c;
      h->slot[(size_t)c] = s;
    }
  }
  // rows, diagonal blocks, q, mu in slot order
  std::vector<int> srp((size_t)n2 + 1, 0), src((size_t)(nnz ? nnz : 0));
  std::vector<double> srv((size_t)(nnz ? nnz : 0)), sdiag(4 * (size_t)nc), sdet((size_t)nc),
      sq((size_t)n2), smu((size_t)nc);
  for (int s = 0; s < nc; s++) {
    const int c = h->order[(size_t)s];
    for (int i = 0; i < 2; i++) {
      const int64_t row = 2 * (int64_t)c + i;
      srp[2 * (size_t)s + i + 1] = (int)(row_ptr[row + 1] - row_ptr[row]);
    }
  }
  for (int64_t i = 0; i < n2; i++) srp[(size_t)i + 1] += srp[(size_t)i];
  for (int s = 0; s < nc; s++) {
    const int c = h->order[(size_t)s];
    for (int i = 0; i < 2; i++) {
      const int64_t row = 2 * (int64_t)c + i;
      int k = srp[2 * (size_t)s + i];
      for (int64_t e = row_ptr[row]; e < row_ptr[row + 1]; e++, k++) {
        const int d = (int)(cols[e] >> 1), comp = (int)(cols[e] & 1);
        src[(size_t)k] = 2 * h->slot[(size_t)d] + comp;
        srv[(size_t)k] = vals[e];
      }
      sq[2 * (size_t)s + i] = q[row];
    }
    for (int e = 0; e < 4; e++) sdiag[4 * (size_t)s + e] = diag[4 * (size_t)c + e];
    // fc2d_nsgs.c:229-233
    const double* B = diag + 4 * (size_t)c;
    sdet[(size_t)s] = B[0] * B[3] - B[1] * B[2];
    smu[(size_t)s] = mu[c];
  }
  std::vector<double> zero((size_t)n2, 0.0);
  if ((rc = dev_upload(&h->cptr, cptr)) || (rc = dev_upload(&h->rp, srp)) ||
      (rc = dev_upload(&h->rc, src)) || (rc = dev_upload(&h->rv, srv)) ||
      (rc = dev_upload(&h->diag, sdiag)) || (rc = dev_upload(&h->det, sdet)) ||
      (rc = dev_upload(&h->q, sq)) || (rc = dev_upload(&h->mu, smu)) ||
      (rc = dev_upload(&h->r, zero)) || (rc = dev_upload(&h->u, zero)) ||
      (rc = dev_upload(&h->light, std::vector<double>((size_t)nc, 0.0)))) {
    fc2d_hip_problem_destroy(h);
    return rc;
  }
  h->hcptr = cptr;
  if (hipMalloc(&h->err, sizeof(double)) != hipSuccess ||
      hipMalloc(&h->out, sizeof(Fc2dOut)) != hipSuccess ||
      hipMalloc(&h->st, sizeof(Fc2dState)) != hipSuccess) {
    set_error("fc2d_hip_problem_create: allocation failed");
    fc2d_hip_problem_destroy(h);
    return NUMERICS_ERR_DEVICE;
  }
  *out = h;
  return 0;
}

int fc2d_hip_problem_num_colours(const fc2d_hip_problem* h) { return h ? h->ncolours : -1; }

int fc2d_hip_problem_order(const fc2d_hip_problem* h, int* order) {
  if (!h || !order) return NUMERICS_ERR_NULL_POINTER;
  std::memcpy(order, h->order.data(), sizeof(int) * (size_t)h->nc);
  return 0;
}

static int copy_in(fc2d_hip_problem* h, const double* host, double* dev) {
  std::vector<double> s(2 * (size_t)h->nc);
  for (int k = 0; k < h->nc; k++) {
    const int c = h->order[(size_t)k];
    s[2 * (size_t)k] = host[2 * (size_t)c];
    s[2 * (size_t)k + 1] = host[2 * (size_t)c + 1];
  }
  FC3D_HIP_TRY(hipMemcpyAsync(dev, s.data(), sizeof(double) * s.size(), hipMemcpyHostToDevice,
                              h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  return 0;
}

static int copy_out(fc2d_hip_problem* h, const double* dev, double* host) {
  std::vector<double> s(2 * (size_t)h->nc);
  FC3D_HIP_TRY(hipMemcpyAsync(s.data(), dev, sizeof(double) * s.size(), hipMemcpyDeviceToHost,
                              h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  for (int k = 0; k < h->nc; k++) {
    const int c = h->order[(size_t)k];
    host[2 * (size_t)c] = s[2 * (size_t)k];
    host[2 * (size_t)c + 1] = s[2 * (size_t)k + 1];
  }
  return 0;
}

int fc2d_hip_set_reaction(fc2d_hip_problem* h, const double* r) {
  if (!h || !r) return NUMERICS_ERR_NULL_POINTER;
  int rc = fc3d::use_device(h->device);
  return rc ? rc : copy_in(h, r, h->r);
}

int fc2d_hip_get_reaction(fc2d_hip_problem* h, double* r) {
  if (!h || !r) return NUMERICS_ERR_NULL_POINTER;
  int rc = fc3d::use_device(h->device);
  return rc ? rc : copy_out(h, h->r, r);
}

int fc2d_hip_get_velocity(fc2d_hip_problem* h, double* u) {
  if (!h || !u) return NUMERICS_ERR_NULL_POINTER;
  int rc = fc3d::use_device(h->device);
  return rc ? rc : copy_out(h, h->u, u);
}

int fc2d_hip_solve(fc2d_hip_problem* h, const fc2d_hip_params* p, fc2d_hip_stats* s) {
  if (!h || !p || !s) return NUMERICS_ERR_NULL_POINTER;
  if (p->max_iter < 0 || !(p->tol >= 0.0)) {
    set_error("fc2d_hip_solve: max_iter %d / tol %g out of range", p->max_iter, p->tol);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  int rc = fc3d::use_device(h->device);
  if (rc) return rc;
  std::memset(s, 0, sizeof *s);
  s->num_colours = h->ncolours;
  // fc2d_nsgs.c:228-245, 318-334: a diagonal block with det < DBL_EPSILON
  // stops the solve before any sweep
  {
    std::vector<double> det((size_t)h->nc);
    FC3D_HIP_TRY(hipMemcpy(det.data(), h->det, sizeof(double) * det.size(),
                           hipMemcpyDeviceToHost));
    for (double d : det)
      if (d < DBL_EPSILON) {
        s->info = 1;
        s->iterations = 0;
        s->residual = INFINITY;
        return 0;
      }
  }
  const int schedule = fc2d_schedule(h, p->schedule);
  if (schedule < 0) {
    set_error("fc2d_hip_solve: schedule %d is not 0, 1 or 2", p->schedule);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  s->schedule = schedule;
  if (schedule == FC2D_SCHEDULE_PER_COLOUR) {
    Fc2dState init;
    std::memset(&init, 0, sizeof init);
    init.notconv = 1;
    init.stop = p->max_iter > 0 ? 0 : 1;
    init.itermax = p->max_iter;
    init.mode = p->error_evaluation;
    init.error = INFINITY;
    init.tolerance = p->tol;
    init.tol = p->tol;
    FC3D_HIP_TRY(hipMemcpyAsync(h->st, &init, sizeof init, hipMemcpyHostToDevice, h->stream));
    if ((rc = build_graph(h))) return rc;
    FC3D_HIP_TRY(hipEventRecord(h->ev[0], h->stream));
    // every replay advances the sweep counter or finds the stop flag set
    const long max_replays = (long)p->max_iter / FC2D_GRAPH_SWEEPS + 2;
    for (long replay = 0;; replay++) {
      if (replay > max_replays) {
        set_error("fc2d_hip_solve: sweep loop did not stop after %ld replays", replay);
        return NUMERICS_ERR_UNKNOWN;
      }
      if (init.stop) break;
      FC3D_HIP_TRY(hipGraphLaunch(h->graph, h->stream));
      FC3D_HIP_TRY(hipMemcpyAsync(&init, h->st, sizeof init, hipMemcpyDeviceToHost, h->stream));
      FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
    }
    hipLaunchKernelGGL(k_fc2d_finish, dim3(1), dim3(FC2D_THREADS), 0, h->stream, h->args(), h->st,
                       h->out);
  } else {
    FC3D_HIP_TRY(hipEventRecord(h->ev[0], h->stream));
    hipLaunchKernelGGL(k_fc2d_nsgs, dim3(1), dim3(FC2D_THREADS), 0, h->stream, h->args(),
                       p->max_iter, p->tol, p->error_evaluation, h->out);
  }
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipEventRecord(h->ev[1], h->stream));
  Fc2dOut o;
  FC3D_HIP_TRY(hipMemcpyAsync(&o, h->out, sizeof o, hipMemcpyDeviceToHost, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  float ms = 0.f;
  hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
  s->info = o.notconv;
  s->iterations = o.iterations;
  s->residual = o.error;
  s->velocity_set = o.velocity_set;
  s->solve_ms = ms;
  return 0;
}

int fc2d_hip_compute_error(fc2d_hip_problem* h, const double* reaction, double* velocity,
                           double norm, double* error) {
  if (!h || !error) return NUMERICS_ERR_NULL_POINTER;
  int rc = fc3d::use_device(h->device);
  if (rc) return rc;
  if (reaction && (rc = copy_in(h, reaction, h->r))) return rc;
  Fc2dArgs a = h->args();
  a.norm_q = norm;
  hipLaunchKernelGGL(k_fc2d_error, dim3(1), dim3(FC2D_THREADS), 0, h->stream, a, h->err);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipMemcpyAsync(error, h->err, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  FC3D_HIP_TRY(hipStreamSynchronize(h->stream));
  if (velocity && (rc = copy_out(h, h->u, velocity))) return rc;
  return 0;
}

}  // extern "C"
