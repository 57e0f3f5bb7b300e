// fc3d_batch.hip — many independent fc3d problems solved in one persistent
// launch.
//
// All problems share one colour-major sliced layout (each problem occupies
// whole slices).  A workgroup takes the next problem from an atomic queue
// (problems ordered by decreasing size) and runs that problem's complete
// coloured NSGS: colour passes separated by workgroup barriers, light error
// and FP64 residual reduced inside the workgroup, the reference convergence
// logic (fc3d_nsgs.c:826-895 with the tolerance manager of
// tolerance_manager.h:203-225) per problem.  No global synchronisation is
// needed.  In mixed precision a problem sweeps in FP32 until its FP32 light
// error is below fp32_tol, then continues in FP64 from that iterate.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "fc3d_internal.hpp"
#include "fc3d_layout.hpp"
#include "fc3d_sweep.hpp"

namespace fc3d {

int colour_on_device(int nc, const int* d_rp, const int* d_ci, int* d_colour, hipStream_t s,
                     int* ncolours_out, std::vector<int>* host_colour);

struct BatchProblem {
  int cptr_off;     // first entry of the problem's colour slice pointers
  int ncolours;
  int slice_begin;
  int slice_end;
  double norm_q;
};

struct BatchArgs {
  SweepPtrs<double> P64;
  SweepPtrs<float> P32;
  const int* cptr;
  const BatchProblem* problems;
  const int* order;
  int nprob;
  int* counter;
  int* out_info;
  int* out_iter;
  int* out_fp32_iter;
  double* out_residual;
};

template <int WPP>
__device__ __forceinline__ double block_reduce(double v, double* sh) {
  v = wave_sum(v);
  if (WPP == 1) return v;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < WPP; w++) t += sh[w];
  return t;
}

// FP64 residual of one problem (fc3d_compute_error.c:70-122).
template <int WPP>
__device__ double problem_residual(const BatchArgs& A, const BatchProblem& d, double* sh) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double e2 = 0.0, r2 = 0.0, u2 = 0.0;
  for (int s = d.slice_begin + wave; s < d.slice_end; s += WPP) {
    double a, b, c;
    residual_slot(A.P64, s, lane, &a, &b, &c);
    e2 += a;
    r2 += b;
    u2 += c;
  }
  e2 = block_reduce<WPP>(e2, sh);
  r2 = block_reduce<WPP>(r2, sh);
  u2 = block_reduce<WPP>(u2, sh);
  double err = sqrt(e2);
  const double scale = fmax(d.norm_q, fmax(sqrt(r2), sqrt(u2)));
  if (fabs(scale) > DBL_EPSILON) err /= scale;
  return err;
}

template <int WPP>
__global__ __launch_bounds__(64 * WPP, 2) void k_batch(BatchArgs A, const DevParams* __restrict__ prm,
                                                    int mixed) {
  __shared__ int sh_task;
  __shared__ double sh[WPP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (;;) {
    if (threadIdx.x == 0) sh_task = atomicAdd(A.counter, 1);
    __syncthreads();
    const int t = sh_task;
    __syncthreads();
    if (t >= A.nprob) return;
    const int p = A.order[t];
    const BatchProblem d = A.problems[p];
    for (int slot = d.slice_begin * 64 + threadIdx.x; slot < d.slice_end * 64; slot += 64 * WPP) {
      A.P64.pli[slot] = 1.0;
      if (mixed) {
        const double4 x = A.P64.r[slot];
        A.P32.r[slot] = make_float4(float(x.x), float(x.y), float(x.z), 0.f);
        A.P32.pli[slot] = 1.0f;
      }
    }
    __syncthreads();
    int iter = 0, fp32_iter = 0, phase = mixed, notconv = 1;
    double working_tol = prm->tol, local_tol = prm->local_tol, incr = INFINITY, full = INFINITY;
    int freq = prm->freq;
    const int mode = prm->mode;
    while (iter < prm->max_iter) {
      double light = 0.0, nr2 = 0.0;
      for (int c = 0; c < d.ncolours; c++) {
        const int sb = A.cptr[d.cptr_off + c], se = A.cptr[d.cptr_off + c + 1];
        for (int s = sb + wave; s < se; s += WPP) {
          double l, n;
          if (phase)
            sweep_slot<float, 8, 0>(A.P32, prm, s, lane, float(prm->fp32_local_tol), &l, &n);
          else
            sweep_slot<double, 4, 0>(A.P64, prm, s, lane, local_tol, &l, &n);
          light += l;
          nr2 += n;
        }
        __syncthreads();  // the next colour reads this colour's reactions
      }
      light = block_reduce<WPP>(light, sh);
      nr2 = block_reduce<WPP>(nr2, sh);
      iter++;
      incr = sqrt(light);
      const double nr = sqrt(nr2);
      if (fabs(nr) > DBL_EPSILON) incr /= nr;
      if (phase) {
        fp32_iter++;
        if (incr < prm->fp32_tol || iter >= prm->max_iter) {
          for (int slot = d.slice_begin * 64 + threadIdx.x; slot < d.slice_end * 64;
               slot += 64 * WPP) {
            const float4 x = A.P32.r[slot];
            A.P64.r[slot] = make_double4(double(x.x), double(x.y), double(x.z), 0.0);
            A.P64.pli[slot] = double(A.P32.pli[slot]);
          }
          __syncthreads();
          phase = 0;
        }
        continue;
      }
      if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT) {
        notconv = incr < working_tol ? 0 : 1;
      } else if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
        if (incr >= working_tol) {
          notconv = 1;
        } else {
          full = problem_residual<WPP>(A, d, sh);
          __syncthreads();
          const double user = prm->tol;
          if (full <= user) {
            notconv = 0;
          } else if (incr < 1e-15 || incr < DBL_EPSILON * 100.0) {
            local_tol = fmax(local_tol / 100.0, DBL_EPSILON * 1e-6);
            notconv = 1;
          } else {
            if (full > 0.0 && incr > 0.0) {
              double nt = incr / full * user;
              const double umin = user * 1e-6, fmin_ = DBL_EPSILON * 1e-6;
              const double mn = umin > fmin_ ? umin : fmin_;
              if (nt < mn) nt = mn;
              if (nt > 0.0 && nt < working_tol) working_tol = nt;
            }
            notconv = 1;
          }
          if (!(working_tol > 0.0)) break;
        }
      } else {
        const bool eval = freq > 0 ? (iter % freq == 0) : true;
        if (eval) {
          full = problem_residual<WPP>(A, d, sh);
          __syncthreads();
          if (mode == SICONOS_NSGS_ERROR_EVALUATION_ADAPTIVE && freq > 0 && full > working_tol)
            freq *= 2;
        } else {
          full = 1e24;
        }
        incr = full;
        notconv = full < working_tol ? 0 : 1;
      }
      if (!notconv) break;
    }
    if (phase) {  // max_iter reached inside the FP32 phase: hand the iterate over
      for (int slot = d.slice_begin * 64 + threadIdx.x; slot < d.slice_end * 64;
           slot += 64 * WPP) {
        const float4 x = A.P32.r[slot];
        A.P64.r[slot] = make_double4(double(x.x), double(x.y), double(x.z), 0.0);
      }
      __syncthreads();
    }
    if (mode == SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL) {
      full = problem_residual<WPP>(A, d, sh);
      notconv = full < prm->tol ? 0 : 1;
    }
    if (threadIdx.x == 0) {
      int info = notconv;
      if (iter == prm->max_iter && notconv > 0) info = NUMERICS_ERR_MAX_ITER;
      A.out_info[p] = info;
      A.out_iter[p] = iter;
      A.out_fp32_iter[p] = fp32_iter;
      A.out_residual[p] = full;
    }
    __syncthreads();
  }
}

template <typename S, typename D>
__global__ void k_convert_b(const S* a, D* b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = D(a[i]);
}

}  // namespace fc3d

using namespace fc3d;

struct fc3d_hip_batch {
  int device = 0;
  int nprob = 0;
  int ntotal = 0;
  int nslices = 0, nslots = 0;
  int64_t nplanes = 0;
  hipStream_t stream = nullptr;
  std::vector<int> nc, coff, slot2orig, orig2slot, cptr, colour;
  std::vector<BatchProblem> problems;
  int *len = nullptr, *dpos = nullptr, *koff = nullptr, *kmax = nullptr, *col = nullptr;
  int *d_cptr = nullptr, *d_order = nullptr, *counter = nullptr;
  int *out_info = nullptr, *out_iter = nullptr, *out_fp32 = nullptr;
  double* out_res = nullptr;
  BatchProblem* d_problems = nullptr;
  double *vals = nullptr, *diag = nullptr, *q = nullptr, *mu = nullptr, *pli = nullptr;
  double4 *r = nullptr, *u = nullptr;
  float *vals32 = nullptr, *diag32 = nullptr, *q32 = nullptr, *mu32 = nullptr, *pli32 = nullptr;
  float4* r32 = nullptr;
  DevParams* prm = nullptr;
  int tuned_wpp[2] = {0, 0};
};

namespace fc3d {

template <typename T>
static int balloc(T** p, size_t n) {
  FC3D_HIP_TRY(hipMalloc(p, sizeof(T) * (n ? n : 1)));
  return 0;
}
template <typename T>
static int bupload(T** p, const std::vector<T>& h) {
  int rc = balloc(p, h.size());
  if (rc) return rc;
  if (!h.empty()) FC3D_HIP_TRY(hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return 0;
}

static BatchArgs batch_args(fc3d_hip_batch* b) {
  BatchArgs A;
  std::memset(&A, 0, sizeof A);
  SweepPtrs<double>& P = A.P64;
  P.len = b->len; P.dpos = b->dpos; P.koff = b->koff; P.kmax = b->kmax; P.col = b->col;
  P.vals = b->vals; P.diag = b->diag; P.q = b->q; P.mu = b->mu; P.r = b->r; P.u = b->u;
  P.pli = b->pli; P.part = nullptr;
  SweepPtrs<float>& Q = A.P32;
  Q.len = b->len; Q.dpos = b->dpos; Q.koff = b->koff; Q.kmax = b->kmax; Q.col = b->col;
  Q.vals = b->vals32; Q.diag = b->diag32; Q.q = b->q32; Q.mu = b->mu32; Q.r = b->r32;
  Q.u = nullptr; Q.pli = b->pli32; Q.part = nullptr;
  A.cptr = b->d_cptr;
  A.problems = b->d_problems;
  A.order = b->d_order;
  A.nprob = b->nprob;
  A.counter = b->counter;
  A.out_info = b->out_info;
  A.out_iter = b->out_iter;
  A.out_fp32_iter = b->out_fp32;
  A.out_residual = b->out_res;
  return A;
}

static int ensure_fp32_batch(fc3d_hip_batch* b) {
  if (b->vals32) return 0;
  const size_t nv = (size_t)b->nplanes * 9 * 64, nd = (size_t)b->nslices * 9 * 64,
               nq = (size_t)b->nslices * 3 * 64, ns = (size_t)b->nslots;
  int rc;
  if ((rc = balloc(&b->vals32, nv)) || (rc = balloc(&b->diag32, nd)) ||
      (rc = balloc(&b->q32, nq)) || (rc = balloc(&b->mu32, ns)) || (rc = balloc(&b->r32, ns)) ||
      (rc = balloc(&b->pli32, ns)))
    return rc;
  auto conv = [&](const double* a, float* c, size_t n) {
    if (n) hipLaunchKernelGGL((k_convert_b<double, float>), dim3((unsigned)((n + 255) / 256)),
                              dim3(256), 0, b->stream, a, c, n);
  };
  conv(b->vals, b->vals32, nv);
  conv(b->diag, b->diag32, nd);
  conv(b->q, b->q32, nq);
  conv(b->mu, b->mu32, ns);
  FC3D_HIP_TRY(hipGetLastError());
  FC3D_HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

// Launches the persistent kernel with `wpp` waves per problem; grid = enough
// workgroups to fill the device at the kernel's occupancy.
static int launch_batch(fc3d_hip_batch* b, int wpp, int mixed) {
  FC3D_HIP_TRY(hipMemsetAsync(b->counter, 0, sizeof(int), b->stream));
  int cus = 0;
  FC3D_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  const int per_cu = wpp == 1 ? 8 : 2;  // 2 waves/SIMD at this kernel's register use
  const int grid = std::max(1, std::min(b->nprob, cus * per_cu));
  const BatchArgs A = batch_args(b);
  if (wpp == 4)
    hipLaunchKernelGGL(k_batch<4>, dim3(grid), dim3(256), 0, b->stream, A, b->prm, mixed);
  else
    hipLaunchKernelGGL(k_batch<1>, dim3(grid), dim3(64), 0, b->stream, A, b->prm, mixed);
  FC3D_HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace fc3d

extern "C" {

int fc3d_hip_batch_create(int nprob, const int* nc, const int64_t* row_ptr,
                          const int64_t* col_idx, const double* blocks, const double* q,
                          const double* mu, int device, fc3d_hip_batch** out) {
  if (!out) return NUMERICS_ERR_NULL_POINTER;
  *out = nullptr;
  if (nprob <= 0 || !nc || !row_ptr || !col_idx || !blocks || !q || !mu) {
    set_error("fc3d_hip_batch_create: invalid arguments");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  std::vector<int> coff((size_t)nprob + 1, 0);
  for (int p = 0; p < nprob; p++) {
    if (nc[p] <= 0) {
      set_error("fc3d_hip_batch_create: problem %d has %d contacts", p, nc[p]);
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
    coff[(size_t)p + 1] = coff[(size_t)p] + nc[p];
  }
  const int ntot = coff[(size_t)nprob];
  const int64_t nnzb = row_ptr[ntot];
  if (row_ptr[0] != 0 || nnzb > 0x7fffffffLL) {
    set_error("fc3d_hip_batch_create: row_ptr[0] must be 0 and blocks < 2^31");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  RowInfo info((size_t)ntot);
  std::string err;
  for (int p = 0; p < nprob; p++) {
    if (int e = validate_rows(coff[(size_t)p], nc[p], row_ptr, col_idx, &info, &err)) {
      set_error("fc3d_hip_batch_create: problem %d: %s", p, err.c_str());
      return e;
    }
  }
  int rc = use_device(device);
  if (rc) return rc;
  fc3d_hip_batch* b = new fc3d_hip_batch();
  b->device = device;
  b->nprob = nprob;
  b->ntotal = ntot;
  b->coff = coff;
  b->nc.assign(nc, nc + nprob);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("fc3d_hip_batch_create: stream creation failed");
    delete b;
    return NUMERICS_ERR_DEVICE;
  }
  auto fail = [&](int code) {
    fc3d_hip_batch_destroy(b);
    return code;
  };
  // the block-diagonal graph of all problems: greedy colouring in global index
  // order equals the per-problem colouring (no edges across problems)
  {
    std::vector<int> rp((size_t)ntot + 1), ci((size_t)(nnzb > 0 ? nnzb : 1));
    for (int i = 0; i <= ntot; i++) rp[(size_t)i] = (int)row_ptr[i];
    for (int64_t k = 0; k < nnzb; k++) ci[(size_t)k] = (int)col_idx[k];
    int *d_rp = nullptr, *d_ci = nullptr, *d_col = nullptr;
    if ((rc = bupload(&d_rp, rp)) || (rc = bupload(&d_ci, ci)) ||
        (rc = balloc(&d_col, (size_t)ntot)))
      return fail(rc);
    int ncol_all = 0;
    rc = colour_on_device(ntot, d_rp, d_ci, d_col, b->stream, &ncol_all, &b->colour);
    hipFree(d_rp);
    hipFree(d_ci);
    hipFree(d_col);
    if (rc) return fail(rc);
  }
  HostLayout L;
  const int window = default_sort_window();
  b->problems.resize((size_t)nprob);
  for (int p = 0; p < nprob; p++) {
    BatchProblem& d = b->problems[(size_t)p];
    d.cptr_off = (int)L.colour_slice_ptr.size();
    d.slice_begin = L.nslices;
    int ncol = 0;
    append_problem(L, coff[(size_t)p], nc[p], row_ptr, col_idx, blocks, q, mu, b->colour.data(),
                   info, window, &ncol);
    d.ncolours = ncol;
    d.slice_end = L.nslices;
    double s = 0.0;
    for (int64_t i = 3 * (int64_t)coff[(size_t)p]; i < 3 * (int64_t)coff[(size_t)p + 1]; i++)
      s += q[i] * q[i];
    d.norm_q = std::sqrt(s);
  }
  if (L.nplanes * 64 > 0x7fffffffLL) {
    set_error("fc3d_hip_batch_create: sliced layout exceeds int32 indexing");
    return fail(NUMERICS_ERR_INVALID_ARGUMENT);
  }
  b->nslices = L.nslices;
  b->nslots = L.nslots;
  b->nplanes = L.nplanes;
  b->slot2orig = L.slot2orig;
  b->orig2slot = L.orig2slot;
  b->cptr = L.colour_slice_ptr;
  std::vector<int> order((size_t)nprob);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nc[x] > nc[y]; });
  if ((rc = bupload(&b->len, L.len)) || (rc = bupload(&b->dpos, L.dpos)) ||
      (rc = bupload(&b->koff, L.koff)) || (rc = bupload(&b->kmax, L.kmax)) ||
      (rc = bupload(&b->col, L.col)) || (rc = bupload(&b->vals, L.vals)) ||
      (rc = bupload(&b->diag, L.diag)) || (rc = bupload(&b->q, L.q)) ||
      (rc = bupload(&b->mu, L.mu)) || (rc = bupload(&b->d_cptr, b->cptr)) ||
      (rc = bupload(&b->d_problems, b->problems)) || (rc = bupload(&b->d_order, order)) ||
      (rc = balloc(&b->counter, 1)) || (rc = balloc(&b->out_info, (size_t)nprob)) ||
      (rc = balloc(&b->out_iter, (size_t)nprob)) || (rc = balloc(&b->out_fp32, (size_t)nprob)) ||
      (rc = balloc(&b->out_res, (size_t)nprob)) || (rc = balloc(&b->r, (size_t)b->nslots)) ||
      (rc = balloc(&b->u, (size_t)b->nslots)) || (rc = balloc(&b->pli, (size_t)b->nslots)) ||
      (rc = balloc(&b->prm, 1)))
    return fail(rc);
  if (hipMemset(b->r, 0, sizeof(double4) * (size_t)b->nslots) != hipSuccess ||
      hipMemset(b->u, 0, sizeof(double4) * (size_t)b->nslots) != hipSuccess) {
    set_error("fc3d_hip_batch_create: memset failed");
    return fail(NUMERICS_ERR_DEVICE);
  }
  *out = b;
  return 0;
}

void fc3d_hip_batch_destroy(fc3d_hip_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  void* ptrs[] = {b->len,  b->dpos,     b->koff,   b->kmax,    b->col,     b->d_cptr, b->d_order,
                  b->counter, b->out_info, b->out_iter, b->out_fp32, b->out_res, b->d_problems,
                  b->vals, b->diag,     b->q,      b->mu,      b->pli,     b->r,      b->u,
                  b->vals32, b->diag32, b->q32,    b->mu32,    b->pli32,   b->r32,    b->prm};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

int fc3d_hip_batch_order(const fc3d_hip_batch* b, int* order) {
  if (!b || !order) return NUMERICS_ERR_NULL_POINTER;
  int k = 0;
  for (int slot = 0; slot < b->nslots; slot++)
    if (b->slot2orig[(size_t)slot] >= 0) order[k++] = b->slot2orig[(size_t)slot];
  return 0;
}

int fc3d_hip_batch_set_reaction(fc3d_hip_batch* b, const double* reaction) {
  if (!b || !reaction) return NUMERICS_ERR_NULL_POINTER;
  int rc = use_device(b->device);
  if (rc) return rc;
  std::vector<double4> buf((size_t)b->nslots, make_double4(0, 0, 0, 0));
  for (int i = 0; i < b->ntotal; i++)
    buf[(size_t)b->orig2slot[(size_t)i]] =
        make_double4(reaction[3 * i], reaction[3 * i + 1], reaction[3 * i + 2], 0.0);
  FC3D_HIP_TRY(hipMemcpy(b->r, buf.data(), sizeof(double4) * buf.size(), hipMemcpyHostToDevice));
  return 0;
}

static int batch_get(const fc3d_hip_batch* b, const double4* d, double* out) {
  int rc = use_device(b->device);
  if (rc) return rc;
  std::vector<double4> buf((size_t)b->nslots);
  FC3D_HIP_TRY(hipMemcpy(buf.data(), d, sizeof(double4) * buf.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < b->ntotal; i++) {
    const double4 v = buf[(size_t)b->orig2slot[(size_t)i]];
    out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
  }
  return 0;
}

int fc3d_hip_batch_get_reaction(const fc3d_hip_batch* b, double* reaction) {
  if (!b || !reaction) return NUMERICS_ERR_NULL_POINTER;
  return batch_get(b, b->r, reaction);
}

int fc3d_hip_batch_get_velocity(const fc3d_hip_batch* b, double* velocity) {
  if (!b || !velocity) return NUMERICS_ERR_NULL_POINTER;
  return batch_get(b, b->u, velocity);
}

int fc3d_hip_batch_solve(fc3d_hip_batch* b, const fc3d_hip_params* p, fc3d_hip_batch_stats* s,
                         int* info, int* iterations, double* residual) {
  if (!b || !p || !s) return NUMERICS_ERR_NULL_POINTER;
  const int ls = p->local_solver;
  if (!(ls == 565 || ls == 551 || ls == 550 || ls == 553 || ls == 552)) {
    set_error("fc3d_hip_batch_solve: local solver %d not supported", ls);
    return NUMERICS_ERR_INVALID_SOLVER;
  }
  const int ls32 = p->fp32_local_solver;
  if (!(ls32 == 0 || ls32 == 565 || ls32 == 551 || ls32 == 550 || ls32 == 553 || ls32 == 552)) {
    set_error("fc3d_hip_batch_solve: FP32 local solver %d not supported", ls32);
    return NUMERICS_ERR_INVALID_SOLVER;
  }
  if (p->error_evaluation < 0 || p->error_evaluation > 3 || p->warm_start >= 2) {
    set_error("fc3d_hip_batch_solve: unsupported evaluation mode or warm start");
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  int rc = use_device(b->device);
  if (rc) return rc;
  std::memset(s, 0, sizeof *s);
  if (p->accel_depth != 0) {
    set_error("fc3d_hip_batch_solve: Anderson acceleration is not available for batches");
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  if (p->freeze_contacts != 0) {
    set_error("fc3d_hip_batch_solve: contact freezing is not available in the batched solver");
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  DevParams d;
  std::memset(&d, 0, sizeof d);
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
  FC3D_HIP_TRY(hipMemcpyAsync(b->prm, &d, sizeof d, hipMemcpyHostToDevice, b->stream));
  if (p->warm_start == 1)
    FC3D_HIP_TRY(hipMemsetAsync(b->r, 0, sizeof(double4) * (size_t)b->nslots, b->stream));
  const int mixed = p->precision == 1;
  if (mixed && (rc = ensure_fp32_batch(b))) return rc;
  // schedule: 1 = one wave per problem, 2 = four waves per problem; 0 = the
  // autotuner's choice (timed on the whole batch, reaction restored)
  int wpp = p->schedule == 2 ? 4 : (p->schedule == 1 ? 1 : 0);
  if (p->schedule < 0 || p->schedule > 2) {
    set_error("fc3d_hip_batch_solve: schedule %d out of range [0, 2]", p->schedule);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  if (wpp == 0) {
    if (!b->tuned_wpp[mixed]) {
      double4* saved = nullptr;
      FC3D_HIP_TRY(hipMalloc(&saved, sizeof(double4) * (size_t)b->nslots));
      FC3D_HIP_TRY(hipMemcpyAsync(saved, b->r, sizeof(double4) * (size_t)b->nslots,
                                  hipMemcpyDeviceToDevice, b->stream));
      DevParams t = d;
      t.max_iter = std::min(p->max_iter, 8);
      t.tol = 0.0;
      FC3D_HIP_TRY(hipMemcpyAsync(b->prm, &t, sizeof t, hipMemcpyHostToDevice, b->stream));
      float best = 1e30f;
      hipEvent_t e0, e1;
      FC3D_HIP_TRY(hipEventCreate(&e0));
      FC3D_HIP_TRY(hipEventCreate(&e1));
      for (int cand : {1, 4}) {
        FC3D_HIP_TRY(hipMemcpyAsync(b->r, saved, sizeof(double4) * (size_t)b->nslots,
                                    hipMemcpyDeviceToDevice, b->stream));
        FC3D_HIP_TRY(hipEventRecord(e0, b->stream));
        if ((rc = launch_batch(b, cand, mixed))) return rc;
        FC3D_HIP_TRY(hipEventRecord(e1, b->stream));
        FC3D_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
          best = ms;
          b->tuned_wpp[mixed] = cand;
        }
      }
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      FC3D_HIP_TRY(hipMemcpyAsync(b->r, saved, sizeof(double4) * (size_t)b->nslots,
                                  hipMemcpyDeviceToDevice, b->stream));
      FC3D_HIP_TRY(hipMemcpyAsync(b->prm, &d, sizeof d, hipMemcpyHostToDevice, b->stream));
      FC3D_HIP_TRY(hipStreamSynchronize(b->stream));
      hipFree(saved);
    }
    wpp = b->tuned_wpp[mixed];
  }
  hipEvent_t t0, t1;
  FC3D_HIP_TRY(hipEventCreate(&t0));
  FC3D_HIP_TRY(hipEventCreate(&t1));
  FC3D_HIP_TRY(hipEventRecord(t0, b->stream));
  if ((rc = launch_batch(b, wpp, mixed))) return rc;
  FC3D_HIP_TRY(hipEventRecord(t1, b->stream));
  FC3D_HIP_TRY(hipEventSynchronize(t1));
  float ms = 0.f;
  hipEventElapsedTime(&ms, t0, t1);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  std::vector<int> hinfo((size_t)b->nprob), hiter((size_t)b->nprob), hfp32((size_t)b->nprob);
  std::vector<double> hres((size_t)b->nprob);
  FC3D_HIP_TRY(hipMemcpy(hinfo.data(), b->out_info, sizeof(int) * hinfo.size(), hipMemcpyDeviceToHost));
  FC3D_HIP_TRY(hipMemcpy(hiter.data(), b->out_iter, sizeof(int) * hiter.size(), hipMemcpyDeviceToHost));
  FC3D_HIP_TRY(hipMemcpy(hfp32.data(), b->out_fp32, sizeof(int) * hfp32.size(), hipMemcpyDeviceToHost));
  FC3D_HIP_TRY(hipMemcpy(hres.data(), b->out_res, sizeof(double) * hres.size(), hipMemcpyDeviceToHost));
  s->solve_ms = ms;
  s->schedule = wpp == 4 ? 2 : 1;
  for (int q = 0; q < b->nprob; q++) {
    if (info) info[q] = hinfo[(size_t)q];
    if (iterations) iterations[q] = hiter[(size_t)q];
    if (residual) residual[q] = hres[(size_t)q];
    s->total_sweeps += hiter[(size_t)q];
    s->fp32_sweeps += hfp32[(size_t)q];
    if (hres[(size_t)q] <= p->tol) {
      s->converged += 1;
      s->converged_contacts += b->nc[(size_t)q];
    }
    s->max_residual = std::max(s->max_residual, hres[(size_t)q]);
  }
  return 0;
}

}  // extern "C"
