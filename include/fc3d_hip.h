/* fc3d_hip.h — C-ABI of the MI355X (gfx950) path: plain pointers and sizes,
 * no HIP or torch types in the signatures.
 *
 * Host arrays are in the caller's contact order ("original order").  A
 * problem handle keeps W, q, mu, the reaction and the velocity resident in
 * HBM in the colour-major slot order described in the library's design notes;
 * set/get functions permute on the way in and out.
 *
 * Every function returns 0 on success or a negative NUMERICS_ERR_* code
 * (NUMERICS_ERR_DEVICE when the HIP runtime or the device fails); the message
 * is available from fc3d_hip_last_error().  Nothing here falls back to the CPU.
 */
#ifndef FC3D_HIP_H
#define FC3D_HIP_H

#include <stdint.h>

/* FC3D_HIP_STANDALONE: the device layer alone, for callers that bring their
 * own numerics headers (the return codes are numerics_errors.h:41-67 plus
 * NUMERICS_ERR_DEVICE = -60). */
#ifdef FC3D_HIP_STANDALONE
#ifndef FC3D_API
#define FC3D_API __attribute__((visibility("default")))
#endif
#else
#include "fc3d_numerics.h"
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of one coloured NSGS solve.  The first block mirrors the
 * iparam/dparam entries the reference fc3d_nsgs reads (fc3d_nsgs.c:560-950,
 * onecontact...c:1017-1056, fc3d_projection.c:350-514); the second block holds
 * the additions of this library. */
typedef struct fc3d_hip_params {
  int max_iter;              /* iparam[SICONOS_IPARAM_MAX_ITER]                   */
  double tol;                /* dparam[SICONOS_DPARAM_TOL]                        */
  int error_evaluation;      /* iparam[13]: FULL, LIGHT, LIGHT_WITH_FULL_FINAL, ADAPTIVE */
  int error_frequency;       /* iparam[14]                                        */
  int filter_local_solution; /* iparam[16]                                        */
  int relaxation;            /* iparam[10]                                        */
  double omega;              /* dparam[10]                                        */
  int local_solver;          /* internal solver id: 565, 551, 550, 553, 552       */
  int local_max_iter;        /* local iparam[0]                                   */
  double local_tol;          /* local dparam[0]                                   */
  int local_formulation;     /* local iparam[10]: AC STD or JeanMoreau STD        */
  int rho_strategy;          /* local iparam[9]                                   */
  double rho_constant;       /* local dparam[3]                                   */
  int ls_max_iter;           /* local iparam[12]                                  */
  int hybrid_strategy;       /* local iparam[14]                                  */
  int hybrid_max_loop;       /* local iparam[15]                                  */
  int hybrid_max_iter;       /* local iparam[16]                                  */
  /* additions */
  int precision;             /* 0: FP64 sweeps; 1: FP32 sweeps + FP64 refinement  */
  int warm_start;            /* 0: resident reaction; 1: zero; 2: learned model,
                                kept only if its FP64 residual after
                                start_trial_sweeps sweeps is below the zero
                                start's; 3: the same against the resident
                                reaction (previous time step)                   */
  int schedule;              /* 0: autotuned; 1..10 explicit (7: single launch)   */
  int sweeps_per_launch;     /* sweeps captured per graph replay (>=1); the FP64
                                sweeps that end a precision = 1 solve replay at
                                most 4 at a time                                 */
  double fp32_tol;           /* FP32 phase hands over when its light error < this */
  double fp32_local_tol;     /* local tolerance used by FP32 sweeps               */
  int instrument;            /* 1: time each sweep-kernel launch with HIP events  */
  /* Anderson acceleration of the sweep map (no reference counterpart):
   * x <- G(x) - dG gamma with the last accel_depth differences, on every
   * accel_every-th sweep, history restarted when |G(x) - x| grows. */
  int accel_depth;           /* 0: off (default); 1..8 history columns           */
  int accel_every;           /* Anderson step on every n-th sweep (default 2)     */
  double accel_reg;          /* Tikhonov weight relative to the Gram trace (1e-10)*/
  /* Sweeps in the velocity form (needs fc3d_hip_problem_set_bodies): the
   * local problem from u = H v with body velocities v = v_free + M^-1 H^T r
   * kept current, instead of the row of W; the same iteration in exact
   * arithmetic.  Convergence is still judged on the FP64 residual with W. */
  int velocity_form;         /* 0: rows of W (default); 1: velocity form          */
  /* Local solver of the FP32 sweeps (precision = 1): 0 keeps local_solver;
   * 550 (undamped AC Newton) is the cheapest that converges on piles.  The
   * FP64 sweeps always use local_solver, and the result is judged on the
   * FP64 residual. */
  int fp32_local_solver;
  /* Contact freezing (iparam[SICONOS_NSGS_FREEZING_CONTACT], fc3d_nsgs.c:
   * 719-800): a contact whose step meets the reference's criteria from the
   * 10th sweep on is skipped for this many sweeps; 0: off.  FP64 sweeps on
   * the rows of W only. */
  int freeze_contacts;
  /* Focused passes (velocity form; no reference counterpart): after every
   * focus_every sweeps from sweep focus_after on, the focus_count contacts
   * whose reactions moved most in the last sweep are taken (all that moved,
   * if fewer did), if they carry at least focus_share of that sweep's sum
   * |r_new - r_old|^2; every contact
   * touching one of their bodies then gets focus_passes extra Gauss-Seidel
   * passes (colour by colour, the sweep's own local step).  On granular
   * piles the last hundreds of sweeps mostly move a few small hyperstatic
   * clusters; the passes settle them.  0 passes: off (default).  The result
   * is judged on the FP64 residual as always. */
  int focus_passes;
  int focus_every;           /* default 16                                         */
  int focus_after;           /* default 64                                         */
  int focus_count;           /* default 256                                        */
  double focus_share;        /* default 0.5                                        */
  /* warm_start 2, 3: sweeps each start gets before the two are compared
   * (default 1, the first-sweep rule) */
  int start_trial_sweeps;
  /* Relaxation of the focused passes only, in (0, 2): each local solution
   * becomes focus_omega * r_local + (1 - focus_omega) * r_old (the
   * reference's relaxation formula, fc3d_nsgs.c:73-78 and 766-767, applied
   * in the focus passes and not in the sweeps).  Default 1.2: the clusters'
   * reactions drift along near-null directions of W, which over-relaxation
   * crosses in fewer passes (on 1M-contact piles 4-16% fewer sweeps than
   * 1.0; 1.6 can stall the solve). */
  double focus_omega;
} fc3d_hip_params;

typedef struct fc3d_hip_stats {
  int info;              /* reference info: 0, 1, or NUMERICS_ERR_MAX_ITER      */
  int iterations;        /* sweeps done (FP32 + FP64)                           */
  int fp32_sweeps;       /* of which in FP32                                    */
  double residual;       /* final fc3d_compute_error value (FP64)               */
  double incr_error;     /* last light error                                    */
  int num_colours;
  int warm_start_used;   /* 0 resident, 1 zero, 2 learned                       */
  double solve_ms;       /* device time from first sweep to final residual      */
  double sweep_kernel_ms;     /* instrumented: summed sweep-kernel durations    */
  int64_t sweep_kernel_launches;
  double sweep_kernel_bytes;  /* algorithmic bytes moved by those launches      */
  int schedule;               /* sweep-kernel schedule used (autotuned when 0 given) */
  /* warm_start = 2, 3: FP64 residual after one sweep from the fallback start
   * (zero, or the resident reaction) and from the learned start */
  double start_residual_fallback;
  double start_residual_learned;
  int accel_steps;            /* Anderson steps taken                            */
  int accel_restarts;         /* history restarts                                */
  int accel_rejects;          /* steps undone by the safeguard (each costs the   *
                               * sweep that judged it)                           */
  /* instrumented: the FP32-sweep share of the three sweep_kernel_* figures */
  double sweep_kernel_ms_fp32;
  int64_t sweep_kernel_launches_fp32;
  double sweep_kernel_bytes_fp32;
  /* instrumented, velocity form: the part of the three *_fp32 figures spent
   * in tail launches (the last small colours of a sweep in one workgroup);
   * the rest is one launch per colour */
  double sweep_tail_ms_fp32;
  int64_t sweep_tail_launches_fp32;
  double sweep_tail_bytes_fp32;
  int start_trial_sweeps;     /* warm_start 2, 3: sweeps of the discarded start  */
  int focus_rounds;           /* focus rounds that ran passes                    */
  int64_t focus_listed;       /* contacts listed over those rounds               */
  int focus_due;              /* focus rounds due (with or without passes)       */
  double focus_last_share;    /* the last due round: share of sum |dr|^2 carried *
                               * by its selection                                */
  int focus_lds_rounds;       /* of focus_rounds, those whose body velocities fit *
                               * in LDS (the rest kept them in HBM)              */
} fc3d_hip_stats;

FC3D_API int fc3d_hip_device_count(void);
FC3D_API const char* fc3d_hip_last_error(void);
/* Defaults equal to solver_options_create(SICONOS_FRICTION_3D_NSGS) with the
 * hybrid local solver (fc3d_nsgs.c:919-950, onecontact...c:1037-1056). */
FC3D_API void fc3d_hip_params_default(fc3d_hip_params* p);

/* ---- resident problem ---------------------------------------------------- */
typedef struct fc3d_hip_problem fc3d_hip_problem;

/* Uploads W (CSR over 3x3 blocks, column-major values, diagonal block present
 * in every row), q and mu; colours the contacts on the GPU and builds the
 * colour-major sliced layout.  Replaces the per-call set-up of fc3d_nsgs
 * (fc3d_nsgs.c:560-630) and of its local solver (onecontact...c:62-187). */
FC3D_API int fc3d_hip_problem_create(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                                     const double* blocks, const double* q, const double* mu,
                                     int device, fc3d_hip_problem** out);
FC3D_API void fc3d_hip_problem_destroy(fc3d_hip_problem* h);
FC3D_API int fc3d_hip_problem_num_colours(const fc3d_hip_problem* h);
FC3D_API int64_t fc3d_hip_problem_nnzb(const fc3d_hip_problem* h);
/* colour of each contact (original order) */
FC3D_API int fc3d_hip_problem_colours(const fc3d_hip_problem* h, int* colour);
/* sweep order: the original index of the contact at each sweep position */
FC3D_API int fc3d_hip_problem_order(const fc3d_hip_problem* h, int* order);
/* New values for an unchanged structure (a new time step of the same contact
 * set): W's blocks in the CSR order given at creation, q and mu in the
 * caller's order; any pointer may be NULL (kept).  New blocks or q drop
 * factors attached with fc3d_hip_problem_set_bodies.  The colouring, layout,
 * tuned schedule and captured graphs stay valid, and so does the resident
 * reaction, which warm_start = 0 or 3 then reuses (LinearOSNS.cpp:873-891
 * keeps the previous step's reaction as the start in the same way). */
FC3D_API int fc3d_hip_problem_update(fc3d_hip_problem* h, const double* blocks, const double* q,
                                     const double* mu);
/* The factors of W = H M^-1 H^T for velocity-form sweeps (no reference
 * counterpart; the factors are the inputs of fc3d_hip_assemble, in the
 * caller's contact order).  Refused when the factors do not reproduce the
 * resident diagonal blocks of W or the resident q = H v_free (to 1e-10 of
 * the terms' magnitude: the sweeps compute H v, so q can hold no other
 * term), or when two contacts of one body share a colour (W must hold a
 * block for every pair of contacts sharing a body).
 * When every body's diagonal M^-1 is (m_t, m_t, m_t, m_r, m_r, m_r) the FP32
 * copy keeps (m_t, m_r) a body, in the padding of its FP32 velocity vectors
 * (the same numbers; FC3D_HIP_MINV_ISO=0 in the environment keeps six). */
FC3D_API int fc3d_hip_problem_set_bodies(fc3d_hip_problem* h, int nbodies, const int* body_a,
                                         const int* body_b, const double* H, const double* Minv,
                                         int minv_diagonal, const double* vfree);
/* Layout of the attached factors: 0 none, 1 general 3x6 blocks of H, 2 the
 * frames layout, taken whenever every contact has rigid-body blocks
 * H_a = [T | T S(l_a)], H_b = -[T | T S(l_b)] (FC3D_HIP_FRAMES=0 in the
 * environment at set_bodies keeps layout 1): T and the lever arms are kept,
 * 15 values per contact instead of 36, and the sweeps rebuild the blocks
 * from them; W_cc is read from W's resident diagonal blocks either way. */
FC3D_API int fc3d_hip_problem_body_layout(const fc3d_hip_problem* h);
FC3D_API int fc3d_hip_set_reaction(fc3d_hip_problem* h, const double* reaction);
FC3D_API int fc3d_hip_get_reaction(const fc3d_hip_problem* h, double* reaction);
FC3D_API int fc3d_hip_get_velocity(const fc3d_hip_problem* h, double* velocity);

/* fc3d_nsgs.c:560-917 in colour order on the resident reaction. */
FC3D_API int fc3d_hip_solve(fc3d_hip_problem* h, const fc3d_hip_params* p, fc3d_hip_stats* s);
/* fc3d_hip_solve with an observer (no reference counterpart: it carries the
 * reference's per-sweep statistics callback, SolverOptions.h:36-44 called at
 * fc3d_nsgs.c:103-110).  The observer runs on the host after every graph
 * replay, i.e. every sweeps_per_launch sweeps (1: after every sweep, as the
 * reference does), with the sweeps done so far and the last light error;
 * while it runs, fc3d_hip_get_reaction and fc3d_hip_get_velocity return the
 * current iterate and the velocity of the last full residual. */
typedef void (*fc3d_hip_sweep_observer)(void* env, int sweeps_done, double incr_error);
FC3D_API int fc3d_hip_solve_observed(fc3d_hip_problem* h, const fc3d_hip_params* p,
                                     fc3d_hip_stats* s, fc3d_hip_sweep_observer observer,
                                     void* env);
/* fc3d_compute_error.c:70-122 on the resident reaction; writes the resident
 * velocity.  norm < 0 uses the stored |q|. */
FC3D_API int fc3d_hip_compute_error(fc3d_hip_problem* h, double norm, double* error);
/* fc3d_compute_error on device arrays, allocating nothing.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every pointer is device memory the caller owns: W as CSR over 3x3 blocks
 * (row_ptr nc + 1 entries, col_idx and blocks nnzb, column-major blocks),
 * q [3nc], mu [nc], reaction [3nc] in; velocity [3nc] (u = q + W r, each
 * row's blocks added in their stored order, as fc3d_compute_error does with
 * SBM_gemv_3x3) and error [1] (fc3d_compute_error.c:70-122 relative to
 * max(norm, |r|, |u|)) out.  The kernels are queued on `stream` (a
 * hipStream_t; NULL: the default stream) and the call returns without
 * waiting.  nc is limited to 4,194,304 contacts (larger is refused): the
 * partial sums use a fixed 1.5 MiB device array of the library on each GPU,
 * so calls that overlap on one GPU race.  A row whose range or column indices
 * fall outside [0, nnzb) or [0, nc) yields a NaN velocity and error. */
FC3D_API int fc3d_hip_compute_error_device(int nc, int64_t nnzb, const int64_t* row_ptr,
                                           const int64_t* col_idx, const double* blocks,
                                           const double* q, const double* mu,
                                           const double* reaction, double norm,
                                           double* velocity, double* error, int device,
                                           void* stream);
/* SparseBlockMatrix.c:360-419: y += W x (host vectors, original order). */
FC3D_API int fc3d_hip_gemv(fc3d_hip_problem* h, const double* x, double* y);

/* ---- unit kernels over independent contacts ------------------------------ */
/* fc3d_AlartCurnier_functions.c:506-753; formulation as local iparam[10]. */
FC3D_API int fc3d_hip_alart_curnier(int n, const double* R, const double* U, const double* mu,
                                    const double* rho, int formulation, double* F, double* A,
                                    double* B);
/* projectionOnCone.c:32-50, in place. */
FC3D_API int fc3d_hip_projection_on_cone(int n, double* r, const double* mu);
/* One local solve per contact (onecontact...c:221-290, 920-1015;
 * fc3d_projection.c:350-603).  state: 4 doubles per contact [PLI rho, AC
 * rho x3] in/out (initialised from W when init_state != 0).  With
 * p->precision = 1 the FP32 sweeps' local solve runs instead
 * (fp32_local_solver, fp32_local_tol, the inputs rounded to float). */
FC3D_API int fc3d_hip_local_solve(int n, const double* W, const double* q, const double* mu,
                                  double* r, const fc3d_hip_params* p, double* state,
                                  int init_state, int* info);

/* ---- many independent problems in one persistent launch -------------------- */
/* No reference counterpart (the reference solves one problem per fc3d_driver
 * call); each problem follows fc3d_nsgs.c:560-917 in its own colour order. */
typedef struct fc3d_hip_batch fc3d_hip_batch;

typedef struct fc3d_hip_batch_stats {
  double solve_ms;            /* device time of the persistent launch            */
  int converged;              /* problems with final residual <= tol             */
  int64_t converged_contacts; /* contacts of those problems                      */
  int64_t total_sweeps;
  int64_t fp32_sweeps;
  double max_residual;
  int schedule;               /* 1: one wave per problem, 2: four waves per problem */
} fc3d_hip_batch_stats;

/* Problem p owns contacts [sum(nc[0..p)), sum(nc[0..p])) of the concatenated
 * CSR (row_ptr has sum(nc) + 1 entries; column indices are global and must
 * stay inside their problem). */
FC3D_API int fc3d_hip_batch_create(int nprob, const int* nc, const int64_t* row_ptr,
                                   const int64_t* col_idx, const double* blocks, const double* q,
                                   const double* mu, int device, fc3d_hip_batch** out);
FC3D_API void fc3d_hip_batch_destroy(fc3d_hip_batch* b);
/* sweep order of every problem, problems one after the other (global indices) */
FC3D_API int fc3d_hip_batch_order(const fc3d_hip_batch* b, int* order);
FC3D_API int fc3d_hip_batch_set_reaction(fc3d_hip_batch* b, const double* reaction);
FC3D_API int fc3d_hip_batch_get_reaction(const fc3d_hip_batch* b, double* reaction);
FC3D_API int fc3d_hip_batch_get_velocity(const fc3d_hip_batch* b, double* velocity);
/* Solves every problem; per-problem info / sweeps / residual (may be NULL). */
FC3D_API int fc3d_hip_batch_solve(fc3d_hip_batch* b, const fc3d_hip_params* p,
                                  fc3d_hip_batch_stats* s, int* info, int* iterations,
                                  double* residual);

/* ---- learned warm start ----------------------------------------------------- */
/* No reference counterpart.  A per-contact MLP (22 features -> 128 -> 128 ->
 * 128 -> 3, GELU) predicts the converged reaction of each contact, scaled by
 * |q_c| / W_cc(0,0), from its local problem.  Forward, backward and weight
 * gradients run on the matrix cores (BF16 operands, FP32 accumulation); the
 * FP32 master weights are updated by Adam on the GPU.
 *
 * Features of contact c (with s = |q_c|, w = W_cc(0,0), neighbours j != c of
 * its block row, n the normal component):
 *   0-2  q_c / s               3-11 W_cc / w (column-major)   12 mu_c
 *   13   log w                 14   log s                     15 neighbour count
 *   16   mean_j |q_j| / s      17   max_j |q_j| / s
 *   18-20 (sum_j W_cj y_j) / s with y_j = (max(0, -q_jn) / W_jj(0,0), 0, 0)
 *   21   fraction of neighbours with q_jn < 0
 * Target: r_c * w / s.
 *
 * Parameters (FP32, FC3D_MLP_NPARAMS): W1[128][32] (inputs padded to 32, the
 * last 10 columns stay zero), b1[128], W2[128][128], b2[128], W3[128][128],
 * b3[128], W4[16][128] (rows 3..15 stay zero), b4[16]. */
#define FC3D_MLP_FEATURES 22
#define FC3D_MLP_HIDDEN 128
#define FC3D_MLP_OUTPUTS 3
#define FC3D_MLP_NPARAMS 39312

typedef struct fc3d_hip_mlp fc3d_hip_mlp;

/* Seeded He-uniform initialisation; feature mean 0 and std 1. */
FC3D_API int fc3d_hip_mlp_create(uint64_t seed, int device, fc3d_hip_mlp** out);
FC3D_API void fc3d_hip_mlp_destroy(fc3d_hip_mlp* m);
FC3D_API int fc3d_hip_mlp_get_params(const fc3d_hip_mlp* m, float* params, float* mean,
                                     float* stdv);
/* Any pointer may be NULL (kept); stdv entries must be positive. */
FC3D_API int fc3d_hip_mlp_set_params(fc3d_hip_mlp* m, const float* params, const float* mean,
                                     const float* stdv);
/* y[n][3] from raw features x[n][22] (standardised with the model's mean/std). */
FC3D_API int fc3d_hip_mlp_forward(fc3d_hip_mlp* m, int n, const float* x, float* y);
/* Mean squared error over n <= 65536 rows and its gradient w.r.t. every
 * parameter (grads: FC3D_MLP_NPARAMS floats, may be NULL). */
FC3D_API int fc3d_hip_mlp_gradients(fc3d_hip_mlp* m, int n, const float* x, const float* t,
                                    float* grads, float* loss);
/* Adam (lr, beta 0.9/0.999, eps 1e-8) over `epochs` seeded shuffles of the n
 * samples in minibatches of `batch` rows (a positive multiple of 64).  With
 * set_normalisation != 0 the feature mean/std are first taken from x.
 * final_loss: mean squared error of the last minibatch. */
FC3D_API int fc3d_hip_mlp_train(fc3d_hip_mlp* m, int n, const float* x, const float* t,
                                int epochs, int batch, float lr, uint64_t seed,
                                int set_normalisation, float* final_loss);
/* Features of every contact (original order, [nc][22]) and, if target is not
 * NULL, the scaled resident reaction ([nc][3]) as a training target. */
FC3D_API int fc3d_hip_problem_features(fc3d_hip_problem* h, float* feat, float* target);
/* The model used by warm_start = 2, 3 on this problem (not owned; NULL detaches). */
FC3D_API int fc3d_hip_problem_set_model(fc3d_hip_problem* h, fc3d_hip_mlp* m);
/* The model fc3d_nsgs_hip attaches when iparam[WARM_START] = 2 (not owned). */
FC3D_API void fc3d_hip_set_default_model(fc3d_hip_mlp* m);
FC3D_API fc3d_hip_mlp* fc3d_hip_default_model(void);

/* ---- colouring ----------------------------------------------------------- */
/* Smallest colour not used by a lower-index neighbour, in index order;
 * returns the number of colours. */
FC3D_API int fc3d_hip_colour(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                             int* colour);

#ifdef __cplusplus
}
#endif
#endif
