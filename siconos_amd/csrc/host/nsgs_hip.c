/* nsgs_hip.c — the SICONOS_FRICTION_3D_NSGS_HIP solver: coloured NSGS on
 * the GPU behind the reference registry and driver.
 *
 * Options keep the meaning fc3d_nsgs gives them (fc3d_nsgs.c:560-950 and the
 * local solver options of onecontact...c:1017-1143, fc3d_projection.c:917-960);
 * the extra slots 20-27 are listed in fc3d_numerics.h.  The sweep order is
 * colour-major (fc3d_hip_problem_order), so the result equals the reference
 * run in that order, not in the natural one. */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "fc3d_hip.h"
#include "fc3d_numerics.h"

int fc3d_resident_acquire(FrictionContactProblem* problem, int device, int mode,
                          fc3d_hip_problem** out, int* token);
void fc3d_resident_release(fc3d_hip_problem* h, int token);

/* The reference's per-sweep statistics callback (fc3d_nsgs.c:103-110) behind
 * the device solve's observer: the current reaction and the velocity of the
 * last full residual are copied to the caller's arrays first. */
typedef struct {
  fc3d_hip_problem* h;
  SolverOptions* options;
  FrictionContactProblem* problem;
  double* reaction;
  double* velocity;
  int failed;
} observer_env;

static void observe_sweep(void* env, int sweeps_done, double incr_error) {
  observer_env* e = (observer_env*)env;
  (void)sweeps_done;
  if (fc3d_hip_get_reaction(e->h, e->reaction) || fc3d_hip_get_velocity(e->h, e->velocity)) {
    e->failed = 1;
    return;
  }
  Callback* cb = e->options->callback;
  cb->collectStatsIteration(cb->env, (size_t)(3 * e->problem->numberOfContacts), e->reaction,
                            e->velocity, incr_error, NULL);
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

/* Maps reference options onto the C-ABI parameters. */
int fc3d_nsgs_hip_params_from_options(const SolverOptions* o, fc3d_hip_params* p) {
  fc3d_hip_params_default(p);
  if (o->numberOfInternalSolvers < 1 || !o->internalSolvers || !o->internalSolvers[0])
    return NUMERICS_ERR_INVALID_OPTION;
  if (o->iparam[SICONOS_NSGS_SHUFFLE] != 0) return NUMERICS_ERR_NOT_IMPLEMENTED;
  if (o->iparam[SICONOS_FRICTION_3D_IPARAM_INTERNAL_ERROR_STRATEGY] !=
      SICONOS_NSGS_INTERNAL_ERROR_STRATEGY_GIVEN_VALUE)
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  const SolverOptions* l = o->internalSolvers[0];
  p->max_iter = o->iparam[SICONOS_IPARAM_MAX_ITER];
  p->tol = o->dparam[SICONOS_DPARAM_TOL];
  p->error_evaluation = o->iparam[SICONOS_NSGS_ERROR_EVALUATION_TYPE];
  p->error_frequency = o->iparam[SICONOS_NSGS_ERROR_EVALUATION_FREQUENCY];
  p->filter_local_solution = o->iparam[SICONOS_NSGS_FILTER_LOCAL_SOLUTION];
  p->relaxation = o->iparam[SICONOS_NSGS_RELAXATION];
  p->omega = o->dparam[SICONOS_NSGS_RELAXATION_VALUE];
  p->local_solver = l->solverId;
  p->local_max_iter = l->iparam[SICONOS_IPARAM_MAX_ITER];
  p->local_tol = l->dparam[SICONOS_DPARAM_TOL];
  p->local_formulation = l->iparam[SICONOS_FRICTION_3D_NSN_FORMULATION];
  p->rho_strategy = l->iparam[SICONOS_FRICTION_3D_NSN_RHO_STRATEGY];
  p->rho_constant = l->dparam[SICONOS_FRICTION_3D_NSN_RHO];
  p->ls_max_iter = l->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH_MAX_ITER];
  p->hybrid_strategy = l->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY];
  p->hybrid_max_loop = l->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP];
  p->hybrid_max_iter = l->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER];
  p->precision = o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_PRECISION];
  p->warm_start = o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_WARM_START];
  p->schedule = o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_SCHEDULE];
  if (o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_SWEEPS_PER_LAUNCH] > 0)
    p->sweeps_per_launch = o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_SWEEPS_PER_LAUNCH];
  if (o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_TOL] > 0.0)
    p->fp32_tol = o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_TOL];
  p->accel_depth = o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_DEPTH];
  if (o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_EVERY] > 0)
    p->accel_every = o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_EVERY];
  if (o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_REG] > 0.0)
    p->accel_reg = o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_REG];
  p->fp32_local_solver = (int)o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_LOCAL_SOLVER];
  if (o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_LOCAL_TOL] > 0.0)
    p->fp32_local_tol = o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_LOCAL_TOL];
  p->freeze_contacts =
      o->iparam[SICONOS_NSGS_FREEZING_CONTACT] > 0 ? o->iparam[SICONOS_NSGS_FREEZING_CONTACT] : 0;
  if (o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_START_TRIAL_SWEEPS] >= 1.0)
    p->start_trial_sweeps = (int)o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_START_TRIAL_SWEEPS];
  return 0;
}

void fc3d_nsgs_hip(FrictionContactProblem* problem, double* reaction, double* velocity,
                   int* info, SolverOptions* options) {
  fc3d_hip_params p;
  int rc = fc3d_nsgs_hip_params_from_options(options, &p);
  if (rc) {
    fprintf(stderr, "fc3d_nsgs_hip: unsupported options (%s)\n", numerics_error_string(rc));
    *info = rc;
    return;
  }
  double t0 = now_ms();
  fc3d_hip_problem* h = NULL;
  int token = -1;
  rc = fc3d_resident_acquire(problem, options->iparam[SICONOS_FRICTION_3D_NSGS_HIP_DEVICE],
                             (int)options->dparam[SICONOS_FRICTION_3D_NSGS_HIP_RESIDENCY], &h,
                             &token);
  if (!rc) rc = fc3d_hip_set_reaction(h, reaction);
  /* warm starts 2 and 3 use the process-wide model (fc3d_hip_set_default_model) */
  if (!rc && p.warm_start >= 2) rc = fc3d_hip_problem_set_model(h, fc3d_hip_default_model());
  double t1 = now_ms();
  fc3d_hip_stats s;
  if (!rc) {
    if (options->callback && options->callback->collectStatsIteration) {
      /* one sweep per replay: the callback sees every sweep, as in the reference */
      observer_env env = {h, options, problem, reaction, velocity, 0};
      p.sweeps_per_launch = 1;
      rc = fc3d_hip_solve_observed(h, &p, &s, observe_sweep, &env);
      if (!rc && env.failed) rc = NUMERICS_ERR_DEVICE;
    } else {
      rc = fc3d_hip_solve(h, &p, &s);
    }
  }
  double t2 = now_ms();
  if (!rc) rc = fc3d_hip_get_reaction(h, reaction);
  if (!rc && p.error_evaluation != SICONOS_NSGS_ERROR_EVALUATION_LIGHT)
    rc = fc3d_hip_get_velocity(h, velocity);
  double t3 = now_ms();
  if (h) fc3d_resident_release(h, token);
  if (rc) {
    fprintf(stderr, "fc3d_nsgs_hip: %s (%s)\n", numerics_error_string(rc), fc3d_hip_last_error());
    *info = rc;
    return;
  }
  *info = s.info;
  options->iparam[SICONOS_IPARAM_ITER_DONE] = s.iterations;
  options->dparam[SICONOS_DPARAM_RESIDU] = s.residual;
  options->iparam[SICONOS_FRICTION_3D_NSGS_HIP_NUM_COLOURS] = s.num_colours;
  options->iparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_SWEEPS] = s.fp32_sweeps;
  options->iparam[SICONOS_FRICTION_3D_NSGS_HIP_WARM_START_USED] = s.warm_start_used;
  options->dparam[SICONOS_DPARAM_TIME_BEFORE_LOOP] = t1 - t0;
  options->dparam[SICONOS_DPARAM_TIME_IN_LOOP] = t2 - t1;
  options->dparam[SICONOS_DPARAM_TIME_AFTER_LOOP] = t3 - t2;
  options->dparam[SICONOS_FRICTION_3D_NSGS_HIP_TIME_SOLVE] = s.solve_ms;
}

void fc3d_nsgs_hip_set_default(SolverOptions* o) {
  /* same NSGS-level defaults as fc3d_nsgs_set_default (fc3d_nsgs.c:919-950) */
  o->iparam[SICONOS_NSGS_ERROR_EVALUATION_TYPE] = SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL;
  o->iparam[SICONOS_FRICTION_3D_IPARAM_INTERNAL_ERROR_STRATEGY] =
      SICONOS_NSGS_INTERNAL_ERROR_STRATEGY_GIVEN_VALUE;
  o->iparam[SICONOS_NSGS_SHUFFLE] = 0;
  o->iparam[SICONOS_NSGS_SHUFFLE_SEED] = 0;
  o->iparam[SICONOS_NSGS_FREEZING_CONTACT] = 0;
  o->iparam[SICONOS_NSGS_FILTER_LOCAL_SOLUTION] = 0;
  o->iparam[SICONOS_NSGS_RELAXATION] = 0;
  o->iparam[SICONOS_NSGS_ERROR_EVALUATION_FREQUENCY] = 0;
  o->dparam[SICONOS_DPARAM_TOL] = 1e-4;
  o->dparam[SICONOS_NSGS_INTERNAL_ERROR_RATIO] = 10.0;
  if (o->numberOfInternalSolvers == 0) {
    o->numberOfInternalSolvers = 1;
    o->internalSolvers = (SolverOptions**)calloc(1, sizeof(SolverOptions*));
  } else {
    solver_options_delete(o->internalSolvers[0]);
  }
  o->internalSolvers[0] = solver_options_create(SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID);
  o->iparam[SICONOS_NSGS_PRINTING_STYLE] = 1;
  /* additions */
  o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_PRECISION] = 0;
  o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_WARM_START] = 0;
  o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_SCHEDULE] = 0;
  o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_SWEEPS_PER_LAUNCH] = 8;
  o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_DEVICE] = 0;
  o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_TOL] = 1e-5;
  o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_DEPTH] = 0;
  o->iparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_EVERY] = 2;
  o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_REG] = 1e-10;
  o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_LOCAL_SOLVER] = 0.0;
  o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_FP32_LOCAL_TOL] = 0.0;
  o->dparam[SICONOS_FRICTION_3D_NSGS_HIP_RESIDENCY] = 0.0;
}

static int nsgs_hip_solve_wrap(void* problem, double* reaction, double* velocity,
                               SolverOptions* options) {
  int info = NUMERICS_OK;
  fc3d_nsgs_hip((FrictionContactProblem*)problem, reaction, velocity, &info, options);
  return info;
}

FC3D_REGISTER_SOLVER(SICONOS_FRICTION_3D_NSGS_HIP, "FC3D_NSGS_HIP",
                     "Coloured non-smooth Gauss-Seidel for 3D friction contact on MI355X", NULL,
                     nsgs_hip_solve_wrap, NULL, NULL, fc3d_nsgs_hip_set_default, 1000, 1e-4, 0)

/* ---- local (one-contact) solvers: option descriptors, solved on the GPU -- */
/* Registered so that solver_options_create(id) yields the reference defaults
 * (onecontact...c:1017-1143, fc3d_projection.c:917-960). */
static void nsn_set_default(SolverOptions* o) {
  o->iparam[SICONOS_FRICTION_3D_NSN_RHO_STRATEGY] =
      SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM_COND;
  o->dparam[SICONOS_FRICTION_3D_NSN_RHO] = 1.0;
  o->iparam[SICONOS_FRICTION_3D_NSN_FORMULATION] = SICONOS_FRICTION_3D_NSN_FORMULATION_ALARTCURNIER_STD;
  o->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH] = SICONOS_FRICTION_3D_NSN_LINESEARCH_NO;
  o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY] = SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_PLI_NSN_LOOP;
  o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP] = 1;
  o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER] = 10;
}

static void nsn_gp_set_default(SolverOptions* o) {
  o->iparam[SICONOS_FRICTION_3D_NSN_RHO_STRATEGY] =
      SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM;
  o->dparam[SICONOS_FRICTION_3D_NSN_RHO] = 1.0;
  o->iparam[SICONOS_FRICTION_3D_NSN_FORMULATION] = SICONOS_FRICTION_3D_NSN_FORMULATION_ALARTCURNIER_STD;
  o->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH] = SICONOS_FRICTION_3D_NSN_LINESEARCH_GOLDSTEINPRICE;
  o->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH_MAX_ITER] = 10;
  o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY] =
      SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_NSN_AND_PLI_NSN_LOOP;
  o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP] = 1;
  o->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER] = 100;
}

static void no_default(SolverOptions* o) { (void)o; }

/* One-contact solve through the driver: problem has one contact. */
static int local_solve_wrap(void* problem, double* reaction, double* velocity,
                            SolverOptions* options) {
  FrictionContactProblem* p = (FrictionContactProblem*)problem;
  if (p->numberOfContacts != 1) return NUMERICS_ERR_INVALID_ARGUMENT;
  double W[9];
  if (p->M->storageType == NM_DENSE) {
    for (int k = 0; k < 9; k++) W[k] = p->M->matrix0[k];
  } else if (p->M->storageType == NM_SPARSE_BLOCK && p->M->matrix1->nbblocks == 1) {
    for (int k = 0; k < 9; k++) W[k] = p->M->matrix1->block[0][k];
  } else {
    return NUMERICS_ERR_NOT_IMPLEMENTED;
  }
  fc3d_hip_params hp;
  fc3d_hip_params_default(&hp);
  hp.local_solver = options->solverId;
  hp.local_max_iter = options->iparam[SICONOS_IPARAM_MAX_ITER];
  hp.local_tol = options->dparam[SICONOS_DPARAM_TOL];
  hp.local_formulation = options->iparam[SICONOS_FRICTION_3D_NSN_FORMULATION];
  hp.rho_strategy = options->iparam[SICONOS_FRICTION_3D_NSN_RHO_STRATEGY];
  hp.rho_constant = options->dparam[SICONOS_FRICTION_3D_NSN_RHO];
  hp.ls_max_iter = options->iparam[SICONOS_FRICTION_3D_NSN_LINESEARCH_MAX_ITER];
  hp.hybrid_strategy = options->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY];
  hp.hybrid_max_loop = options->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP];
  hp.hybrid_max_iter = options->iparam[SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER];
  double state[4];
  int info = 0;
  int rc = fc3d_hip_local_solve(1, W, p->q, p->mu, reaction, &hp, state, 1, &info);
  if (rc) return rc;
  for (int i = 0; i < 3; i++) {
    velocity[i] = p->q[i];
    for (int j = 0; j < 3; j++) velocity[i] += W[i + 3 * j] * reaction[j];
  }
  return info > 0 ? 1 : 0;
}

FC3D_REGISTER_SOLVER(SICONOS_FRICTION_3D_ONECONTACT_NSN, "ONECONE_NSN",
                     "One-contact nonsmooth Newton solver (GPU)", NULL, local_solve_wrap, NULL,
                     NULL, nsn_set_default, 100, 1e-14, 1)
FC3D_REGISTER_SOLVER(SICONOS_FRICTION_3D_ONECONTACT_NSN_GP, "ONECONE_NSN_GP",
                     "One-contact nonsmooth Newton with Goldstein-Price line search (GPU)", NULL,
                     local_solve_wrap, NULL, NULL, nsn_gp_set_default, 100, 1e-14, 1)
FC3D_REGISTER_SOLVER(SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID, "ONECONE_NSN_GP_HYBRID",
                     "Hybrid PLI-NSN one-contact solver (GPU)", NULL, local_solve_wrap, NULL,
                     NULL, nsn_gp_set_default, 100, 1e-14, 1)
FC3D_REGISTER_SOLVER(SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnCone, "OC_PROJ",
                     "Projection on cone (local solver, GPU)", NULL, local_solve_wrap, NULL, NULL,
                     no_default, 1000, 1e-14, 1)
FC3D_REGISTER_SOLVER(SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnConeWithLocalIteration,
                     "OC_PROJ_LI", "Projection on cone with local iteration (local solver, GPU)",
                     NULL, local_solve_wrap, NULL, NULL, no_default, 1000, 1e-14, 1)
