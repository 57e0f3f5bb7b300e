// fc3d_runtime.hip — error reporting and device selection for the C-ABI.
#include "fc3d_internal.hpp"

namespace fc3d {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
  va_end(ap);
}

int use_device(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    set_error("no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    return NUMERICS_ERR_DEVICE;
  }
  if (device < 0 || device >= n) {
    set_error("device ordinal %d out of range (%d devices)", device, n);
    return NUMERICS_ERR_INVALID_ARGUMENT;
  }
  FC3D_HIP_TRY(hipSetDevice(device));
  return 0;
}

}  // namespace fc3d

extern "C" {

int fc3d_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* fc3d_hip_last_error(void) { return fc3d::g_last_error; }

void fc3d_hip_params_default(fc3d_hip_params* p) {
  p->max_iter = 1000;
  p->tol = 1e-4;
  p->error_evaluation = SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL;
  p->error_frequency = 0;
  p->filter_local_solution = 0;
  p->relaxation = 0;
  p->omega = 0.0;
  p->local_solver = SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID;
  p->local_max_iter = 100;
  p->local_tol = 1e-14;
  p->local_formulation = SICONOS_FRICTION_3D_NSN_FORMULATION_ALARTCURNIER_STD;
  p->rho_strategy = SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM;
  p->rho_constant = 1.0;
  p->ls_max_iter = 10;
  p->hybrid_strategy = SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_NSN_AND_PLI_NSN_LOOP;
  p->hybrid_max_loop = 1;
  p->hybrid_max_iter = 100;
  p->precision = 0;
  p->warm_start = 0;
  p->schedule = 0;
  p->sweeps_per_launch = 8;
  p->fp32_tol = 1e-5;
  p->fp32_local_tol = 1e-6;
  p->instrument = 0;
  p->accel_depth = 0;
  p->accel_every = 2;
  p->accel_reg = 1e-10;
  p->velocity_form = 0;
  p->fp32_local_solver = 0;
  p->freeze_contacts = 0;
  p->focus_passes = 0;
  p->focus_every = 16;
  p->focus_after = 64;
  p->focus_count = 256;
  p->focus_share = 0.5;
  p->start_trial_sweeps = 1;
  p->focus_omega = 1.2;
}

}  // extern "C"
