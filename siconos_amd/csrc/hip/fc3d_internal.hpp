// fc3d_internal.hpp — shared host-side helpers of the HIP library (error
// reporting, checked HIP calls, device selection).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "fc3d_hip.h"

namespace fc3d {

void set_error(const char* fmt, ...);

// Selects `device` for the calling thread; fails if no device is usable.
int use_device(int device);

}  // namespace fc3d

#define FC3D_HIP_TRY(expr)                                                              \
  do {                                                                                  \
    hipError_t fc3d_e_ = (expr);                                                        \
    if (fc3d_e_ != hipSuccess) {                                                        \
      fc3d::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,                      \
                      hipGetErrorString(fc3d_e_));                                      \
      return NUMERICS_ERR_DEVICE;                                                       \
    }                                                                                   \
  } while (0)

#define FC3D_WAVE 64
