// fc3d_mlp.hpp — shape of the warm-start network and the device-side entry
// points the solver uses (fc3d_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fc3d_hip.h"

namespace fc3d {

constexpr int MLP_NF = FC3D_MLP_FEATURES;  // raw features per contact (22)
constexpr int MLP_K1 = 32;                 // features padded to two MFMA K-steps
constexpr int MLP_H = FC3D_MLP_HIDDEN;     // hidden width (128)
constexpr int MLP_NO = FC3D_MLP_OUTPUTS;   // outputs (3)
constexpr int MLP_NOP = 16;                // outputs padded to one MFMA tile
constexpr int MLP_NPARAMS = MLP_H * MLP_K1 + MLP_H + 2 * (MLP_H * MLP_H + MLP_H) +
                            MLP_NOP * MLP_H + MLP_NOP;
static_assert(MLP_NPARAMS == FC3D_MLP_NPARAMS, "parameter count");

// y[n][3] = MLP(feat[n][22]) on the device, on stream s.  X is a scratch
// buffer of (n rounded up to 64) x 32 bf16 values.
int mlp_forward_device(fc3d_hip_mlp* m, int n, const float* feat, unsigned short* X, float* y,
                       hipStream_t s);
int mlp_device(const fc3d_hip_mlp* m);

}  // namespace fc3d
