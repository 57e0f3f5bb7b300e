// fc3d_fp32.hip — the FP32 velocity-form and focused-pass kernels, built with
// -ffp-contract=fast (the list and the reason: fc3d_fp32_kernels.hpp).
#include "fc3d_fp32_kernels.hpp"

FC3D_FP32_KERNELS()
