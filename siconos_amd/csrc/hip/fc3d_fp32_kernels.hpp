// fc3d_fp32_kernels.hpp — the FP32 kernels of the velocity form and of the
// focused passes, which are compiled in their own translation unit
// (fc3d_fp32.hip, built with -ffp-contract=fast so that each a * b + c of
// their arithmetic is one fused multiply-add).  fc3d_problem.hip is built
// with -ffp-contract=off, which keeps the FP64 sweeps bit for bit equal to
// the reference's C; it declares these instantiations extern and launches
// them from there.  The FP32 sweeps are refined and judged in FP64, so the
// rounding of their products is not part of any result.
#pragma once
#include "fc3d_internal.hpp"
#include "fc3d_focus.hpp"
#include "fc3d_velocity.hpp"

#define FC3D_FP32_KERNELS(SPEC)                                                                  \
  SPEC template __global__ void fc3d::k_sweep_vel<float, false>(                                 \
      fc3d::SweepPtrs<float>, fc3d::BodyPtrs<float>, const fc3d::DevParams*,                     \
      const fc3d::SolveState*, int, int, int);                                                   \
  SPEC template __global__ void fc3d::k_sweep_vel<float, true>(                                  \
      fc3d::SweepPtrs<float>, fc3d::BodyPtrs<float>, const fc3d::DevParams*,                     \
      const fc3d::SolveState*, int, int, int);                                                   \
  SPEC template __global__ void fc3d::k_sweep_vel_tail<float, false>(                            \
      fc3d::SweepPtrs<float>, fc3d::BodyPtrs<float>, const fc3d::DevParams*, fc3d::SolveState*,  \
      const int*, int, int, int, int, int);                                                      \
  SPEC template __global__ void fc3d::k_sweep_vel_tail<float, true>(                             \
      fc3d::SweepPtrs<float>, fc3d::BodyPtrs<float>, const fc3d::DevParams*, fc3d::SolveState*,  \
      const int*, int, int, int, int, int);                                                      \
  SPEC template __global__ void fc3d::k_focus_passes_staged<float>(                              \
      fc3d::FocusStage<float>, fc3d::BodyPtrs<float>, const fc3d::DevParams*,                    \
      const fc3d::SolveState*, fc3d::FocusState*, const int*, int, int, int, int, int);          \
  SPEC template __global__ void fc3d::k_focus_passes<float, false>(                              \
      fc3d::SweepPtrs<float>, fc3d::BodyPtrs<float>, const fc3d::DevParams*,                     \
      const fc3d::SolveState*, fc3d::FocusState*, const int*, const int*, int, int, int, int);   \
  SPEC template __global__ void fc3d::k_focus_passes<float, true>(                               \
      fc3d::SweepPtrs<float>, fc3d::BodyPtrs<float>, const fc3d::DevParams*,                     \
      const fc3d::SolveState*, fc3d::FocusState*, const int*, const int*, int, int, int, int);
