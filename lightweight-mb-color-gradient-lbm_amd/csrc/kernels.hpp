// Host-side launchers of the device kernels (internal C++ interface of the library).
#pragma once

#include <hip/hip_runtime.h>

#include "step_common.hpp"

namespace cglbm {

// Direct-gather step (any layout).
hipError_t launch_gather(const StepArgs &a, int q, hipStream_t s);
// Boundary planes of B from B and Aold (jet inlet/outlet or wall copies); no-op when periodic.
hipError_t launch_ybc(const BcArgs &a, hipStream_t s);

// z-marching step (compact layout, fused boundary planes). Returns hipErrorInvalidValue when the
// shape is outside what the kernel supports (caller falls back to the gather path).
hipError_t launch_march(const StepArgs &a, int q, hipStream_t s);
bool march_supports(const StepArgs &a, int q);

// FP64 diagnostics: partial sums per block into `partials` (capacity kReduceBlocks * 6), then the
// fixed-order final sum into `result` (6 doubles).
inline constexpr int kReduceBlocks = 512;
hipError_t launch_reduce(const float *const *A, int nx, int ny, int nz, int y0, uint32_t py, uint32_t pz,
                         double *partials, double *result, hipStream_t s);

}  // namespace cglbm
