// z-marching fused step (placeholder until the tuned kernel lands).
#include "kernels.hpp"

namespace cglbm {

bool march_supports(const StepArgs &, int) { return false; }

hipError_t launch_march(const StepArgs &, int, hipStream_t) { return hipErrorInvalidValue; }

}  // namespace cglbm
