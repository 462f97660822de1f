// D3Q19 instantiations of the z-marching step (march_kernel.inl).
#include "march_kernel.inl"

namespace cglbm {
hipError_t launch_march_q19(const MarchArgs &a, MarchShape m, hipStream_t s) { return march::launch_march_q<19>(a, m, s); }
hipError_t launch_batch_q19(const BatchLaunch &L, hipStream_t s) { return march::launch_batch_q<19>(L, s); }
}  // namespace cglbm
