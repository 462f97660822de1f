// D3Q19 instantiations of the z-marching step for FP32 states addressed through one 64-bit pointer
// per field (fields in separate allocations, or a state of 4 GiB or more; march_kernel.inl).
// Single domains only.
#include "march_kernel.inl"

namespace cglbm {
hipError_t launch_march_q19_f32ptr(const MarchArgs &a, MarchShape m, hipStream_t s) {
    return march::launch_march_qs<19, kStoreF32Ptr>(a, m, s);
}
}  // namespace cglbm
