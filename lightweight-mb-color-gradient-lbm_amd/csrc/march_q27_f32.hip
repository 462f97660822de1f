// D3Q27 instantiations of the z-marching step, FP32 storage (march_kernel.inl).
#include "march_kernel.inl"

namespace cglbm {
hipError_t launch_march_q27_f32(const MarchArgs &a, MarchShape m, hipStream_t s) {
    return march::launch_march_qs<27, kStoreF32>(a, m, s);
}
hipError_t launch_batch_q27_f32(const BatchLaunch &L, hipStream_t s) { return march::launch_batch_qs<27, kStoreF32>(L, s); }
}  // namespace cglbm
