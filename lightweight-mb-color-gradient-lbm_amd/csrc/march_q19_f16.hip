// D3Q19 instantiations of the z-marching step, FP16-deviation storage (march_kernel.inl).
#include "march_kernel.inl"

namespace cglbm {
hipError_t launch_march_q19_f16(const MarchArgs &a, MarchShape m, hipStream_t s) {
    return march::launch_march_qs<19, kStoreF16>(a, m, s);
}
hipError_t launch_batch_q19_f16(const BatchLaunch &L, hipStream_t s) { return march::launch_batch_qs<19, kStoreF16>(L, s); }
}  // namespace cglbm
