// D3Q27 instantiations of the z-marching step for FP32 states of 4 GiB or more (64-bit
// addressing, march_kernel.inl). Single domains only.
#include "march_kernel.inl"

namespace cglbm {
hipError_t launch_march_q27_f32big(const MarchArgs &a, MarchShape m, hipStream_t s) {
    return march::launch_march_qs<27, kStoreF32Big>(a, m, s);
}
}  // namespace cglbm
