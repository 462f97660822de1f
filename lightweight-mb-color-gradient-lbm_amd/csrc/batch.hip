// Device side of batches: the step-base counter advanced after each bundle of steps.
#include "kernels.hpp"

namespace cglbm {

namespace {
__global__ void advance_kernel(int64_t *base, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) base[0] += n;
}
}  // namespace

hipError_t launch_advance(int64_t *base, int n, hipStream_t s) {
    advance_kernel<<<1, 64, 0, s>>>(base, n);
    return hipGetLastError();
}

}  // namespace cglbm
