// Conversion between an FP16-deviation state (storage.hpp) and 11 FP32 compact fields, used by
// the domain's host copies and initialisers (never by the step, which reads and writes the
// stored format directly). Field f of the FP32 side starts at f * fs floats.
#include "kernels.hpp"
#include "storage.hpp"

namespace cglbm {

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) pack_f16_kernel(const float *__restrict__ src, float *__restrict__ state,
                                                            int64_t fs, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    const float v = src[f * fs + i];
    if (f < 2) {
        state[f * fs + i] = v;
    } else {
        uint16_t *h = reinterpret_cast<uint16_t *>(state);
        h[(f + 2) * fs + i] = f16_encode(f, v);
    }
}

__global__ void __launch_bounds__(kThreads) unpack_f16_kernel(const float *__restrict__ state, float *__restrict__ dst,
                                                              int64_t fs, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    if (f < 2) {
        dst[f * fs + i] = state[f * fs + i];
    } else {
        const uint16_t *h = reinterpret_cast<const uint16_t *>(state);
        dst[f * fs + i] = f16_decode(f, h[(f + 2) * fs + i]);
    }
}

dim3 grid_for(int64_t n) { return dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads), 11, 1); }

}  // namespace

hipError_t launch_pack_f16(const float *src, float *state, int64_t fs, int64_t n, hipStream_t s) {
    if (n <= 0 || n > fs) return hipErrorInvalidValue;
    pack_f16_kernel<<<grid_for(n), kThreads, 0, s>>>(src, state, fs, n);
    return hipGetLastError();
}

hipError_t launch_unpack_f16(const float *state, float *dst, int64_t fs, int64_t n, hipStream_t s) {
    if (n <= 0 || n > fs) return hipErrorInvalidValue;
    unpack_f16_kernel<<<grid_for(n), kThreads, 0, s>>>(state, dst, fs, n);
    return hipGetLastError();
}

}  // namespace cglbm
