// Conversion between an FP16-deviation state (storage.hpp) and FP32 compact fields, used by the
// domain's host copies and initialisers (never by the step, which reads and writes the stored
// format directly). A call converts fields [f0, f1); field f of the FP32 side starts at
// (f - f0) * xs floats.
#include "kernels.hpp"
#include "storage.hpp"

namespace cglbm {

namespace {

constexpr int kThreads = 256;

// Fields f0 + blockIdx.y; FP32 side field f at src/dst + (f - f0) * xs.
__global__ void __launch_bounds__(kThreads) pack_f16_kernel(const float *__restrict__ src, int64_t xs,
                                                            float *__restrict__ state, int64_t fs, int64_t n, int f0) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int f = f0 + static_cast<int>(blockIdx.y);
    if (i >= n) return;
    const float v = src[(f - f0) * xs + i];
    if (f < 2) {
        state[f * fs + i] = v;
    } else {
        uint16_t *h = reinterpret_cast<uint16_t *>(state);
        h[(f + 2) * fs + i] = f16_encode(f, v);
    }
}

__global__ void __launch_bounds__(kThreads) unpack_f16_kernel(const float *__restrict__ state, int64_t fs,
                                                              float *__restrict__ dst, int64_t xs, int64_t n, int f0) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int f = f0 + static_cast<int>(blockIdx.y);
    if (i >= n) return;
    float v;
    if (f < 2) {
        v = state[f * fs + i];
    } else {
        const uint16_t *h = reinterpret_cast<const uint16_t *>(state);
        v = f16_decode(f, h[(f + 2) * fs + i]);
    }
    dst[(f - f0) * xs + i] = v;
}

bool bad_range(int64_t n, int64_t fs, int64_t xs, int f0, int f1) {
    return n <= 0 || n > fs || n > xs || f0 < 0 || f1 > 11 || f0 >= f1;
}

dim3 grid_for(int64_t n, int nf) { return dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads), nf, 1); }

}  // namespace

hipError_t launch_pack_f16(const float *src, int64_t xs, float *state, int64_t fs, int64_t n, int f0, int f1,
                           hipStream_t s) {
    if (bad_range(n, fs, xs, f0, f1)) return hipErrorInvalidValue;
    pack_f16_kernel<<<grid_for(n, f1 - f0), kThreads, 0, s>>>(src, xs, state, fs, n, f0);
    return hipGetLastError();
}

hipError_t launch_unpack_f16(const float *state, int64_t fs, float *dst, int64_t xs, int64_t n, int f0, int f1,
                             hipStream_t s) {
    if (bad_range(n, fs, xs, f0, f1)) return hipErrorInvalidValue;
    unpack_f16_kernel<<<grid_for(n, f1 - f0), kThreads, 0, s>>>(state, fs, dst, xs, n, f0);
    return hipGetLastError();
}

}  // namespace cglbm
