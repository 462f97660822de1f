// Conversion between an FP16-deviation state (storage.hpp) and FP32 compact fields, used by the
// domain's host copies and initialisers (never by the step, which reads and writes the stored
// format directly). A call converts fields [f0, f1); field f of the FP32 side starts at
// (f - f0) * xs floats. Also the scatter of a compact FP32 state into caller arrays of any layout
// (the launch-level initialisers).
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

// Compact FP32 fields (field f at src + f * n, cell (x, y, z) at x + nx (y + nyt z)) into the 11
// arrays of `dst` at cell x + y py + z pz (pointers already carry the layout origin).
__global__ void __launch_bounds__(kThreads) scatter_fields_kernel(const float *__restrict__ src, int64_t n,
                                                                  FieldPtrs dst, int nx, int nyt, uint32_t py,
                                                                  uint32_t pz) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int f = static_cast<int>(blockIdx.y);
    if (i >= n) return;
    const int x = static_cast<int>(i % nx);
    const int64_t r = i / nx;
    const int y = static_cast<int>(r % nyt);
    const int z = static_cast<int>(r / nyt);
    dst.p[f][static_cast<uint32_t>(x) + py * static_cast<uint32_t>(y) + pz * static_cast<uint32_t>(z)] = src[f * n + i];
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

hipError_t launch_scatter_fields(const float *src, const FieldPtrs &dst, int nx, int nyt, int nz, uint32_t py,
                                 uint32_t pz, hipStream_t s) {
    const int64_t n = static_cast<int64_t>(nx) * nyt * nz;
    if (n <= 0) return hipErrorInvalidValue;
    scatter_fields_kernel<<<grid_for(n, 11), kThreads, 0, s>>>(src, n, dst, nx, nyt, py, pz);
    return hipGetLastError();
}

}  // namespace cglbm
