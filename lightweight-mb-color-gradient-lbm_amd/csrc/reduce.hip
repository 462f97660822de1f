// FP64 per-phase mass and momentum diagnostics (the reference has none; lbm.cuh stores the
// moments these sum). Deterministic: every cell goes to a fixed block and lane, each lane sums
// its cells in a fixed order in double, blocks combine with a fixed tree, and one block adds the
// block partials in index order. The result therefore does not depend on timing or on what else
// runs on the GPU.
#include "kernels.hpp"
#include "storage.hpp"

namespace cglbm {

namespace {

constexpr int kThreads = 256;

// Field f (2..10) at cell id: FP32, or a scaled binary16 (storage.hpp).
template <int ST>
__device__ __forceinline__ float load_field(const void *p, int f, uint32_t id) {
    if constexpr (ST == kStoreF16) return f16_decode(f, static_cast<const uint16_t *>(p)[id]);
    else return static_cast<const float *>(p)[id];
}

// The 11 fields of one state (FP32 arrays, or for FP16 storage the binary16 fields 2..10).
struct ReduceFields {
    const void *p[11];
};

// A cell counts as bad when a phase density is negative or not finite, the density is not
// positive, or any velocity or Pi component is not finite (a saturated binary16 field decodes to
// +-inf, so FP16 overflow is counted too).
template <int ST>
__global__ void __launch_bounds__(kThreads) reduce_partials_kernel(ReduceFields F, int nx, int ny, int nz, int y0,
                                                                   uint32_t py, uint32_t pz, double *partials) {
    __shared__ double sh[6][kThreads];
    const float *rr = static_cast<const float *>(F.p[0]);
    const float *rb = static_cast<const float *>(F.p[1]);
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    const int64_t nb = static_cast<int64_t>(gridDim.x);
    const int64_t per_block = (n + nb - 1) / nb;
    const int64_t lo = per_block * blockIdx.x;
    const int64_t hi = lo + per_block < n ? lo + per_block : n;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t c = lo + threadIdx.x; c < hi; c += kThreads) {
        const int x = static_cast<int>(c % nx);
        const int64_t r = c / nx;
        const int y = static_cast<int>(r % ny) + y0;
        const int z = static_cast<int>(r / ny);
        const uint32_t id = static_cast<uint32_t>(x) + py * static_cast<uint32_t>(y) + pz * static_cast<uint32_t>(z);
        const double a = rr[id], b = rb[id];
        const double rho = a + b;
        const double ux = load_field<ST>(F.p[2], 2, id), uy = load_field<ST>(F.p[3], 3, id),
                     uz = load_field<ST>(F.p[4], 4, id);
        s[0] += a;
        s[1] += b;
        s[2] += rho * ux;
        s[3] += rho * uy;
        s[4] += rho * uz;
        bool pi_ok = true;
#pragma unroll
        for (int f = 5; f < 11; ++f) pi_ok = pi_ok && isfinite(load_field<ST>(F.p[f], f, id));
        const bool bad = !(a >= 0.0) || !(b >= 0.0) || !(rho > 0.0) || !isfinite(ux) || !isfinite(uy) || !isfinite(uz) ||
                         !pi_ok;
        s[5] += bad ? 1.0 : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
#pragma unroll
            for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partials[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// One block: lane l sums partials l, l + 256, ... in order, then a fixed tree combines lanes.
__global__ void __launch_bounds__(kThreads) reduce_final_kernel(const double *partials, int nblocks, double *result) {
    __shared__ double sh[6][kThreads];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += kThreads) {
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] += partials[b * 6 + k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
#pragma unroll
            for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) result[threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace

hipError_t launch_reduce(const float *const *A, int nx, int ny, int nz, int y0, uint32_t py, uint32_t pz,
                         double *partials, double *result, hipStream_t s) {
    ReduceFields F;
    for (int f = 0; f < 11; ++f) F.p[f] = A[f];
    reduce_partials_kernel<kStoreF32><<<kReduceBlocks, kThreads, 0, s>>>(F, nx, ny, nz, y0, py, pz, partials);
    reduce_final_kernel<<<1, kThreads, 0, s>>>(partials, kReduceBlocks, result);
    return hipGetLastError();
}

hipError_t launch_reduce_f16(const float *state, int64_t fs, int nx, int ny, int nz, int y0, uint32_t py, uint32_t pz,
                             double *partials, double *result, hipStream_t s) {
    const char *b = reinterpret_cast<const char *>(state);
    ReduceFields F;
    for (int f = 0; f < 11; ++f) F.p[f] = b + field_offset(kStoreF16, f, fs);
    reduce_partials_kernel<kStoreF16><<<kReduceBlocks, kThreads, 0, s>>>(F, nx, ny, nz, y0, py, pz, partials);
    reduce_final_kernel<<<1, kThreads, 0, s>>>(partials, kReduceBlocks, result);
    return hipGetLastError();
}

}  // namespace cglbm
