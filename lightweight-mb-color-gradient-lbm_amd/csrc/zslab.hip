// Device side of the halo exchange of a domain split into z-slabs: packing the planes a slab's two
// neighbours need into contiguous buffers, and writing received buffers into the halo planes.
//
// A slab's update reads planes z_lo-2 .. z_hi+1 (march_kernel.inl): the 11 moments of the planes
// next to it (z_lo-1, z_hi: sources of its first and last target planes) and rhor, rhob of the
// planes beyond (z_lo-2, z_hi+1: the two-hop colour gradient). So each face moves 13 plane-fields:
//   buffer[k], k = 0..10: field k of the plane next to the face; k = 11, 12: rhor, rhob of the plane
//   beyond it.
// The lower face sends planes z_lo (fields) and z_lo+1 (rhor, rhob), the upper face z_hi-1 and
// z_hi-2; received buffers land in z_lo-1, z_lo-2 and z_hi, z_hi+1. One launch handles both faces;
// cells are addressed through the layout, so any pitches work.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

namespace cglbm {
namespace {

struct HaloPlanes {
    float *f[11];
    uint32_t py, pz;
    int nx, nyt;
    int z_next[2], z_beyond[2];  // per face (0 = lower, 1 = upper)
    float *buf[2];
};

// One thread moves V consecutive x cells (V = 4 when rows, pitches and pointers allow 16-byte
// accesses, else 1); i indexes (face, plane-field, y, x / V).
template <bool PACK, int V>
__global__ void __launch_bounds__(256) zslab_halo_kernel(HaloPlanes h) {
    using Vec = typename std::conditional<V == 4, float4, float>::type;
    const uint32_t row = static_cast<uint32_t>(h.nx) / V;
    const uint32_t plane = row * static_cast<uint32_t>(h.nyt);
    const uint32_t total = 2u * 13u * plane;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const uint32_t face = i / (13u * plane);
        const uint32_t r = i - face * 13u * plane;
        const uint32_t k = r / plane;
        const uint32_t c = r - k * plane;
        const uint32_t y = c / row, xv = c - y * row;
        const int field = k < 11u ? static_cast<int>(k) : static_cast<int>(k) - 11;
        const int z = k < 11u ? h.z_next[face] : h.z_beyond[face];
        Vec *cell = reinterpret_cast<Vec *>(h.f[field] + V * xv + h.py * y + h.pz * static_cast<uint32_t>(z));
        Vec *slot = reinterpret_cast<Vec *>(h.buf[face]) + r;
        if constexpr (PACK) *slot = *cell;
        else *cell = *slot;
    }
}

template <bool PACK>
hipError_t launch_halo(const HaloPlanes &h, bool vec4, hipStream_t s) {
    const int v = vec4 ? 4 : 1;
    const int64_t total = 2ll * 13ll * (h.nx / v) * h.nyt;
    const int grid = static_cast<int>((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
    if (vec4) zslab_halo_kernel<PACK, 4><<<grid, 256, 0, s>>>(h);
    else zslab_halo_kernel<PACK, 1><<<grid, 256, 0, s>>>(h);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_zslab_halo(const StepArgs &a, int z_lo, int z_hi, float *lo, float *hi, bool pack, hipStream_t s) {
    HaloPlanes h{};
    for (int f = 0; f < 11; ++f) h.f[f] = const_cast<float *>(a.A[f]);
    h.py = a.py;
    h.pz = a.pz;
    h.nx = a.nx;
    h.nyt = a.nyt;
    if (pack) {
        h.z_next[0] = z_lo;
        h.z_beyond[0] = z_lo + 1;
        h.z_next[1] = z_hi - 1;
        h.z_beyond[1] = z_hi - 2;
    } else {
        h.z_next[0] = z_lo - 1;
        h.z_beyond[0] = z_lo - 2;
        h.z_next[1] = z_hi;
        h.z_beyond[1] = z_hi + 1;
    }
    h.buf[0] = lo;
    h.buf[1] = hi;
    bool vec4 = a.nx % 4 == 0 && a.py % 4 == 0 && a.pz % 4 == 0;
    for (int f = 0; f < 11; ++f) vec4 = vec4 && reinterpret_cast<uintptr_t>(h.f[f]) % 16 == 0;
    vec4 = vec4 && reinterpret_cast<uintptr_t>(lo) % 16 == 0 && reinterpret_cast<uintptr_t>(hi) % 16 == 0;
    return pack ? launch_halo<true>(h, vec4, s) : launch_halo<false>(h, vec4, s);
}

}  // namespace cglbm
