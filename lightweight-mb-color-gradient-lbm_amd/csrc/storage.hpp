// Moment storage formats.
//
// FP32 (88 B per update): the 11 fields as binary32, as the reference stores them (memory.cuh:14-29).
//
// FP16 deviations (52 B per update): rhor and rhob stay binary32. They carry the phase field and
// the conserved per-phase masses. The nine other moments are stored as binary16 deviations from
// their rest values (zero), scaled by a power of two:
//     ux, uy, uz                          h = half(u  * 2^10)  (|u|  < 64 representable)
//     Pixx Pixy Piyy Piyz Pizz Pixz       h = half(Pi * 2^12)  (|Pi| < 16 representable)
// Encoding rounds to nearest even. Decoding is exact: half -> float, times the inverse power of
// two. All arithmetic stays binary32. A step is therefore the FP32 step applied to the decoded
// state, followed by the rounding of those nine fields. The scales keep the typical magnitudes
// (|u| ~ 1e-5..1e-1, |Pi| ~ 1e-7..1e-2) in the normal binary16 range; beyond the limits above, a
// field saturates to +-inf and the diagnostics count the cell as non-finite.
//
// Layout of one FP16 state with field stride fs (cells, a multiple of 64):
//     byte 0          rhor   fs x binary32
//     byte 4 fs       rhob   fs x binary32
//     byte 2 fs (f+2) field f = 2..10, fs x binary16
// for 26 fs bytes in total.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cglbm {

enum : int { kStoreF32 = 0, kStoreF16 = 1 };
// FP32 storage addressed through one 64-bit pointer per field: fields in separate allocations (one
// array per moment, as the reference allocates them), or a single-domain state of 4 GiB or more,
// which the 32-bit buffer offsets of the marching kernel cannot span. Internal to the kernels.
enum : int { kStoreF32Ptr = 2 };

__host__ __device__ constexpr float f16_scale(int f) { return f < 5 ? 1024.0f : 4096.0f; }
__host__ __device__ constexpr float f16_inv_scale(int f) { return f < 5 ? (1.0f / 1024.0f) : (1.0f / 4096.0f); }

// Bytes of one state in the given format for field stride fs (cells).
__host__ __device__ constexpr int64_t state_bytes(int store, int64_t fs) {
    return store == kStoreF16 ? 26 * fs : 44 * fs;
}
// Byte offset of field f within a state.
__host__ __device__ constexpr int64_t field_offset(int store, int f, int64_t fs) {
    return (store == kStoreF16 && f >= 2) ? 2 * fs * (f + 2) : 4 * fs * f;
}
__host__ __device__ constexpr int field_bytes(int store, int f) { return (store == kStoreF16 && f >= 2) ? 2 : 4; }

__device__ __forceinline__ uint16_t f16_encode(int f, float v) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v * f16_scale(f)));
}
// The product of the exact conversion and a power of two is exact (binary16 values times 2^-10 or
// 2^-12 stay normal binary32), so one fused multiply-add with -0 (which keeps the product's sign of
// zero) gives the same bits as the conversion followed by the multiplication, in one mixed-precision
// instruction.
__device__ __forceinline__ float f16_decode(int f, uint16_t h) {
    return __builtin_fmaf(static_cast<float>(__builtin_bit_cast(_Float16, h)), f16_inv_scale(f), -0.0f);
}

}  // namespace cglbm
