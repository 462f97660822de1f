// Lattice tables and binary32 constants, usable in host and gfx950 device code.
//
// Direction order, weights, B_i, Hermite values and 1/|c| follow stencil_ct.cuh:16-216 (D3Q27)
// and 227-417 (D3Q19); D3Q19 is the first 19 directions of the D3Q27 order. The order is part
// of the result (populations are accumulated i = 0..Q-1, lbm.cuh:73-237), so it is kept. The
// constants are formed with the same binary32 expressions as constants.cuh:56-93; constexpr
// evaluation rounds to binary32 exactly as the reference's compile-time evaluation does.
#pragma once

#include <cstdint>

#ifndef __HIP_DEVICE_COMPILE__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif

namespace cglbm {

inline constexpr int kCx[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
inline constexpr int kCy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
inline constexpr int kCz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

__host__ __device__ constexpr int norm2(int i) { return kCx[i] * kCx[i] + kCy[i] * kCy[i] + kCz[i] * kCz[i]; }

// Opposite direction: the tables list directions in +/- pairs (1,2), (3,4), ...
__host__ __device__ constexpr int opposite(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

inline constexpr float kCs2 = static_cast<float>(1.0 / 3.0);
inline constexpr float kCs4 = kCs2 * kCs2;
inline constexpr float kCs6 = kCs4 * kCs2;
inline constexpr float kInvCs2 = 1.0f / kCs2;
inline constexpr float kInvCs4 = 1.0f / kCs4;
inline constexpr float kInvCs6 = 1.0f / kCs6;
inline constexpr float kInv2Cs2 = 1.0f / (2.0f * kCs2);
inline constexpr float kInv2Cs4 = 1.0f / (2.0f * kCs4);
inline constexpr float kInv6Cs6 = 1.0f / (6.0f * kCs6);
inline constexpr float kInv2Cs6 = 1.0f / (2.0f * kCs6);

template <int Q>
__host__ __device__ constexpr float weight(int i) {
    const int s = norm2(i);
    if constexpr (Q == 27)
        return s == 0 ? static_cast<float>(8.0 / 27.0)
                      : s == 1 ? static_cast<float>(2.0 / 27.0)
                               : s == 2 ? static_cast<float>(1.0 / 54.0) : static_cast<float>(1.0 / 216.0);
    else
        return s == 0 ? static_cast<float>(1.0 / 3.0)
                      : s == 1 ? static_cast<float>(1.0 / 18.0) : static_cast<float>(1.0 / 36.0);
}

template <int Q>
__host__ __device__ constexpr float bcoef(int i) {
    if (i == 0) return Q == 27 ? static_cast<float>(-10.0 / 27.0) : static_cast<float>(-1.0 / 3.0);
    return weight<Q>(i);
}

__host__ __device__ constexpr float invcnorm(int i) {
    const int s = norm2(i);
    return s == 0 ? 0.0f : s == 1 ? 1.0f : s == 2 ? static_cast<float>(0.7071067811865475)
                                                   : static_cast<float>(0.5773502691896258);
}

// Second- and third-order Hermite values, binary32 (stencil_ct.cuh:96-168).
__host__ __device__ constexpr float hxx(int i) { return static_cast<float>(kCx[i] * kCx[i]) - kCs2; }
__host__ __device__ constexpr float hxy(int i) { return static_cast<float>(kCx[i] * kCy[i]); }
__host__ __device__ constexpr float hyy(int i) { return static_cast<float>(kCy[i] * kCy[i]) - kCs2; }
__host__ __device__ constexpr float hyz(int i) { return static_cast<float>(kCy[i] * kCz[i]); }
__host__ __device__ constexpr float hzz(int i) { return static_cast<float>(kCz[i] * kCz[i]) - kCs2; }
__host__ __device__ constexpr float hxz(int i) { return static_cast<float>(kCx[i] * kCz[i]); }
__host__ __device__ constexpr float hxxy(int i) { return static_cast<float>(kCx[i] * kCx[i] * kCy[i]) - kCs2 * static_cast<float>(kCy[i]); }
__host__ __device__ constexpr float hxxz(int i) { return static_cast<float>(kCx[i] * kCx[i] * kCz[i]) - kCs2 * static_cast<float>(kCz[i]); }
__host__ __device__ constexpr float hxyy(int i) { return static_cast<float>(kCx[i] * kCy[i] * kCy[i]) - kCs2 * static_cast<float>(kCx[i]); }
__host__ __device__ constexpr float hxzz(int i) { return static_cast<float>(kCx[i] * kCz[i] * kCz[i]) - kCs2 * static_cast<float>(kCx[i]); }
__host__ __device__ constexpr float hyyz(int i) { return static_cast<float>(kCy[i] * kCy[i] * kCz[i]) - kCs2 * static_cast<float>(kCz[i]); }
__host__ __device__ constexpr float hyzz(int i) { return static_cast<float>(kCy[i] * kCz[i] * kCz[i]) - kCs2 * static_cast<float>(kCy[i]); }
__host__ __device__ constexpr float hxyz(int i) { return static_cast<float>(kCx[i] * kCy[i] * kCz[i]); }

// Per-domain binary32 coefficients, formed on the host with the reference's expressions
// (constants.cuh:78-93, forces.cuh:114-117,132).
struct Coeffs {
    float oms;                  // 1 - 1/taub
    float Acoef;                // (sigma / (4 cs4)) / taub
    float beta;                 // beta_recolor
    float one_minus_half_omega; // 1 - omega/2 (body force)
    float grav;                 // gravity magnitude along -y
    float atwood;
};

inline Coeffs make_coeffs(float tau, float sigma, float beta, float gravity, float atwood) {
    Coeffs c{};
    const float omega = 1.0f / tau;
    c.oms = 1.0f - omega;
    const float s4 = sigma / (4.0f * kCs4);
    c.Acoef = s4 / tau;
    c.beta = beta;
    c.one_minus_half_omega = 1.0f - 0.5f * omega;
    c.grav = gravity;
    c.atwood = atwood;
    return c;
}

// Counter-based uniform in [0,1) keyed by (seed, x, z, step, component); the same integer
// operations run on host and device (inlet perturbation of the jet).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline float inlet_uniform(uint32_t seed, int x, int z, int64_t step, int comp) {
    const uint64_t k = (static_cast<uint64_t>(seed) << 32) ^
                       (static_cast<uint64_t>(static_cast<uint32_t>(x)) * 0x9E3779B97F4A7C15ull) ^
                       (static_cast<uint64_t>(static_cast<uint32_t>(z)) * 0xC2B2AE3D27D4EB4Full) ^
                       (static_cast<uint64_t>(step) * 0x165667B19E3779F9ull) ^ static_cast<uint64_t>(comp);
    return static_cast<float>(mix64(k) >> 40) * (1.0f / 16777216.0f);
}

}  // namespace cglbm
