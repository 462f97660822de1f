// Shared device-side definitions of the fused colour-gradient step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "lattice.hpp"

namespace cglbm {

enum : int { kYPeriodic = 0, kYJet = 1, kYWalls = 2 };

// Kernel arguments of one update. Pointers already include the layout origin; cell (x, y, z)
// (y = stored index) is at x + y*py + z*pz.
struct StepArgs {
    const float *A[11];
    float *B[11];
    int nx, ny, nz, nyt;
    uint32_t py, pz;
    int ybc, outlet, conservative, forced;
    Coeffs c;
    float jet_x0, jet_z0, jet_r, jet_u, jet_noise;
    uint32_t seed;
    int64_t step;
};

struct BcArgs {
    float *B[11];
    const float *Aold[11];
    int nx, ny, nz, nyt;
    uint32_t py, pz;
    int ybc, outlet;
    float jet_x0, jet_z0, jet_r, jet_u, jet_noise;
    uint32_t seed;
    int64_t step;
};

template <int I, int N, typename F>
__device__ __forceinline__ void unroll(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        unroll<I + 1, N>(f);
    }
}

// feq (pdf.cuh:16-36), written out in full so every rounding step matches.
template <int Q, int I>
__device__ __forceinline__ float feq_full(float rho, float ux, float uy, float uz) {
    constexpr float cx = static_cast<float>(kCx[I]), cy = static_cast<float>(kCy[I]), cz = static_cast<float>(kCz[I]);
    constexpr float wi = weight<Q>(I);
    const float cu = ux * cx + uy * cy + uz * cz;
    const float usq = ux * ux + uy * uy + uz * uz;
    const float cu2 = cu * cu;
    const float A2eq = (cu * kInvCs2) - (usq * kInv2Cs2) + (cu * cu) * kInv2Cs4;
    const float A3eq = cu * (cu2 * kInv6Cs6 - usq * kInv2Cs4);
    return wi * rho * (1.0f + A2eq + A3eq);
}

// fneqr (pdf.cuh:38-79).
template <int Q, int I>
__device__ __forceinline__ float fneqr_full(float Pxx, float Pxy, float Pyy, float Pyz, float Pzz, float Pxz,
                                            float ux, float uy, float uz) {
    constexpr float wi = weight<Q>(I);
    const float A2neq = (Pxx * hxx(I) + 2.0f * Pxy * hxy(I) + Pyy * hyy(I) + 2.0f * Pyz * hyz(I) + Pzz * hzz(I) +
                         2.0f * Pxz * hxz(I)) * kInv2Cs4;
    const float a3xxy = Pxx * uy + 2.0f * Pxy * ux;
    const float a3xxz = Pxx * uz + 2.0f * Pxz * ux;
    const float a3xyy = Pyy * ux + 2.0f * Pxy * uy;
    const float a3xzz = Pzz * ux + 2.0f * Pxz * uz;
    const float a3yyz = Pyy * uz + 2.0f * Pyz * uy;
    const float a3yzz = Pzz * uy + 2.0f * Pyz * uz;
    const float a3xyz = Pxy * uz + Pxz * uy + Pyz * ux;
    const float A3neq = (a3xxy * hxxy(I) + a3xxz * hxxz(I) + a3xyy * hxyy(I) + a3xzz * hxzz(I) + a3yyz * hyyz(I) +
                         a3yzz * hyzz(I) + 2.0f * a3xyz * hxyz(I)) * kInv2Cs6;
    return wi * (A2neq + A3neq);
}

// guo_force (pdf.cuh:94-136), dormant in the reference; used by the body-force option.
template <int Q, int I>
__device__ __forceinline__ float guo_full(float Fx, float Fy, float Fz, float ux, float uy, float uz) {
    constexpr float wi = weight<Q>(I);
    constexpr float cx = static_cast<float>(kCx[I]), cy = static_cast<float>(kCy[I]), cz = static_cast<float>(kCz[I]);
    const float h = 0.5f;
    const float A1 = (Fx * cx + Fy * cy + Fz * cz) * kInvCs2;
    const float A2 = (Fx * ux * hxx(I) + Fy * uy * hyy(I) + Fz * uz * hzz(I) + (Fx * uy + Fy * ux) * hxy(I) +
                      (Fy * uz + Fz * uy) * hyz(I) + (Fx * uz + Fz * ux) * hxz(I)) * kInvCs4;
    const float A3 = ((Fx * ux * uy + h * Fy * ux * ux) * hxxy(I) + (Fx * ux * uz + h * Fz * ux * ux) * hxxz(I) +
                      (Fy * ux * uy + h * Fx * uy * uy) * hxyy(I) + (Fz * ux * uz + h * Fx * uz * uz) * hxzz(I) +
                      (Fy * uy * uz + h * Fz * uy * uy) * hyyz(I) + (Fz * uy * uz + h * Fy * uz * uz) * hyzz(I) +
                      (Fx * uy * uz + Fy * ux * uz + Fz * ux * uy) * hxyz(I)) * kInvCs6;
    return wi * (A1 + A2 + A3);
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

}  // namespace cglbm
