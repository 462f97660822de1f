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
__host__ __device__ __forceinline__ float guo_full(float Fx, float Fy, float Fz, float ux, float uy, float uz) {
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

// guo_full(0, Fy, 0, ux, uy, uz), the only force the build applies (gravity along y), without the
// terms that are a product with the zero force components or with a zero Hermite coefficient.
// Those terms are +0 or -0 for finite u; adding one to a partial sum leaves every nonzero sum as
// it is, so each surviving term enters the same sums in the same order and with the same
// rounding as in guo_full. The two can differ only in the sign of a zero result, which the step
// then adds to a nonzero equilibrium (feq > 0 for rho > 0), so the populations are the same
// bits. Without fast-math the compiler may not drop 0 * x itself (x could be inf or NaN).
template <int Q, int I>
__host__ __device__ __forceinline__ float guo_y(float Fy, float ux, float uy, float uz) {
    constexpr float wi = weight<Q>(I);
    constexpr float cy = static_cast<float>(kCy[I]);
    const float h = 0.5f;
    // A2 = ((Fy uy hyy + Fy ux hxy) + Fy uz hyz) / cs^4
    float A2 = (Fy * uy) * hyy(I);
    if constexpr (hxy(I) != 0.0f) A2 = A2 + (Fy * ux) * hxy(I);
    if constexpr (hyz(I) != 0.0f) A2 = A2 + (Fy * uz) * hyz(I);
    A2 = A2 * kInvCs4;
    // A3 terms in guo_full's order: xxy, xyy, yyz, yzz, xyz
    float A3 = 0.0f;
    bool any = false;
    auto add = [&](float t) {
        A3 = any ? A3 + t : t;
        any = true;
    };
    if constexpr (hxxy(I) != 0.0f) add(((h * Fy) * ux * ux) * hxxy(I));
    if constexpr (hxyy(I) != 0.0f) add((Fy * ux * uy) * hxyy(I));
    if constexpr (hyyz(I) != 0.0f) add((Fy * uy * uz) * hyyz(I));
    if constexpr (hyzz(I) != 0.0f) add(((h * Fy) * uz * uz) * hyzz(I));
    if constexpr (hxyz(I) != 0.0f) add((Fy * ux * uz) * hxyz(I));
    float sum;
    if constexpr (cy != 0.0f) sum = ((Fy * cy) * kInvCs2) + A2;
    else sum = A2;
    if (any) sum = sum + A3 * kInvCs6;
    return wi * sum;
}

// guo_y of direction I and of its opposite I + 1 from one set of products: the opposite direction
// has the same weight and second-order Hermite values and negated first- and third-order ones
// (binary32-exact: -a - (-b) = -(a - b)), so its first- and third-order parts are the negated
// parts of direction I. Each result has guo_y's bits, up to the sign of an all-zero result (Fy = 0),
// which, as above, changes no population.
template <int Q, int I>
__host__ __device__ __forceinline__ void guo_y_pair(float Fy, float ux, float uy, float uz, float &Fi, float &Fo) {
    constexpr float wi = weight<Q>(I);
    constexpr float cy = static_cast<float>(kCy[I]);
    const float h = 0.5f;
    float A2 = (Fy * uy) * hyy(I);
    if constexpr (hxy(I) != 0.0f) A2 = A2 + (Fy * ux) * hxy(I);
    if constexpr (hyz(I) != 0.0f) A2 = A2 + (Fy * uz) * hyz(I);
    A2 = A2 * kInvCs4;
    float A3 = 0.0f;
    bool any = false;
    auto add = [&](float t) {
        A3 = any ? A3 + t : t;
        any = true;
    };
    if constexpr (hxxy(I) != 0.0f) add(((h * Fy) * ux * ux) * hxxy(I));
    if constexpr (hxyy(I) != 0.0f) add((Fy * ux * uy) * hxyy(I));
    if constexpr (hyyz(I) != 0.0f) add((Fy * uy * uz) * hyyz(I));
    if constexpr (hyzz(I) != 0.0f) add(((h * Fy) * uz * uz) * hyzz(I));
    if constexpr (hxyz(I) != 0.0f) add((Fy * ux * uz) * hxyz(I));
    float si = A2, so = A2;
    if constexpr (cy != 0.0f) {
        const float A1 = (Fy * cy) * kInvCs2;
        si = A1 + A2;
        so = -A1 + A2;
    }
    if (any) {
        const float A3s = A3 * kInvCs6;
        si = si + A3s;
        so = so + -A3s;
    }
    Fi = wi * si;
    Fo = wi * so;
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

}  // namespace cglbm
