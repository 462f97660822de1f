// Direct-gather form of the fused step: one thread per target cell pulls every direction from
// global memory and rebuilds the source population with the reference's full expressions
// (lbm.cuh:20-262), including the per-direction colour gradient of forces.cuh:20-60. It is the
// simple, layout-agnostic path (any pitch, the reference's own ghost-column layout included) and
// the GPU-side cross-check of the z-marching kernel. It is not the fast path.
#include "step_common.hpp"

namespace cglbm {

namespace {

template <int Q, int YBC>
__device__ __forceinline__ void colour_gradient_at(const StepArgs &a, int xs, int ys, int zs, float &Fx, float &Fy,
                                                   float &Fz) {
    const float *__restrict__ rr = a.A[0];
    const float *__restrict__ rb = a.A[1];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    unroll<1, Q>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr float wj = weight<Q>(j);
        constexpr float cx = static_cast<float>(kCx[j]), cy = static_cast<float>(kCy[j]), cz = static_cast<float>(kCz[j]);
        const int xn = wrap(xs + kCx[j], a.nx);
        const int zn = wrap(zs + kCz[j], a.nz);
        int yn = ys + kCy[j];
        if constexpr (YBC == kYPeriodic) {
            yn = wrap(yn, a.ny);
        } else {
            if (yn >= a.ny + 1) yn = a.ny;  // force_y_outlet_limited (mathUtilities.cuh:25-33)
        }
        const uint32_t n = static_cast<uint32_t>(xn) + a.py * static_cast<uint32_t>(yn) + a.pz * static_cast<uint32_t>(zn);
        const float psi_n = (rr[n] - rb[n]) / (rr[n] + rb[n]);
        sx += wj * psi_n * cx;
        sy += wj * psi_n * cy;
        sz += wj * psi_n * cz;
    });
    Fx = kInvCs2 * sx;
    Fy = kInvCs2 * sy;
    Fz = kInvCs2 * sz;
}

template <int Q, int YBC>
__global__ void __launch_bounds__(256) rcs_gather_kernel(StepArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int yi = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= a.nx || yi >= a.ny || z >= a.nz) return;
    const int y = (YBC == kYPeriodic) ? yi : yi + 1;
    const uint32_t id = static_cast<uint32_t>(x) + a.py * static_cast<uint32_t>(y) + a.pz * static_cast<uint32_t>(z);

    float sumr = 0.0f, sumb = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
    float Axx = 0.0f, Axy = 0.0f, Ayy = 0.0f, Ayz = 0.0f, Azz = 0.0f, Axz = 0.0f;
    const Coeffs c = a.c;

    unroll<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr float cx = static_cast<float>(kCx[i]), cy = static_cast<float>(kCy[i]), cz = static_cast<float>(kCz[i]);
        const int xs = wrap(x - kCx[i], a.nx);
        const int zs = wrap(z - kCz[i], a.nz);
        int ys = y - kCy[i];
        if constexpr (YBC == kYPeriodic) ys = wrap(ys, a.ny);
        const uint32_t src = static_cast<uint32_t>(xs) + a.py * static_cast<uint32_t>(ys) + a.pz * static_cast<uint32_t>(zs);
        float fr, fb;
        bool ghost = false;
        if constexpr (YBC != kYPeriodic) ghost = (ys == 0 || ys == a.ny + 1);
        if (ghost) {
            const uint32_t idF = static_cast<uint32_t>(xs) + a.py * static_cast<uint32_t>(ys == 0 ? 1 : a.ny) +
                                 a.pz * static_cast<uint32_t>(zs);
            const float rrB = a.A[0][src], rbB = a.A[1][src];
            const float rhoB = rrB + rbB;
            const float inv = 1.0f / rhoB;
            const float aR = rrB * inv, aB = rbB * inv;
            const float gieq = feq_full<Q, i>(rhoB, a.A[2][src], a.A[3][src], a.A[4][src]);
            const float gineq = fneqr_full<Q, i>(a.A[5][idF], a.A[6][idF], a.A[7][idF], a.A[8][idF], a.A[9][idF],
                                                 a.A[10][idF], a.A[2][idF], a.A[3][idF], a.A[4][idF]);
            const float gi = gieq + c.oms * gineq;
            fr = aR * gi;
            fb = aB * gi;
        } else {
            const float rr = a.A[0][src], rb = a.A[1][src];
            const float rho_s = rr + rb;
            const float inv = 1.0f / (rr + rb);
            const float aR = rr * inv, aB = rb * inv;
            const float ux = a.A[2][src], uy = a.A[3][src], uz = a.A[4][src];
            const float gieq = feq_full<Q, i>(rho_s, ux, uy, uz);
            float gineq = fneqr_full<Q, i>(a.A[5][src], a.A[6][src], a.A[7][src], a.A[8][src], a.A[9][src],
                                           a.A[10][src], ux, uy, uz);
            float omega1;
            if (a.forced) {
                const float Fby = -(c.grav * (c.atwood * (rb - rr)));
                const float Fi = guo_full<Q, i>(0.0f, Fby, 0.0f, ux, uy, uz);
                gineq = gineq - 0.5f * Fi;
                omega1 = gieq + c.oms * gineq + c.one_minus_half_omega * Fi;
            } else {
                omega1 = gieq + c.oms * gineq;
            }
            const float ind = 4.0f * rr * rb * inv * inv;
            float gi = omega1;
            float Delta = 0.0f;
            if (ind > 1.0e-4f) {
                float Fx, Fy, Fz;
                colour_gradient_at<Q, YBC>(a, xs, ys, zs, Fx, Fy, Fz);
                const float af = sqrtf(Fx * Fx + Fy * Fy + Fz * Fz);
                const bool small = af <= 1.0e-4f;
                if (!(small && a.conservative)) {
                    float c2 = 0.0f;
                    if (i != 0 && !small) {
                        const float Fc = Fx * cx + Fy * cy + Fz * cz;
                        c2 = (Fc * Fc) / (af * af);
                    }
                    gi += 0.5f * c.Acoef * af * (weight<Q>(i) * c2 - bcoef<Q>(i));
                }
                if constexpr (i != 0) {
                    const float rho = rr + rb;
                    float cphi = 0.0f;
                    if (!small) {
                        const float dot = Fx * cx + Fy * cy + Fz * cz;
                        cphi = dot * invcnorm(i) / af;
                    }
                    const float coeff = c.beta * (rr * rb) / (rho * rho);
                    const float fieq = weight<Q>(i) * rho;
                    Delta = coeff * cphi * fieq;
                }
            }
            fr = aR * gi + Delta;
            fb = aB * gi - Delta;
        }
        sumr += fr;
        sumb += fb;
        const float g = fr + fb;
        jx += g * cx;
        jy += g * cy;
        jz += g * cz;
        Axx += g * hxx(i);
        Axy += g * hxy(i);
        Ayy += g * hyy(i);
        Ayz += g * hyz(i);
        Azz += g * hzz(i);
        Axz += g * hxz(i);
    });

    const float rho_t = sumr + sumb;
    const float inv_rho = 1.0f / rho_t;
    float vx, vy, vz;
    if (a.forced) {
        const float Fty = -(c.grav * (c.atwood * (sumb - sumr)));
        vx = jx * inv_rho;
        vy = (jy + 0.5f * Fty) * inv_rho;
        vz = jz * inv_rho;
    } else {
        vx = jx * inv_rho;
        vy = jy * inv_rho;
        vz = jz * inv_rho;
    }
    a.B[0][id] = sumr;
    a.B[1][id] = sumb;
    a.B[2][id] = vx;
    a.B[3][id] = vy;
    a.B[4][id] = vz;
    a.B[5][id] = Axx - rho_t * vx * vx;
    a.B[6][id] = Axy - rho_t * vx * vy;
    a.B[7][id] = Ayy - rho_t * vy * vy;
    a.B[8][id] = Ayz - rho_t * vy * vz;
    a.B[9][id] = Azz - rho_t * vz * vz;
    a.B[10][id] = Axz - rho_t * vx * vz;
}

// Boundary planes of B (jet/boundary_conditions.cuh:16-109, or the wall copies of
// rti/rtiIn.cuh:77-119), one thread per (x, z) column.
__global__ void __launch_bounds__(256) ybc_kernel(BcArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= a.nx || z >= a.nz) return;
    const uint32_t base = static_cast<uint32_t>(x) + a.pz * static_cast<uint32_t>(z);
    const uint32_t i0 = base, i1 = base + a.py;
    const uint32_t iT = base + a.py * static_cast<uint32_t>(a.ny + 1), iF = base + a.py * static_cast<uint32_t>(a.ny);
    if (a.ybc == kYWalls) {
        const int copy[8] = {0, 1, 5, 6, 7, 8, 9, 10};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int f = copy[k];
            a.B[f][i0] = a.B[f][i1];
            a.B[f][iT] = a.B[f][iF];
        }
#pragma unroll
        for (int f = 2; f < 5; ++f) {
            a.B[f][i0] = 0.0f;
            a.B[f][iT] = 0.0f;
        }
        return;
    }
    const float dx = static_cast<float>(x) - a.jet_x0;
    const float dz = static_cast<float>(z) - a.jet_z0;
    const bool in_jet = (dx * dx + dz * dz) <= (a.jet_r * a.jet_r);
    const float m = in_jet ? 1.0f : 0.0f;
    float U = a.jet_u, Ux = 0.0f, Uz = 0.0f;
    const bool noisy = a.jet_noise != 0.0f;
    if (noisy) {
        const float eps = a.jet_noise;
        const float r0 = inlet_uniform(a.seed, x, z, a.step, 0);
        const float r1 = inlet_uniform(a.seed, x, z, a.step, 1);
        const float r2 = inlet_uniform(a.seed, x, z, a.step, 2);
        Ux = U * (eps * (2.0f * r1 - 1.0f));
        Uz = U * (eps * (2.0f * r2 - 1.0f));
        U = U * (1.0f + eps * (2.0f * r0 - 1.0f));
        Ux = m * Ux;
        Uz = m * Uz;
    }
    // Outlet reads the interior plane of B before anything here writes B.
    const float rrF = a.B[0][iF], rbF = a.B[1][iF];
    const float uxF = a.B[2][iF], uyF = a.B[3][iF], uzF = a.B[4][iF];
    const float rrO = a.Aold[0][iT], rbO = a.Aold[1][iT];
    const float uxO = a.Aold[2][iT], uyO = a.Aold[3][iT], uzO = a.Aold[4][iT];
    float PF[6], P1[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        PF[k] = a.B[5 + k][iF];
        P1[k] = a.B[5 + k][i1];
    }
    a.B[0][i0] = (1.0f - m) * 1.0f;
    a.B[1][i0] = m * 1.0f;
    a.B[2][i0] = noisy ? Ux : 0.0f;
    a.B[3][i0] = m * U;
    a.B[4][i0] = noisy ? Uz : 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) a.B[5 + k][i0] = P1[k];
    const float uc = a.outlet == 1 ? 1.0f : (uyF < 0.0f ? 0.0f : (uyF > 1.0f ? 1.0f : uyF));
    const float rhoO = rrO + rbO;
    if (rhoO <= 1.0e-20f) {
        a.B[0][iT] = rrF;
        a.B[1][iT] = rbF;
        a.B[2][iT] = uxF;
        a.B[3][iT] = uyF;
        a.B[4][iT] = uzF;
    } else {
        a.B[0][iT] = rrO - uc * (rrO - rrF);
        a.B[1][iT] = rbO - uc * (rbO - rbF);
        a.B[2][iT] = uxO - uc * (uxO - uxF);
        a.B[3][iT] = uyO - uc * (uyO - uyF);
        a.B[4][iT] = uzO - uc * (uzO - uzF);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) a.B[5 + k][iT] = PF[k];
}

}  // namespace

hipError_t launch_gather(const StepArgs &a, int q, hipStream_t s) {
    const dim3 block(64, 4, 1);
    const dim3 grid((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
    const bool per = a.ybc == kYPeriodic;
    if (q == 19) {
        if (per) rcs_gather_kernel<19, kYPeriodic><<<grid, block, 0, s>>>(a);
        else rcs_gather_kernel<19, kYJet><<<grid, block, 0, s>>>(a);
    } else {
        if (per) rcs_gather_kernel<27, kYPeriodic><<<grid, block, 0, s>>>(a);
        else rcs_gather_kernel<27, kYJet><<<grid, block, 0, s>>>(a);
    }
    return hipGetLastError();
}

hipError_t launch_ybc(const BcArgs &a, hipStream_t s) {
    if (a.ybc == kYPeriodic) return hipSuccess;
    const dim3 block(64, 4, 1);
    const dim3 grid((a.nx + 63) / 64, (a.nz + 3) / 4, 1);
    ybc_kernel<<<grid, block, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace cglbm
