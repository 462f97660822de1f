// Population engine of the push-form step: everything one source cell does, shared by the
// single-domain z-marching kernel and the batched kernel.
//
// The arithmetic is the reference's (lbm.cuh:20-262 with pdf.cuh:16-163, forces.cuh:20-137,
// collisionOperators.cuh:16-80): every binary32 rounding step of its expressions is kept in the
// same order. Terms that are exactly zero for a direction (a product with c_a = 0 or H = 0) are
// skipped, which changes no value. Opposite directions i, i+1 share |c.u|^2, the cubic
// equilibrium term, the regularized second-order part, Omega2 and the recolouring magnitude
// (they differ only by exact sign flips), so that work is done once per pair.
#pragma once

#include "step_common.hpp"

namespace cglbm {

inline constexpr float kC1 = 1.0f - kCs2;  // H_aa for |c_a| = 1 (stencil_ct.cuh:96-99)

struct Mom {
    float rr, rb, ux, uy, uz, pxx, pxy, pyy, pyz, pzz, pxz;
};

__device__ __forceinline__ Mom load_mom(const float *const *A, uint32_t id) {
    Mom m;
    m.rr = A[0][id];
    m.rb = A[1][id];
    m.ux = A[2][id];
    m.uy = A[3][id];
    m.uz = A[4][id];
    m.pxx = A[5][id];
    m.pxy = A[6][id];
    m.pyy = A[7][id];
    m.pyz = A[8][id];
    m.pzz = A[9][id];
    m.pxz = A[10][id];
    return m;
}

// acc + s*v for a compile-time sign s in {-1, +1} (exact: adding -v is subtracting v).
template <int S>
__device__ __forceinline__ float addsgn(float acc, float v) {
    if constexpr (S > 0) return acc + v;
    else return acc - v;
}

// Direction pairs: pair k holds directions 2k+1 and 2k+2 (opposites, stencil_ct.cuh:16-23).
template <int Q>
__host__ __device__ constexpr uint32_t pair_mask_where(int axis) {
    uint32_t m = 0;
    for (int k = 0; k < (Q - 1) / 2; ++k) {
        const int i = 2 * k + 1;
        const int c = axis == 0 ? kCx[i] : (axis == 1 ? kCy[i] : kCz[i]);
        if (c != 0) m |= 1u << k;
    }
    return m;
}
template <int Q>
__host__ __device__ constexpr uint32_t pair_mask_all() { return (1u << ((Q - 1) / 2)) - 1u; }

// Per-source quantities shared by all directions.
template <int Q>
struct SrcTerms {
    float T2, U4;
    float wr0, wr1, wr2, wr3;  // w(|c|^2 class) * rho
    float pxx1, pxxc, pyy1, pyyc, pzz1, pzzc, p2xy, p2yz, p2xz;
    float xxy1, xxyc, xxz1, xxzc, xyy1, xyyc, xzz1, xzzc, yyz1, yyzc, yzz1, yzzc, a2xyz;

    __device__ __forceinline__ void build_feq(float rho, float ux, float uy, float uz) {
        const float usq = ux * ux + uy * uy + uz * uz;
        T2 = usq * kInv2Cs2;
        U4 = usq * kInv2Cs4;
        wr0 = weight<Q>(0) * rho;
        wr1 = weight<Q>(1) * rho;
        wr2 = weight<Q>(7) * rho;
        if constexpr (Q == 27) wr3 = weight<Q>(19) * rho;
    }
    __device__ __forceinline__ void build_neq(float Pxx, float Pxy, float Pyy, float Pyz, float Pzz, float Pxz,
                                              float vx, float vy, float vz) {
        pxx1 = Pxx * kC1;
        pxxc = Pxx * kCs2;
        pyy1 = Pyy * kC1;
        pyyc = Pyy * kCs2;
        pzz1 = Pzz * kC1;
        pzzc = Pzz * kCs2;
        p2xy = 2.0f * Pxy;
        p2yz = 2.0f * Pyz;
        p2xz = 2.0f * Pxz;
        const float a3xxy = Pxx * vy + p2xy * vx;
        const float a3xxz = Pxx * vz + p2xz * vx;
        const float a3xyy = Pyy * vx + p2xy * vy;
        const float a3xzz = Pzz * vx + p2xz * vz;
        const float a3yyz = Pyy * vz + p2yz * vy;
        const float a3yzz = Pzz * vy + p2yz * vz;
        xxy1 = a3xxy * kC1;
        xxyc = a3xxy * kCs2;
        xxz1 = a3xxz * kC1;
        xxzc = a3xxz * kCs2;
        xyy1 = a3xyy * kC1;
        xyyc = a3xyy * kCs2;
        xzz1 = a3xzz * kC1;
        xzzc = a3xzz * kCs2;
        yyz1 = a3yyz * kC1;
        yyzc = a3yyz * kCs2;
        yzz1 = a3yzz * kC1;
        yzzc = a3yzz * kCs2;
        if constexpr (Q == 27) {
            const float a3xyz = Pxy * vz + Pxz * vy + Pyz * vx;
            a2xyz = 2.0f * a3xyz;
        }
    }
    template <int I>
    __device__ __forceinline__ float wr() const {
        constexpr int s = norm2(I);
        if constexpr (s == 0) return wr0;
        else if constexpr (s == 1) return wr1;
        else if constexpr (s == 2) return wr2;
        else return wr3;
    }
};

// H_aab = c_b (c_a^2 - cs2) in binary32: 0 when c_b = 0, c_b * (1 - cs2) when |c_a| = 1,
// -c_b * cs2 when c_a = 0 (stencil_ct.cuh:134-168).
template <int CA, int CB>
struct H3 {
    static constexpr int sign = (CB == 0) ? 0 : ((CA != 0) ? CB : -CB);
    static constexpr bool one = (CA != 0);
};

// Second-order regularized term of direction I before the weight (pdf.cuh:66), order
// xx, xy, yy, yz, zz, xz; the opposite direction has the same value.
template <int Q, int I>
__device__ __forceinline__ float a2neq(const SrcTerms<Q> &t) {
    constexpr int cx = kCx[I], cy = kCy[I], cz = kCz[I];
    float acc = (cx != 0) ? t.pxx1 : -t.pxxc;
    if constexpr (cx * cy != 0) acc = addsgn<cx * cy>(acc, t.p2xy);
    acc = (cy != 0) ? acc + t.pyy1 : acc - t.pyyc;
    if constexpr (cy * cz != 0) acc = addsgn<cy * cz>(acc, t.p2yz);
    acc = (cz != 0) ? acc + t.pzz1 : acc - t.pzzc;
    if constexpr (cx * cz != 0) acc = addsgn<cx * cz>(acc, t.p2xz);
    return acc * kInv2Cs4;
}

template <int S, bool ONE>
__device__ __forceinline__ void a3_term(float &acc, bool &started, float v1, float vc) {
    if constexpr (S != 0) {
        const float v = ONE ? v1 : vc;
        if (!started) {
            acc = (S > 0) ? v : -v;
            started = true;
        } else {
            acc = addsgn<S>(acc, v);
        }
    }
}

// Third-order regularized term of direction I (pdf.cuh:68-76), order xxy, xxz, xyy, xzz, yyz,
// yzz, xyz; the opposite direction's is its negation.
template <int Q, int I>
__device__ __forceinline__ float a3neq(const SrcTerms<Q> &t) {
    constexpr int cx = kCx[I], cy = kCy[I], cz = kCz[I];
    float acc = 0.0f;
    bool started = false;
    a3_term<H3<cx, cy>::sign, H3<cx, cy>::one>(acc, started, t.xxy1, t.xxyc);
    a3_term<H3<cx, cz>::sign, H3<cx, cz>::one>(acc, started, t.xxz1, t.xxzc);
    a3_term<H3<cy, cx>::sign, H3<cy, cx>::one>(acc, started, t.xyy1, t.xyyc);
    a3_term<H3<cz, cx>::sign, H3<cz, cx>::one>(acc, started, t.xzz1, t.xzzc);
    a3_term<H3<cy, cz>::sign, H3<cy, cz>::one>(acc, started, t.yyz1, t.yyzc);
    a3_term<H3<cz, cy>::sign, H3<cz, cy>::one>(acc, started, t.yzz1, t.yzzc);
    if constexpr (Q == 27) {
        constexpr int sxyz = cx * cy * cz;
        if constexpr (sxyz != 0) {
            if (!started) {
                acc = (sxyz > 0) ? t.a2xyz : -t.a2xyz;
                started = true;
            } else {
                acc = addsgn<sxyz>(acc, t.a2xyz);
            }
        }
    }
    return acc * kInv2Cs6;
}

// c_I . v in the reference's left-to-right order with zero terms dropped.
template <int I>
__device__ __forceinline__ float cdot(float vx, float vy, float vz) {
    constexpr int cx = kCx[I], cy = kCy[I], cz = kCz[I];
    float acc = 0.0f;
    bool started = false;
    if constexpr (cx != 0) {
        acc = (cx > 0) ? vx : -vx;
        started = true;
    }
    if constexpr (cy != 0) {
        if (started) acc = addsgn<cy>(acc, vy);
        else {
            acc = (cy > 0) ? vy : -vy;
            started = true;
        }
    }
    if constexpr (cz != 0) {
        if (started) acc = addsgn<cz>(acc, vz);
        else acc = (cz > 0) ? vz : -vz;
    }
    return acc;
}

// Interface data of one source: colour gradient, |F|, the Omega2 prefactor (0.5 A)|F| and the
// recolouring coefficient beta rr rb / rho^2 (forces.cuh:114-137, collisionOperators.cuh:60-80).
struct Iface {
    bool on;     // interface indicator 4 rr rb / rho^2 > 1e-4 (lbm.cuh:181-187)
    bool small;  // |F| <= 1e-4
    bool omega;  // Omega2 applied (false in the conservative variant with small |F|, and for a
                 // non-interface lane that runs the interface form with its wave)
    float Fx, Fy, Fz, af, af2, ha, coeff;
    float rinv_af2, rinv_af;  // RN(1 / |F|^2), RN(1 / |F|): shared by the per-direction divisions
};

// a / b correctly rounded, given y = RN(1 / b) shared by many divisions by b. Markstein's
// correction: with q = RN(a y) and the exact remainder r = a - b q (one fma), RN(q + r y) is the
// correctly rounded quotient provided r does not underflow. The quotient of positive b takes the
// sign of a, which the corrected value has for every nonzero a; zero numerators, for which the
// correction returns +0, get a's sign of zero back that way (a / b for finite nonzero b). Numerators
// below 2^-100, where the remainder could underflow, take the IEEE division. Four operations
// instead of the eleven of the IEEE division sequence; bitwise equal to a / b (checked against the
// IEEE quotient for every significand of b and the operand ranges of the colour-gradient terms).
// The tiny-numerator case is a wave-uniform branch that stays a branch: left to itself the
// compiler evaluates the IEEE division on every call and selects, which costs more than the
// division it replaces; the empty volatile asm cannot be speculated, so the division runs only in
// a wave that has such a lane.
__device__ __forceinline__ float div_shared(float a, float b, float y) {
    float q = a * y;
    const float r = fmaf(-q, b, a);
    q = __builtin_copysignf(fmaf(r, y, q), a);
    const bool tiny = fabsf(a) < 0x1p-100f && a != 0.0f;
    if (__builtin_amdgcn_ballot_w64(tiny) != 0) {
        asm volatile("");
        if (tiny) q = a / b;
    }
    return q;
}

// a / b given y = RN(1 / b), without div_shared's zero and tiny-numerator corrections: used only
// for the Omega2 direction term c2 = (F.c)^2 / |F|^2 of a source with |F| > 1e-4, where it gives
// the same om as the correctly rounded quotient. The numerator (F.c)^2 is never -0, and a zero
// numerator gives +0 here as in IEEE division. The result differs from RN(a / b) only when the
// remainder underflows, which needs a < 2^-100; the quotient is then below 2^-100 / 1e-8 (|F|^2 >
// 1e-8), and this form returns a value within a few ulps of it. Either value, times the weight,
// is below 1e-23, less than half an ulp of every B_i (at least 1/216, the D3Q27 corner weight;
// half an ulp 2.3e-10), so w c2 - B_i rounds to -B_i exactly and om is the same bits.
__device__ __forceinline__ float div_markstein(float a, float b, float y) {
    float q = a * y;
    const float r = fmaf(-q, b, a);
    return fmaf(r, y, q);
}

__device__ __forceinline__ Iface iface_begin(const Mom &s) {
    Iface f;
    const float inv = 1.0f / (s.rr + s.rb);
    const float ind = 4.0f * s.rr * s.rb * inv * inv;
    f.on = ind > 1.0e-4f;
    f.small = false;
    f.omega = true;
    f.Fx = f.Fy = f.Fz = f.af = f.af2 = f.ha = f.coeff = 0.0f;
    f.rinv_af2 = f.rinv_af = 0.0f;
    return f;
}

// Also prepares the shared reciprocals of the interface divisions (div_markstein, div_shared) and
// the per-lane mask of the wave-uniform interface form (D3Q19; D3Q27 runs its interface form under
// a per-lane branch, where the mask is true).
template <int Q>
__device__ __forceinline__ void iface_finish(Iface &f, const Mom &s, const Coeffs &c, int conservative) {
    f.af = sqrtf(f.Fx * f.Fx + f.Fy * f.Fy + f.Fz * f.Fz);
    f.af2 = f.af * f.af;
    f.small = f.af <= 1.0e-4f;
    f.omega = !(f.small && conservative);
    f.ha = 0.5f * c.Acoef * f.af;
    const float rho = s.rr + s.rb;
    f.coeff = c.beta * (s.rr * s.rb) / (rho * rho);
    f.rinv_af2 = 1.0f / f.af2;
    f.rinv_af = 1.0f / f.af;
    f.omega = f.omega && f.on;
}

// Colour gradient F = (1/cs2) sum_{j>=1} w_j psi(x + c_j) c_j (forces.cuh:20-60) with psi read
// through `psi_at(cx, cy, cz)`, accumulated in the reference's j order.
template <int Q, typename PsiAt>
__device__ __forceinline__ void colour_gradient(PsiAt &&psi_at, float &Fx, float &Fy, float &Fz) {
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    bool stx = false, sty = false, stz = false;
    unroll<1, Q>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr int cx = kCx[j], cy = kCy[j], cz = kCz[j];
        const float tj = weight<Q>(j) * psi_at(J);
        if constexpr (cx != 0) {
            if (stx) ax = addsgn<cx>(ax, tj);
            else {
                ax = (cx > 0) ? tj : -tj;
                stx = true;
            }
        }
        if constexpr (cy != 0) {
            if (sty) ay = addsgn<cy>(ay, tj);
            else {
                ay = (cy > 0) ? tj : -tj;
                sty = true;
            }
        }
        if constexpr (cz != 0) {
            if (stz) az = addsgn<cz>(az, tj);
            else {
                az = (cz > 0) ? tj : -tj;
                stz = true;
            }
        }
    });
    Fx = kInvCs2 * ax;
    Fy = kInvCs2 * ay;
    Fz = kInvCs2 * az;
}

// All populations of a regular source restricted to the pairs in PMASK (and the rest direction
// when REST), handed to sink(integral_constant<I>, fr, fb). IFACE selects the interface form
// (Omega2 and recolouring, lbm.cuh:187-208) at compile time so the per-pair work has no branches;
// inside it the |F| <= 1e-4 and conservative cases are selects. Adding the +0 that a skipped
// Omega2 or recolouring contributes changes no value.
template <int Q, uint32_t PMASK, bool REST, bool FORCED, bool IFACE, typename Sink>
__device__ __forceinline__ void emit_regular(const Mom &s, const Coeffs &c, const Iface &fi, Sink &&sink) {
    constexpr bool forced = FORCED;
    const float rho = s.rr + s.rb;
    const float inv = 1.0f / (s.rr + s.rb);
    const float aR = s.rr * inv, aB = s.rb * inv;
    SrcTerms<Q> t;
    t.build_feq(rho, s.ux, s.uy, s.uz);
    t.build_neq(s.pxx, s.pxy, s.pyy, s.pyz, s.pzz, s.pxz, s.ux, s.uy, s.uz);
    const float Fby = forced ? -(c.grav * (c.atwood * (s.rb - s.rr))) : 0.0f;

    if constexpr (REST) {
        const float feq0 = t.wr0 * (1.0f - t.T2);
        const float A2n = ((-t.pxxc - t.pyyc) - t.pzzc) * kInv2Cs4;
        float gneq = weight<Q>(0) * A2n;
        float g;
        if constexpr (forced) {
            const float Fi = guo_y<Q, 0>(Fby, s.ux, s.uy, s.uz);
            gneq = gneq - 0.5f * Fi;
            g = feq0 + c.oms * gneq + c.one_minus_half_omega * Fi;
        } else {
            g = feq0 + c.oms * gneq;
        }
        if constexpr (IFACE) g += fi.omega ? fi.ha * (-bcoef<Q>(0)) : -0.0f;
        sink(std::integral_constant<int, 0>{}, aR * g, aB * g);
    }
    unroll<0, (Q - 1) / 2>([&](auto K) {
        constexpr int k = decltype(K)::value;
        if constexpr ((PMASK >> k) & 1u) {
            constexpr int I = 2 * k + 1;
            constexpr int O = I + 1;
            constexpr float w = weight<Q>(I);
            const float wr = t.template wr<I>();
            const float cu = cdot<I>(s.ux, s.uy, s.uz);
            const float t1 = cu * kInvCs2;
            const float cu2 = cu * cu;
            const float t3 = cu2 * kInv2Cs4;
            const float A2i = (t1 - t.T2) + t3;
            const float A2o = (-t1 - t.T2) + t3;
            const float inner = cu2 * kInv6Cs6 - t.U4;
            const float A3 = cu * inner;
            const float feqi = wr * ((1.0f + A2i) + A3);
            const float feqo = wr * ((1.0f + A2o) - A3);
            const float An2 = a2neq<Q, I>(t);
            const float An3 = a3neq<Q, I>(t);
            float gneqi = w * (An2 + An3);
            float gneqo = w * (An2 - An3);
            float gi, go;
            if constexpr (forced) {
                float Fii, Fio;
                guo_y_pair<Q, I>(Fby, s.ux, s.uy, s.uz, Fii, Fio);
                gneqi = gneqi - 0.5f * Fii;
                gneqo = gneqo - 0.5f * Fio;
                gi = feqi + c.oms * gneqi + c.one_minus_half_omega * Fii;
                go = feqo + c.oms * gneqo + c.one_minus_half_omega * Fio;
            } else {
                gi = feqi + c.oms * gneqi;
                go = feqo + c.oms * gneqo;
            }
            if constexpr (IFACE) {
                const float Fc = cdot<I>(fi.Fx, fi.Fy, fi.Fz);
                const float c2q = div_markstein(Fc * Fc, fi.af2, fi.rinv_af2);
                const float c2 = fi.small ? 0.0f : c2q;
                // a skipped Omega2 adds -0, which leaves every value as it is
                const float om = fi.omega ? fi.ha * (w * c2 - bcoef<Q>(I)) : -0.0f;
                gi += om;
                go += om;
                const float cq = div_shared(Fc * invcnorm(I), fi.af, fi.rinv_af);
                const float cphi = fi.small ? 0.0f : cq;
                const float D = fi.coeff * cphi * wr;
                const float ri = aR * gi, bi = aB * gi, ro = aR * go, bo = aB * go;
                // A non-interface lane running with an interface wave adds no recolouring term: it
                // adds -0 where the term is added and subtracts +0 where it is subtracted, both of
                // which leave every value (signed zeros included) as it is.
                const float Dp = fi.on ? D : -0.0f;
                const float Dm = fi.on ? D : 0.0f;
                sink(std::integral_constant<int, I>{}, ri + Dp, bi - Dm);
                sink(std::integral_constant<int, O>{}, ro - Dm, bo + Dp);
            } else {
                sink(std::integral_constant<int, I>{}, aR * gi, aB * gi);
                sink(std::integral_constant<int, O>{}, aR * go, aB * go);
            }
        }
    });
}

// One member of each pair in PMASK, for sources of which only one direction of a pair reaches a
// target that is accumulated (halo sources beside the tile, the first and last source plane of a
// chunk): direction I = 2k + 1 when bit k of OMASK is clear, its opposite O = I + 1 when it is
// set, and the other one of the two for a lane whose `flip` is 0x80000000 (flip 0 keeps the
// choice). Member O is member I's expressions with c.u, the third-order regularized term and the
// recolouring term negated: emit_regular forms O from the same intermediates with these exact sign
// flips (-t1 for t1, -A3, An2 - An3, go - D), x + (-y) is x - y in binary32, and a product or
// quotient with one operand negated is the negated result, so every population is bitwise the one
// emit_regular hands to the sink. sink(integral_constant<I>, integral_constant<bool, OMASK bit>,
// fr, fb); the member the lane took is O exactly when that bit differs from `flip`. No body-force
// form (the forced kernels emit both members).
template <int Q, uint32_t PMASK, uint32_t OMASK, bool IFACE, typename Sink>
__device__ __forceinline__ void emit_one(const Mom &s, const Coeffs &c, const Iface &fi, uint32_t flip, Sink &&sink) {
    const float rho = s.rr + s.rb;
    const float inv = 1.0f / (s.rr + s.rb);
    const float aR = s.rr * inv, aB = s.rb * inv;
    SrcTerms<Q> t;
    t.build_feq(rho, s.ux, s.uy, s.uz);
    t.build_neq(s.pxx, s.pxy, s.pyy, s.pyz, s.pzz, s.pxz, s.ux, s.uy, s.uz);
    unroll<0, (Q - 1) / 2>([&](auto K) {
        constexpr int k = decltype(K)::value;
        if constexpr ((PMASK >> k) & 1u) {
            constexpr int I = 2 * k + 1;
            constexpr bool OM = (OMASK >> k) & 1u;
            constexpr float w = weight<Q>(I);
            const uint32_t sg = flip ^ (OM ? 0x80000000u : 0u);
            const auto neg = [sg](float v) { return __uint_as_float(__float_as_uint(v) ^ sg); };
            const float wr = t.template wr<I>();
            const float cu = neg(cdot<I>(s.ux, s.uy, s.uz));
            const float t1 = cu * kInvCs2;
            const float cu2 = cu * cu;
            const float t3 = cu2 * kInv2Cs4;
            const float A2 = (t1 - t.T2) + t3;
            const float inner = cu2 * kInv6Cs6 - t.U4;
            const float A3 = cu * inner;
            const float feq = wr * ((1.0f + A2) + A3);
            const float An2 = a2neq<Q, I>(t);
            const float An3 = neg(a3neq<Q, I>(t));
            float g = feq + c.oms * (w * (An2 + An3));
            const auto IC = std::integral_constant<int, I>{};
            const auto OC = std::integral_constant<bool, OM>{};
            if constexpr (IFACE) {
                const float Fc = cdot<I>(fi.Fx, fi.Fy, fi.Fz);
                const float c2q = div_markstein(Fc * Fc, fi.af2, fi.rinv_af2);
                const float c2 = fi.small ? 0.0f : c2q;
                g += fi.omega ? fi.ha * (w * c2 - bcoef<Q>(I)) : -0.0f;
                const float cq = div_shared(Fc * invcnorm(I), fi.af, fi.rinv_af);
                const float cphi = fi.small ? 0.0f : cq;
                const float D = neg(fi.coeff * cphi * wr);
                // as emit_regular: a non-interface lane of an interface wave adds -0 / subtracts +0
                const float Dp = fi.on ? D : -0.0f;
                const float Dm = fi.on ? D : 0.0f;
                sink(IC, OC, aR * g + Dp, aB * g - Dm);
            } else {
                sink(IC, OC, aR * g, aB * g);
            }
        }
    });
}

// Ghost-plane source (lbm.cuh:79-151): equilibrium from the ghost's rho and u, regularized part
// from the adjacent interior cell F; no Omega2, recolouring or body force.
template <int Q, uint32_t PMASK, bool REST, typename Sink>
__device__ __forceinline__ void emit_ghost(const Mom &g, const Mom &F, const Coeffs &c, Sink &&sink) {
    const float rhoB = g.rr + g.rb;
    const float inv = 1.0f / rhoB;
    const float aR = g.rr * inv, aB = g.rb * inv;
    SrcTerms<Q> t;
    t.build_feq(rhoB, g.ux, g.uy, g.uz);
    t.build_neq(F.pxx, F.pxy, F.pyy, F.pyz, F.pzz, F.pxz, F.ux, F.uy, F.uz);
    if constexpr (REST) {
        const float feq0 = t.wr0 * (1.0f - t.T2);
        const float A2n = ((-t.pxxc - t.pyyc) - t.pzzc) * kInv2Cs4;
        const float gg = feq0 + c.oms * (weight<Q>(0) * A2n);
        sink(std::integral_constant<int, 0>{}, aR * gg, aB * gg);
    }
    unroll<0, (Q - 1) / 2>([&](auto K) {
        constexpr int k = decltype(K)::value;
        if constexpr ((PMASK >> k) & 1u) {
            constexpr int I = 2 * k + 1;
            constexpr int O = I + 1;
            constexpr float w = weight<Q>(I);
            const float wr = t.template wr<I>();
            const float cu = cdot<I>(g.ux, g.uy, g.uz);
            const float t1 = cu * kInvCs2;
            const float cu2 = cu * cu;
            const float t3 = cu2 * kInv2Cs4;
            const float A2i = (t1 - t.T2) + t3;
            const float A2o = (-t1 - t.T2) + t3;
            const float inner = cu2 * kInv6Cs6 - t.U4;
            const float A3 = cu * inner;
            const float feqi = wr * ((1.0f + A2i) + A3);
            const float feqo = wr * ((1.0f + A2o) - A3);
            const float An2 = a2neq<Q, I>(t);
            const float An3 = a3neq<Q, I>(t);
            const float gi = feqi + c.oms * (w * (An2 + An3));
            const float go = feqo + c.oms * (w * (An2 - An3));
            sink(std::integral_constant<int, I>{}, aR * gi, aB * gi);
            sink(std::integral_constant<int, O>{}, aR * go, aB * go);
        }
    });
}

// Target-side accumulation in the reference order i = 0..Q-1 (lbm.cuh:206-261) from the
// populations pulled into target cell t, read through pop(I) -> float2(fr, fb). Returns the 11
// new moments; `forced` adds the half-force velocity shift of the body-force option.
template <int Q, bool FORCED, typename Pop>
__device__ __forceinline__ void accumulate(Pop &&pop, const Coeffs &c, float out[11]) {
    constexpr bool forced = FORCED;
    const float2 p0 = pop(std::integral_constant<int, 0>{});
    float sumr = p0.x, sumb = p0.y;
    const float g0 = sumr + sumb;
    float jx = 0.0f, jy = 0.0f, jz = 0.0f;
    float Axx = g0 * hxx(0), Ayy = g0 * hyy(0), Azz = g0 * hzz(0);
    float Axy = 0.0f, Ayz = 0.0f, Axz = 0.0f;
    unroll<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int cx = kCx[i], cy = kCy[i], cz = kCz[i];
        const float2 p = pop(I);
        sumr += p.x;
        sumb += p.y;
        const float g = p.x + p.y;
        if constexpr (cx != 0) jx = addsgn<cx>(jx, g);
        if constexpr (cy != 0) jy = addsgn<cy>(jy, g);
        if constexpr (cz != 0) jz = addsgn<cz>(jz, g);
        Axx += g * hxx(i);
        if constexpr (cx * cy != 0) Axy = addsgn<cx * cy>(Axy, g);
        Ayy += g * hyy(i);
        if constexpr (cy * cz != 0) Ayz = addsgn<cy * cz>(Ayz, g);
        Azz += g * hzz(i);
        if constexpr (cx * cz != 0) Axz = addsgn<cx * cz>(Axz, g);
    });
    const float rho_t = sumr + sumb;
    const float inv_rho = 1.0f / rho_t;
    float vx, vy, vz;
    if constexpr (forced) {
        const float Fty = -(c.grav * (c.atwood * (sumb - sumr)));
        vx = jx * inv_rho;
        vy = (jy + 0.5f * Fty) * inv_rho;
        vz = jz * inv_rho;
    } else {
        vx = jx * inv_rho;
        vy = jy * inv_rho;
        vz = jz * inv_rho;
    }
    out[0] = sumr;
    out[1] = sumb;
    out[2] = vx;
    out[3] = vy;
    out[4] = vz;
    out[5] = Axx - rho_t * vx * vx;
    out[6] = Axy - rho_t * vx * vy;
    out[7] = Ayy - rho_t * vy * vy;
    out[8] = Ayz - rho_t * vy * vz;
    out[9] = Azz - rho_t * vz * vz;
    out[10] = Axz - rho_t * vx * vz;
}

}  // namespace cglbm
