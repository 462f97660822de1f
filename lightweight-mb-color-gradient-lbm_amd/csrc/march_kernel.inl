// z-marching fused step kernel template (included by march_q19.hip and march_q27.hip).
//
// Same arithmetic as lbm.cuh:20-262 (see march_core.hpp), organised for the hardware:
//
//  * Push form: a source cell builds its post-collision populations once, opposite directions
//    sharing their common work.
//  * A work-group owns a 32 x TY column tile and marches along z over a chunk of planes. Its own
//    waves hold one source/target column per lane. Populations that stay in a column (rest, +z,
//    -z) stay in registers; the others cross to the target column through LDS, one float2 per
//    (direction, target).
//  * Ring waves build the halo sources around the tile, each only for the direction pairs that
//    can enter it: the row ring (rows -1 and TY) the pairs with c_y != 0, the column ring
//    (columns -1, 32 and the corners) the pairs with c_x != 0 (two waves; the 16-row tile splits
//    each ring's pairs over two waves, four in all, so the halo work spreads over the four SIMDs).
//    The first and last source plane of a chunk build only the pairs with c_z != 0. The ring
//    waves also form the phase field psi = (rr - rb)/(rr + rb) once per cell, two planes ahead,
//    into an LDS ring (halo 2).
//  * Target plane t gets its +z, 0 and -z contributions at source planes t-1, t, t+1; the partial
//    sets wait in registers and plane t is accumulated in the reference order i = 0..Q-1.
//  * NBUF = 2 double-buffers the LDS exchange so each plane needs one barrier; NBUF = 1 uses
//    half the LDS and two barriers per plane.
//  * The y boundary planes of the jet/RTI cases are written by the lanes that finish rows 1 and
//    ny, so a step is one launch.
//  * Blocks are remapped so that each XCD gets a contiguous range of tiles.
//  * A work-group's item is a column tile and a z-range: uniform chunks of the launch shape, or
//    the entries of an explicit list (MarchArgs::items) that the planner (plan.hip) cuts to equal
//    cost from the interface map and item timings a recording step (REC) writes.
#pragma once

#include <cstdlib>

#include "kernels.hpp"
#include "march_core.hpp"
#include "storage.hpp"

#ifndef CGLBM_MARCH_PREFETCH
#define CGLBM_MARCH_PREFETCH 1
#endif
#ifndef CGLBM_MARCH_WAVES
#define CGLBM_MARCH_WAVES 0
#endif

#if CGLBM_MARCH_WAVES > 0
#define CGLBM_MARCH_OCC __attribute__((amdgpu_waves_per_eu(CGLBM_MARCH_WAVES, CGLBM_MARCH_WAVES)))
#else
#define CGLBM_MARCH_OCC
#endif

namespace cglbm {
namespace march {

constexpr int TX = 32;
constexpr int kPsiSlots = 5;

__host__ __device__ constexpr bool cross_col(int i) { return !(kCx[i] == 0 && kCy[i] == 0); }
__host__ __device__ constexpr int slot_of(int i) {
    int s = 0;
    for (int j = 0; j < i; ++j) s += cross_col(j) ? 1 : 0;
    return s;
}
template <int Q>
__host__ __device__ constexpr int nslots() {
    int s = 0;
    for (int j = 0; j < Q; ++j) s += cross_col(j) ? 1 : 0;
    return s;
}

// Ring waves per work-group. The 16-row tile runs four (8 own waves + 4 = 3 per SIMD): the row
// and the column ring each split their direction pairs over two waves, so the halo work spreads
// over all four SIMDs instead of loading two of them. Smaller tiles run two (rows, columns).
template <int TY>
__host__ __device__ constexpr int ring_waves() {
    return TY == 16 ? 4 : 2;
}
template <int TY>
__host__ __device__ constexpr int block_threads() {
    return TX * TY + 64 * ring_waves<TY>();
}
// Pairs whose member I = 2k + 1 has a negative component along `axis`: for them the member with
// the positive component is O = I + 1 (the OMASK of emit_one).
template <int Q>
__host__ __device__ constexpr uint32_t pair_mask_neg(int axis) {
    uint32_t m = 0;
    for (int k = 0; k < (Q - 1) / 2; ++k) {
        const int i = 2 * k + 1;
        const int c = axis == 0 ? kCx[i] : (axis == 1 ? kCy[i] : kCz[i]);
        if (c < 0) m |= 1u << k;
    }
    return m;
}
// Half h (0 or 1) of the pairs in mask m: the first ceil(n/2) set bits, or the rest.
__host__ __device__ constexpr uint32_t mask_half(uint32_t m, int h) {
    int n = 0;
    for (int k = 0; k < 32; ++k) n += (m >> k) & 1u;
    uint32_t lo = 0;
    int c = 0;
    for (int k = 0; k < 32; ++k)
        if ((m >> k) & 1u) {
            if (c < (n + 1) / 2) lo |= 1u << k;
            ++c;
        }
    return h == 0 ? lo : (m & ~lo);
}

// Exchange planes: per buffer, one TY x TX plane of (red, blue) populations for each direction
// that leaves the column (a slot). Both layouts give every cell one step outside the tile a
// location that no lane reads, so a tile source pushes every direction without a bounds test:
//  * margin (M = 1): slot s is a (TY + 2) x (TX + 2) array, tile cell (x, y) at row y + 1,
//    column x + 1;
//  * packed (M = 0, the 16-row D3Q19 tile, where a full margin would not let two buffers fit the
//    160 KB LDS): the slots are stacked rows of TX + 1 cells, the last a padding column, with one
//    padding row before each slot and after the last and one padding cell at the start. A push
//    one cell past x = TX - 1 lands in the row's padding column, one before x = 0 in the padding
//    column of the row before, one row outside the slot in a padding row.
// Ring sources (outside the tile) still test their pushes: some land two cells out.
template <int Q, int TY>
__host__ __device__ constexpr int margin() {
    return (Q == 19 && TY == 16) ? 0 : 1;
}
template <int Q, int TY>
struct XferLayout {
    static constexpr int M = margin<Q, TY>();
    static constexpr int NS = nslots<Q>();
    static constexpr int RS = M ? TX + 2 : TX + 1;  // row stride (cells)
    static constexpr int SR = M ? TY + 2 : TY + 1;  // rows per slot
    static constexpr int SIZE = M ? NS * SR * RS : 1 + (NS * SR + 1) * RS;
    // cell (x, y) of slot s, for -1 <= x <= TX and -1 <= y <= TY
    __host__ __device__ static constexpr int at(int s, int x, int y) { return (RS + 1) + s * SR * RS + y * RS + x; }
};
// Compile-time proof of the layout: every cell within one step of the tile maps inside the
// buffer, and onto a tile cell exactly when it is one.
template <int Q, int TY>
constexpr bool xfer_layout_ok() {
    using XL = XferLayout<Q, TY>;
    for (int s = 0; s < XL::NS; ++s)
        for (int y = -1; y <= TY; ++y)
            for (int x = -1; x <= TX; ++x) {
                const int i = XL::at(s, x, y);
                if (i < 0 || i >= XL::SIZE) return false;
                const int t = i - (XL::RS + 1);
                const int r = t % (XL::SR * XL::RS);
                const bool tile = t >= 0 && t / (XL::SR * XL::RS) < XL::NS && r / XL::RS < TY && r % XL::RS < TX;
                if (tile != (x >= 0 && x < TX && y >= 0 && y < TY)) return false;
            }
    return true;
}
template <int Q, int TY, int NBUF>
struct Shared {
    static_assert(xfer_layout_ok<Q, TY>(), "exchange layout");
    float2 xfer[NBUF][XferLayout<Q, TY>::SIZE];
    float psi[kPsiSlots][TY + 4][TX + 4];
};

template <int Q, int TY, int NBUF>
constexpr size_t shared_bytes() {
    return sizeof(Shared<Q, TY, NBUF>);
}

template <int Q>
struct Held {
    float2 v[Q];
};

// One state as a buffer resource. FP32 storage: field f of cell `id` at byte 4 (f fs + id). FP16
// storage (storage.hpp): rhor/rhob as FP32, field f >= 2 as a scaled binary16 at byte
// 2 (fs (f + 2) + id). Loads and stores take the cell offset in a VGPR and the field offset in an
// SGPR, so the 11 accesses of a cell need no per-field address arithmetic. The field index is a
// compile-time constant at every call site after unrolling, so the format branch folds away.
template <int ST>
struct FieldBuf {
    __amdgpu_buffer_rsrc_t r;
    uint32_t fs4;
    __device__ __forceinline__ FieldBuf(const float *base, uint32_t fs, uint32_t span) {
        r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), static_cast<short>(0), static_cast<int>(span),
                                              0x00020000);
        fs4 = fs * 4u;
    }
    // the state read (src) and the state written (dst) by a step
    __device__ __forceinline__ static FieldBuf src(const MarchArgs &a) { return FieldBuf(a.A0, a.fsA, a.spanA); }
    __device__ __forceinline__ static FieldBuf dst(const MarchArgs &a) { return FieldBuf(a.B0, a.fsB, a.spanB); }
    __device__ __forceinline__ float ld(int f, uint32_t id) const {
        if (ST == kStoreF16 && f >= 2) {
            const uint32_t soff = static_cast<uint32_t>(f + 2) * (fs4 >> 1);
            return f16_decode(f, static_cast<uint16_t>(__builtin_amdgcn_raw_buffer_load_b16(r, id * 2u, soff, 0)));
        }
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, id * 4u, static_cast<uint32_t>(f) * fs4, 0));
    }
    __device__ __forceinline__ void st(int f, uint32_t id, float v) const {
        if (ST == kStoreF16 && f >= 2) {
            const uint32_t soff = static_cast<uint32_t>(f + 2) * (fs4 >> 1);
            __builtin_amdgcn_raw_buffer_store_b16(f16_encode(f, v), r, id * 2u, soff, 0);
            return;
        }
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, id * 4u, static_cast<uint32_t>(f) * fs4, 0);
    }
    __device__ __forceinline__ Mom mom(uint32_t id) const {
        Mom m;
        m.rr = ld(0, id);
        m.rb = ld(1, id);
        m.ux = ld(2, id);
        m.uy = ld(3, id);
        m.uz = ld(4, id);
        m.pxx = ld(5, id);
        m.pxy = ld(6, id);
        m.pyy = ld(7, id);
        m.pyz = ld(8, id);
        m.pzz = ld(9, id);
        m.pxz = ld(10, id);
        return m;
    }
};

// FP32 state addressed through one 64-bit pointer per field (fields in separate allocations, or a
// state of 4 GiB or more): the same accesses, field f at p[f] + cell. Each field base is uniform
// and the cell offset a 32-bit VGPR, so the loads keep the scalar-base plus vector-offset form.
// Every access of the kernel stays inside the fields, as with the buffer form.
template <>
struct FieldBuf<kStoreF32Ptr> {
    float *p[11];
    __device__ __forceinline__ static FieldBuf src(const MarchArgs &a) {
        FieldBuf b;
#pragma unroll
        for (int f = 0; f < 11; ++f) b.p[f] = const_cast<float *>(a.Af[f]);
        return b;
    }
    __device__ __forceinline__ static FieldBuf dst(const MarchArgs &a) {
        FieldBuf b;
#pragma unroll
        for (int f = 0; f < 11; ++f) b.p[f] = a.Bf[f];
        return b;
    }
    __device__ __forceinline__ float ld(int f, uint32_t id) const { return p[f][id]; }
    __device__ __forceinline__ void st(int f, uint32_t id, float v) const { p[f][id] = v; }
    __device__ __forceinline__ Mom mom(uint32_t id) const {
        Mom m;
        m.rr = ld(0, id);
        m.rb = ld(1, id);
        m.ux = ld(2, id);
        m.uy = ld(3, id);
        m.uz = ld(4, id);
        m.pxx = ld(5, id);
        m.pxy = ld(6, id);
        m.pyy = ld(7, id);
        m.pyz = ld(8, id);
        m.pzz = ld(9, id);
        m.pxz = ld(10, id);
        return m;
    }
};

// Boundary planes of the new state (jet/boundary_conditions.cuh:16-109 or the wall copies of
// rti/rtiIn.cuh:77-119), written by the lane that has just finished row y = 1 (inlet/bottom) or
// y = ny (outlet/top) of column (gx, zt).
template <int ST>
__device__ __forceinline__ void boundary_planes(const MarchArgs &a, const FieldBuf<ST> &A, const FieldBuf<ST> &B, int gx,
                                                int zt, int rown, const float out[11]) {
    auto cell = [&](int row) {
        return static_cast<uint32_t>(gx) + a.py * static_cast<uint32_t>(row) + a.pz * static_cast<uint32_t>(zt);
    };
    if (rown == 0) {
        const uint32_t i0 = cell(0);
        if (a.ybc == kYWalls) {
            B.st(0, i0, out[0]);
            B.st(1, i0, out[1]);
            B.st(2, i0, 0.0f);
            B.st(3, i0, 0.0f);
            B.st(4, i0, 0.0f);
        } else {
            const int zg = zt + a.zoff;  // global plane of a z-slab
            const float dx = static_cast<float>(gx) - a.jet_x0;
            const float dz = static_cast<float>(zg) - a.jet_z0;
            const float m = ((dx * dx + dz * dz) <= (a.jet_r * a.jet_r)) ? 1.0f : 0.0f;
            float U = a.jet_u, Ux = 0.0f, Uz = 0.0f;
            const bool noisy = a.jet_noise != 0.0f;
            if (noisy) {
                const float eps = a.jet_noise;
                const float q0 = inlet_uniform(a.seed, gx, zg, a.step, 0);
                const float q1 = inlet_uniform(a.seed, gx, zg, a.step, 1);
                const float q2 = inlet_uniform(a.seed, gx, zg, a.step, 2);
                Ux = U * (eps * (2.0f * q1 - 1.0f));
                Uz = U * (eps * (2.0f * q2 - 1.0f));
                U = U * (1.0f + eps * (2.0f * q0 - 1.0f));
                Ux = m * Ux;
                Uz = m * Uz;
            }
            B.st(0, i0, (1.0f - m) * 1.0f);
            B.st(1, i0, m * 1.0f);
            B.st(2, i0, noisy ? Ux : 0.0f);
            B.st(3, i0, m * U);
            B.st(4, i0, noisy ? Uz : 0.0f);
        }
#pragma unroll
        for (int f = 5; f < 11; ++f) B.st(f, i0, out[f]);
    }
    if (rown == a.ny - 1) {
        const uint32_t iT = cell(a.ny + 1);
        if (a.ybc == kYWalls) {
            B.st(0, iT, out[0]);
            B.st(1, iT, out[1]);
            B.st(2, iT, 0.0f);
            B.st(3, iT, 0.0f);
            B.st(4, iT, 0.0f);
        } else {
            const float rrO = A.ld(0, iT), rbO = A.ld(1, iT);
            const float uxO = A.ld(2, iT), uyO = A.ld(3, iT), uzO = A.ld(4, iT);
            const float uyF = out[3];
            const float uc = a.outlet == 1 ? 1.0f : (uyF < 0.0f ? 0.0f : (uyF > 1.0f ? 1.0f : uyF));
            const float rhoO = rrO + rbO;
            if (rhoO <= 1.0e-20f) {
#pragma unroll
                for (int f = 0; f < 5; ++f) B.st(f, iT, out[f]);
            } else {
                B.st(0, iT, rrO - uc * (rrO - out[0]));
                B.st(1, iT, rbO - uc * (rbO - out[1]));
                B.st(2, iT, uxO - uc * (uxO - out[2]));
                B.st(3, iT, uyO - uc * (uyO - out[3]));
                B.st(4, iT, uzO - uc * (uzO - out[4]));
            }
        }
#pragma unroll
        for (int f = 5; f < 11; ++f) B.st(f, iT, out[f]);
    }
}

enum : int { kRoleOwn = 0, kRoleRows = 1, kRoleCols = 2, kRoleNone = 3 };

// v mod n (result in [0, n)) for the indices of a tile and its halo, which lie within one
// extent of [0, n): one conditional add or subtract. Extents smaller than the tile can put v
// further out; those lanes take the general remainder (a branch no lane of a larger domain takes,
// so the integer division it needs is not executed there).
__device__ __forceinline__ int wrap_near(int v, int n) {
    if (v < 0) v += n;
    else if (v >= n) v -= n;
    if (v < 0 || v >= n) v = ((v % n) + n) % n;
    return v;
}

// One work item: tile (tix, tiy), target planes z0 .. z1-1 (z0 < z1), of the domain described by
// `a`, executed by the whole work-group. Ends with the work-group synchronised so the caller may
// run another item with the same LDS. REC: also write the interface map a.wmap (a separate
// instantiation, so steps that do not record carry none of its code).
template <int Q, int YBC, int TY, int NBUF, bool FORCED, int ST, bool REC = false>
__device__ __forceinline__ void march_tile(const MarchArgs &a, Shared<Q, TY, NBUF> &sh, int tix, int tiy, int z0,
                                           int z1) {
    constexpr int NOWN = TX * TY;
    constexpr int RW = ring_waves<TY>();
    constexpr int NRING = 64 * RW;
    constexpr int PX = TX + 4, PY = TY + 4, NPSI = PX * PY;
    constexpr int NPASS = (NPSI + NRING - 1) / NRING;  // psi cells per ring lane
    constexpr uint32_t ALL = pair_mask_all<Q>();
    constexpr uint32_t MX = pair_mask_where<Q>(0), MY = pair_mask_where<Q>(1), MZ = pair_mask_where<Q>(2);

    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int Lc = z1 - z0;
    const int x0 = tix * TX, r0 = tiy * TY;
    const int t = threadIdx.x;
    const Coeffs c = a.c;
    const FieldBuf<ST> Abuf = FieldBuf<ST>::src(a);
    const FieldBuf<ST> Bbuf = FieldBuf<ST>::dst(a);

    // ---- role and source cell of this lane -----------------------------------------------
    // ring waves: RW == 2 -> rows, columns; RW == 4 -> rows (pair half 0), rows (half 1),
    // columns (half 0), columns (half 1)
    int role, sx, sy;
    const int rw = t >= NOWN ? (t - NOWN) / 64 : 0;
    const int half = RW == 4 ? (rw & 1) : 0;
    const bool rows_wave = RW == 4 ? rw < 2 : rw == 0;
    if (t < NOWN) {
        role = kRoleOwn;
        sx = t % TX;
        sy = t / TX;
    } else if (rows_wave) {
        const int l = (t - NOWN) % 64;
        role = kRoleRows;
        sx = l % TX;
        sy = l < TX ? -1 : TY;
    } else {
        const int l = (t - NOWN) % 64;
        role = kRoleCols;
        if (l < TY) {
            sx = -1;
            sy = l;
        } else if (l < 2 * TY) {
            sx = TX;
            sy = l - TY;
        } else if (l < 2 * TY + 4) {
            const int k = l - 2 * TY;
            sx = (k & 1) ? TX : -1;
            sy = (k & 2) ? TY : -1;
        } else {
            role = kRoleNone;
            sx = 0;
            sy = 0;
        }
    }
    const bool ring = t >= NOWN;
    const int gx = wrap_near(x0 + sx, nx);
    const int rsrc = r0 + sy;  // interior row of the source
    int srow = -1;             // stored row of the source, -1 = not needed
    bool sghost = false;
    if (role != kRoleNone) {
        if constexpr (YBC == kYPeriodic) {
            srow = wrap_near(rsrc, ny);
        } else if (rsrc >= -1 && rsrc <= ny) {
            srow = rsrc + 1;
            sghost = (rsrc == -1 || rsrc == ny);
        }
    }
    const int frow = sghost ? (rsrc == -1 ? 1 : ny) : 0;  // adjacent interior row of a ghost source
    const bool own_target = role == kRoleOwn && (x0 + sx < nx) && (rsrc < ny);

    auto cell = [&](int x, int row, int z) -> uint32_t {
        return static_cast<uint32_t>(x) + a.py * static_cast<uint32_t>(row) + a.pz * static_cast<uint32_t>(z);
    };
    auto wrapz = [&](int z) { return wrap_near(z, nz); };

    // Cell (x, row) the moment loads of this lane address. Every lane loads, so the loads carry no
    // lane mask and overwrite their registers whole instead of merging into the previous plane's
    // values. A lane with no source reads a cell the work-group loads anyway (its values are not
    // used): an idle column-ring lane the first column-ring cell, a source row beyond the ghost
    // planes of a bounded y the nearest ghost plane.
    uint32_t lxy;
    {
        int ax = gx, arow = srow;
        if (role == kRoleNone) {
            ax = wrap_near(x0 - 1, nx);
            arow = YBC == kYPeriodic ? r0 : r0 + 1;
        } else if (srow < 0) {
            arow = rsrc < -1 ? 0 : ny + 1;
        }
        lxy = static_cast<uint32_t>(ax) + a.py * static_cast<uint32_t>(arow);
    }
    auto lcell = [&](int z) -> uint32_t { return lxy + a.pz * static_cast<uint32_t>(z); };

    // ---- psi cells of a ring lane ----------------------------------------------------------
    // psi_row: stored row of the psi cell, -1 outside a bounded y (psi 0), -2 no psi cell (no
    // store). Every ring lane loads, unmasked as the moment loads: a lane past the psi ring reads
    // the ring's last cell, a psi cell outside a bounded y the nearest ghost plane.
    int psi_row[NPASS];
    uint32_t psi_xy[NPASS];
    float psi_r[NPASS], psi_b[NPASS];
#pragma unroll
    for (int k = 0; k < NPASS; ++k) {
        const int ci = (t - NOWN) + k * NRING;
        const int ca = ci < NPSI ? ci : NPSI - 1;
        const int px = ca % PX - 2, py = ca / PX - 2;
        const int rr_ = r0 + py;
        int arow;
        if constexpr (YBC == kYPeriodic) arow = wrap_near(rr_, ny);
        else arow = rr_ < -1 ? 0 : (rr_ > ny ? ny + 1 : rr_ + 1);
        psi_xy[k] = static_cast<uint32_t>(wrap_near(x0 + px, nx)) + a.py * static_cast<uint32_t>(arow);
        if (!ring || ci >= NPSI) psi_row[k] = -2;
        else if constexpr (YBC == kYPeriodic) psi_row[k] = arow;
        else psi_row[k] = (rr_ >= -1 && rr_ <= ny) ? arow : -1;
    }
    auto psi_load_into = [&](int z, float (&r)[NPASS], float (&b)[NPASS]) {
        const uint32_t zo = a.pz * static_cast<uint32_t>(wrapz(z));
#pragma unroll
        for (int k = 0; k < NPASS; ++k) {
            r[k] = Abuf.ld(0, psi_xy[k] + zo);
            b[k] = Abuf.ld(1, psi_xy[k] + zo);
        }
    };
    auto psi_store_from = [&](int slot, const float (&r)[NPASS], const float (&b)[NPASS]) {
#pragma unroll
        for (int k = 0; k < NPASS; ++k) {
            if (psi_row[k] > -2) {
                const int ci = (t - NOWN) + k * NRING;
                const float v = psi_row[k] >= 0 ? (r[k] - b[k]) / (r[k] + b[k]) : 0.0f;
                sh.psi[slot][ci / PX][ci % PX] = v;
            }
        }
    };
    auto psi_load = [&](int z) { psi_load_into(z, psi_r, psi_b); };
    auto psi_store = [&](int slot) { psi_store_from(slot, psi_r, psi_b); };
    auto psi_slot = [&](int z) { return (z - (z0 - 2)) % kPsiSlots; };

    // psi of neighbour c_j of this lane's source (forces.cuh:43-50; bounded y limited as
    // mathUtilities.cuh:25-33) for source plane p.
    auto psi_reader = [&](int p) {
        const int sm = psi_slot(p - 1), s0 = psi_slot(p), sp = psi_slot(p + 1);
        return [&, sm, s0, sp](auto J) -> float {
            constexpr int j = decltype(J)::value;
            constexpr int cx = kCx[j], cy = kCy[j], cz = kCz[j];
            int prow = sy + cy;
            if constexpr (YBC != kYPeriodic) {
                if constexpr (cy > 0) {
                    if (rsrc + cy > ny - 1) prow = (ny - 1) - r0;
                }
            }
            const int slot = cz == 0 ? s0 : (cz > 0 ? sp : sm);
            return sh.psi[slot][prow + 2][sx + cx + 2];
        };
    };

    Held<Q> H0, H1, H2;
#pragma unroll
    for (int i = 0; i < Q; ++i) H0.v[i] = H1.v[i] = H2.v[i] = make_float2(0.0f, 0.0f);

    // ---- prologue: psi planes z0-2 .. z0+1 (ring lanes), first moments ---------------------
    // All four planes' psi loads and the first source plane's moments are in flight at once, so
    // the prologue costs one memory latency (measured 3% on the batched step of 256 small
    // domains, 1% on a 128^3 domain).
    Mom nxt{};
    if (ring) {
        float pr[4][NPASS], pb[4][NPASS];
#pragma unroll
        for (int q = 0; q < 4; ++q) psi_load_into(z0 - 2 + q, pr[q], pb[q]);
#if CGLBM_MARCH_PREFETCH
        nxt = Abuf.mom(lcell(wrapz(z0 - 1)));
#endif
#pragma unroll
        for (int q = 0; q < 4; ++q) psi_store_from(q, pr[q], pb[q]);
    } else {
#if CGLBM_MARCH_PREFETCH
        nxt = Abuf.mom(lcell(wrapz(z0 - 1)));
#endif
    }
    __syncthreads();

    const int kend = Lc + 1;
    // One step: source plane p = z0 - 1 + k, target plane p - 1 completes. OWN selects the own
    // waves' code at compile time; HA, HB, HC hold the populations received so far by targets
    // p - 1, p and p + 1 (own waves only).
    auto step = [&](int k, auto OWN, Held<Q> &HA, Held<Q> &HB, Held<Q> &HC) {
        constexpr bool own = decltype(OWN)::value;
        const int p = z0 - 1 + k;
        const int pz = wrapz(p);
        const bool edge = (k == 0) || (k == kend);
        const int xb = NBUF == 2 ? (k & 1) : 0;
#if CGLBM_MARCH_PREFETCH
        // The next plane's moments are prefetched; the last step loads its own plane again (cached
        // and unused): an unconditional load keeps the prefetch out of a merge with the previous
        // values, which would wait for the loads.
        const Mom cur = nxt;
        nxt = Abuf.mom(lcell(wrapz(k < kend ? p + 1 : p)));
#else
        const Mom cur = Abuf.mom(lcell(pz));
#endif
        // the adjacent interior cell of a ghost source is loaded in the ghost branch, so its 11
        // values are live only there

        // ---- sources ----------------------------------------------------------------------
        [[maybe_unused]] bool ion = false;  // this lane's source took the interface form (REC)
        if (srow >= 0) {
            using XL = XferLayout<Q, TY>;
            constexpr int MG = XL::M;
            auto lds_push = [&](auto I, float fr, float fb) {
                constexpr int i = decltype(I)::value;
                if constexpr (cross_col(i)) {
                    const int txx = sx + kCx[i], tyy = sy + kCy[i];
                    if (txx >= -MG && txx < TX + MG && tyy >= -MG && tyy < TY + MG)
                        sh.xfer[xb][XL::at(slot_of(i), txx, tyy)] = make_float2(fr, fb);
                }
            };
            auto own_sink = [&](auto I, float fr, float fb) {
                constexpr int i = decltype(I)::value;
                if constexpr (cross_col(i)) {
                    // at most one cell outside the tile: a location nobody reads (XferLayout)
                    sh.xfer[xb][XL::at(slot_of(i), sx + kCx[i], sy + kCy[i])] = make_float2(fr, fb);
                } else if constexpr (kCz[i] == 0) {
                    HB.v[i] = make_float2(fr, fb);
                } else if constexpr (kCz[i] > 0) {
                    HC.v[i] = make_float2(fr, fb);
                } else {
                    HA.v[i] = make_float2(fr, fb);
                }
            };
            auto run = [&](auto MASKC, auto RESTC, auto &sink) {
                constexpr uint32_t M = decltype(MASKC)::value;
                constexpr bool R = decltype(RESTC)::value;
                if (YBC != kYPeriodic && sghost) {
                    emit_ghost<Q, M, R>(cur, Abuf.mom(cell(gx, frow, pz)), c, sink);
                } else {
                    Iface f = iface_begin(cur);
                    // D3Q19: wave-uniform choice; a wave with any interface source runs the
                    // interface form on every lane (non-interface lanes add no Omega2 or
                    // recolouring term) instead of both forms under divergent masks. D3Q27
                    // (longer interface form, higher register pressure) branches per lane.
                    const bool iface = Q == 19 ? (__builtin_amdgcn_ballot_w64(f.on) != 0) : f.on;
                    if constexpr (REC) ion = f.on;
                    if (iface) {
                        colour_gradient<Q>(psi_reader(p), f.Fx, f.Fy, f.Fz);
                        iface_finish<Q>(f, cur, c, a.conservative);
                        emit_regular<Q, M, R, FORCED, true>(cur, c, f, sink);
                    } else {
                        emit_regular<Q, M, R, FORCED, false>(cur, c, f, sink);
                    }
                }
            };
            // one member per pair (emit_one): the halo sources push only the member that enters the
            // tile (c_y toward the tile for the row ring, c_x for the column ring: a lane on the far
            // side takes the other member), the first and last source plane of a chunk only the
            // member that reaches the chunk's first or last target plane (c_z = +1, -1)
            const uint32_t ring_flip = (role == kRoleRows ? sy == TY : sx == TX) ? 0x80000000u : 0u;
            auto ring_push1 = [&](auto I, auto OMC, float fr, float fb) {
                constexpr int i = decltype(I)::value, o = i + 1;
                static_assert(cross_col(i) && cross_col(o), "ring pairs leave the column");
                const bool takeO = decltype(OMC)::value != (ring_flip != 0u);
                const int txx = sx + (takeO ? kCx[o] : kCx[i]), tyy = sy + (takeO ? kCy[o] : kCy[i]);
                if (txx >= -MG && txx < TX + MG && tyy >= -MG && tyy < TY + MG)
                    sh.xfer[xb][XL::at(takeO ? slot_of(o) : slot_of(i), txx, tyy)] = make_float2(fr, fb);
            };
            auto own_push1 = [&](auto I, auto OMC, float fr, float fb) {
                constexpr int d = decltype(I)::value + (decltype(OMC)::value ? 1 : 0);
                own_sink(std::integral_constant<int, d>{}, fr, fb);
            };
            auto run1 = [&](auto MASKC, auto OMASKC, uint32_t flip, auto &sink3, auto &sink1) {
                constexpr uint32_t M = decltype(MASKC)::value;
                constexpr uint32_t OM = decltype(OMASKC)::value;
                if (YBC != kYPeriodic && sghost) {
                    emit_ghost<Q, M, false>(cur, Abuf.mom(cell(gx, frow, pz)), c, sink3);
                } else {
                    Iface f = iface_begin(cur);
                    const bool iface = Q == 19 ? (__builtin_amdgcn_ballot_w64(f.on) != 0) : f.on;
                    if (iface) {
                        colour_gradient<Q>(psi_reader(p), f.Fx, f.Fy, f.Fz);
                        iface_finish<Q>(f, cur, c, a.conservative);
                        emit_one<Q, M, OM, true>(cur, c, f, flip, sink1);
                    } else {
                        emit_one<Q, M, OM, false>(cur, c, f, flip, sink1);
                    }
                }
            };
            constexpr uint32_t NX_ = pair_mask_neg<Q>(0), NY_ = pair_mask_neg<Q>(1), NZ_ = pair_mask_neg<Q>(2);
            auto ring_run = [&](auto MASKC, auto OMASKC) {
                if constexpr (FORCED) run(MASKC, std::false_type{}, lds_push);
                else run1(MASKC, OMASKC, ring_flip, lds_push, ring_push1);
            };
            if constexpr (own) {
                if (!edge) {
                    run(std::integral_constant<uint32_t, ALL>{}, std::true_type{}, own_sink);
                } else if constexpr (FORCED) {
                    run(std::integral_constant<uint32_t, MZ>{}, std::false_type{}, own_sink);
                } else if (k == 0) {  // first source plane: the c_z = +1 members
                    run1(std::integral_constant<uint32_t, MZ>{}, std::integral_constant<uint32_t, MZ & NZ_>{}, 0u,
                         own_sink, own_push1);
                } else {  // last source plane: the c_z = -1 members
                    run1(std::integral_constant<uint32_t, MZ>{}, std::integral_constant<uint32_t, MZ & ~NZ_>{}, 0u,
                         own_sink, own_push1);
                }
            } else if (RW == 2) {
                if (role == kRoleRows) {
                    if (edge) ring_run(std::integral_constant<uint32_t, MY & MZ>{}, std::integral_constant<uint32_t, NY_>{});
                    else ring_run(std::integral_constant<uint32_t, MY>{}, std::integral_constant<uint32_t, NY_>{});
                } else {
                    if (edge) ring_run(std::integral_constant<uint32_t, MX & MZ>{}, std::integral_constant<uint32_t, NX_>{});
                    else ring_run(std::integral_constant<uint32_t, MX>{}, std::integral_constant<uint32_t, NX_>{});
                }
            } else if (role == kRoleRows) {
                if (half == 0) {
                    if (edge) ring_run(std::integral_constant<uint32_t, mask_half(MY & MZ, 0)>{}, std::integral_constant<uint32_t, NY_>{});
                    else ring_run(std::integral_constant<uint32_t, mask_half(MY, 0)>{}, std::integral_constant<uint32_t, NY_>{});
                } else {
                    if (edge) ring_run(std::integral_constant<uint32_t, mask_half(MY & MZ, 1)>{}, std::integral_constant<uint32_t, NY_>{});
                    else ring_run(std::integral_constant<uint32_t, mask_half(MY, 1)>{}, std::integral_constant<uint32_t, NY_>{});
                }
            } else {
                if (half == 0) {
                    if (edge) ring_run(std::integral_constant<uint32_t, mask_half(MX & MZ, 0)>{}, std::integral_constant<uint32_t, NX_>{});
                    else ring_run(std::integral_constant<uint32_t, mask_half(MX, 0)>{}, std::integral_constant<uint32_t, NX_>{});
                } else {
                    if (edge) ring_run(std::integral_constant<uint32_t, mask_half(MX & MZ, 1)>{}, std::integral_constant<uint32_t, NX_>{});
                    else ring_run(std::integral_constant<uint32_t, mask_half(MX, 1)>{}, std::integral_constant<uint32_t, NX_>{});
                }
            }
        }
        if constexpr (REC) {
            if (own && !edge) {  // wave-uniform condition
                const bool any = __builtin_amdgcn_ballot_w64(ion) != 0;
                if ((t & 63) == 0) {
                    const int ntx = (nx + TX - 1) / TX;
                    const size_t col = static_cast<size_t>(tiy) * ntx + tix;
                    a.wmap[(col * nz + p) * (TY / 2) + t / 64] = any ? 1 : 0;
                }
            }
        }
        __syncthreads();

        // ---- ring lanes: psi two planes ahead (plane p + 3, first read at step k + 2) ------
        // (loading it before the sources instead, to hide the load, measured no faster on
        // single domains and 2.4% slower on batches: the registers it holds across the sources)
        if (!own && k + 3 <= kend + 1) {
            psi_load(p + 3);
            psi_store(psi_slot(p + 3));
        }

        // ---- targets ----------------------------------------------------------------------
        if constexpr (own) {
            unroll<0, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                if constexpr (cross_col(i)) {
                    const float2 v = sh.xfer[xb][XferLayout<Q, TY>::at(slot_of(i), sx, sy)];
                    if constexpr (kCz[i] == 0) HB.v[i] = v;
                    else if constexpr (kCz[i] > 0) HC.v[i] = v;
                    else HA.v[i] = v;
                }
            });
            if (k >= 2 && own_target) {
                const int zt = wrapz(p - 1);
                float out[11];
                accumulate<Q, FORCED>([&](auto I) { return HA.v[decltype(I)::value]; }, c, out);
                const uint32_t id = cell(gx, srow, zt);
#pragma unroll
                for (int f = 0; f < 11; ++f) Bbuf.st(f, id, out[f]);
                if constexpr (YBC != kYPeriodic) boundary_planes(a, Abuf, Bbuf, gx, zt, rsrc, out);
            }
        }
        if constexpr (NBUF == 1) __syncthreads();
    };
    // Own waves: after a step, target p becomes p - 1 and p + 1 becomes p, so the three sets
    // rotate with period 3; the loop is unrolled by three with the sets renamed instead of
    // copied. Ring waves hold no targets. Both loops run kend + 1 steps and so the same
    // barriers.
    if (role == kRoleOwn) {
        for (int k = 0; k <= kend; k += 3) {
            step(k, std::true_type{}, H0, H1, H2);
            if (k + 1 > kend) break;
            step(k + 1, std::true_type{}, H1, H2, H0);
            if (k + 2 > kend) break;
            step(k + 2, std::true_type{}, H2, H0, H1);
        }
    } else {
        for (int k = 0; k <= kend; ++k) step(k, std::false_type{}, H0, H1, H2);
    }
    if constexpr (NBUF == 2) __syncthreads();
}

template <int Q, int YBC, int TY, int NBUF, bool FORCED, int ST, bool REC>
__global__ void __launch_bounds__(block_threads<TY>()) CGLBM_MARCH_OCC
    march_kernel(MarchArgs a, int ntx, int nty, int nchunk, int nblocks) {
    __shared__ Shared<Q, TY, NBUF> sh;
    [[maybe_unused]] uint64_t t_start = 0;
    if constexpr (REC) t_start = wall_clock64();
    const int b = blockIdx.x;
    const int per = gridDim.x / 8;
    const int logical = (b % 8) * per + b / 8;  // XCD-contiguous tile ranges
    if (logical >= nblocks) return;
    int tix, tiy, z0, z1;
    if (a.items != nullptr) {
        const int4 it = a.items[logical];
        tix = it.x;
        tiy = it.y;
        z0 = it.z;
        z1 = it.w;
        if (z0 >= z1) return;  // idle item
    } else {
        tix = logical % ntx;
        tiy = (logical / ntx) % nty;
        int chunk = logical / (ntx * nty), n = nchunk, lo = a.zlo, hi = a.zhi;
        if (a.zface > 0) {  // faces of a slab: half the chunks per face
            n = nchunk / 2;
            if (chunk < n) hi = a.zlo + a.zface;
            else { lo = a.zhi - a.zface; chunk -= n; }
        }
        const int span = hi - lo;
        z0 = lo + static_cast<int>((static_cast<int64_t>(chunk) * span) / n);
        z1 = lo + static_cast<int>((static_cast<int64_t>(chunk + 1) * span) / n);
    }
    march_tile<Q, YBC, TY, NBUF, FORCED, ST, REC>(a, sh, tix, tiy, z0, z1);
    if constexpr (REC) {
        // march_tile ends with the work-group synchronised: every wave is done with the item
        if (a.itime != nullptr && threadIdx.x == 0) a.itime[logical] = static_cast<uint32_t>(wall_clock64() - t_start);
    }
}

// Batched form: many domains of one kernel class in one launch. Each work item names a domain
// descriptor and a tile/chunk of it; the grid has one block per item (rounded up to a multiple of
// 8) and blocks take items in XCD-contiguous ranges. The step counter (parity of the ping-pong and
// the inlet perturbation key) is base[0] + k, base living in device memory so a captured
// multi-step graph can be replayed.
template <int Q, int YBC, int TY, int NBUF, bool FORCED, int ST>
__global__ void __launch_bounds__(block_threads<TY>()) CGLBM_MARCH_OCC
    march_batch_kernel(const BatchDesc *__restrict__ descs, const int4 *__restrict__ items, int nitems,
                       const int64_t *__restrict__ base, int k) {
    __shared__ Shared<Q, TY, NBUF> sh;
    const int nper = (nitems + 7) / 8;
    const int b = blockIdx.x;
    const int logical = (b % 8) * nper + b / 8;
    if (b >= nper * 8 || logical >= nitems) return;
    const int4 it = items[logical];
    const BatchDesc &d = descs[it.x];
    const int64_t kk = base[0] + k;
    MarchArgs a = d.args;
    const int64_t steps_done = d.step0 + kk;  // member steps completed before this launch
    const int parity = d.cur0 ^ static_cast<int>(kk & 1);
    a.A0 = parity ? d.buf1 : d.buf0;
    a.B0 = parity ? d.buf0 : d.buf1;
    a.step = steps_done + 1;
    a.items = nullptr;
    a.itime = nullptr;
    const int z0 = static_cast<int>((static_cast<int64_t>(it.w) * a.nz) / d.nchunk);
    const int z1 = static_cast<int>((static_cast<int64_t>(it.w + 1) * a.nz) / d.nchunk);
    march_tile<Q, YBC, TY, NBUF, FORCED, ST>(a, sh, it.y, it.z, z0, z1);
}

template <int Q, int YBC, int TY, int NBUF, int ST>
hipError_t launch_shape(const MarchArgs &a, int L, hipStream_t s) {
    const int ntx = (a.nx + TX - 1) / TX;
    const int nty = (a.ny + TY - 1) / TY;
    const int nchunk = a.zface > 0 ? 2 * ((a.zface + L - 1) / L) : (a.zhi - a.zlo + L - 1) / L;
    const int nblocks = a.items != nullptr ? a.nitems : ntx * nty * nchunk;
    const int grid = (nblocks + 7) / 8 * 8;
    // The body-force variant exists for bounded y (the RTI walls); a periodic forced domain takes
    // the gather path (march_supports).
    if constexpr (YBC != kYPeriodic) {
        if (a.forced) {
            if (a.wmap)
                march_kernel<Q, YBC, TY, NBUF, true, ST, true>
                    <<<grid, block_threads<TY>(), 0, s>>>(a, ntx, nty, nchunk, nblocks);
            else
                march_kernel<Q, YBC, TY, NBUF, true, ST, false>
                    <<<grid, block_threads<TY>(), 0, s>>>(a, ntx, nty, nchunk, nblocks);
            return hipGetLastError();
        }
    }
    if (a.wmap)
        march_kernel<Q, YBC, TY, NBUF, false, ST, true><<<grid, block_threads<TY>(), 0, s>>>(a, ntx, nty, nchunk, nblocks);
    else
        march_kernel<Q, YBC, TY, NBUF, false, ST, false><<<grid, block_threads<TY>(), 0, s>>>(a, ntx, nty, nchunk, nblocks);
    return hipGetLastError();
}

// Double-buffered exchange must fit the 160 KB LDS of a CU.
template <int Q, int TY>
constexpr bool fits2() {
    return shared_bytes<Q, TY, 2>() <= 160 * 1024;
}

template <int Q, int YBC, int ST>
hipError_t launch_q(const MarchArgs &a, MarchShape m, hipStream_t s) {
    const bool one = m.nbuf == 1;
    switch (m.ty) {
    case 4: return one ? launch_shape<Q, YBC, 4, 1, ST>(a, m.L, s) : launch_shape<Q, YBC, 4, 2, ST>(a, m.L, s);
    case 16:
        if constexpr (fits2<Q, 16>()) {
            if (!one) return launch_shape<Q, YBC, 16, 2, ST>(a, m.L, s);
        }
        return launch_shape<Q, YBC, 16, 1, ST>(a, m.L, s);
    default: return one ? launch_shape<Q, YBC, 8, 1, ST>(a, m.L, s) : launch_shape<Q, YBC, 8, 2, ST>(a, m.L, s);
    }
}

// One stencil and storage format (instantiated per translation unit: march_q19.hip, ...).
template <int Q, int ST>
hipError_t launch_march_qs(const MarchArgs &a, MarchShape m, hipStream_t s) {
    return a.ybc == kYPeriodic ? launch_q<Q, kYPeriodic, ST>(a, m, s) : launch_q<Q, kYJet, ST>(a, m, s);
}

template <int Q, int YBC, int TY, int NBUF, bool FORCED, int ST>
hipError_t launch_batch_shape(const BatchLaunch &L, hipStream_t s) {
    march_batch_kernel<Q, YBC, TY, NBUF, FORCED, ST>
        <<<L.grid, block_threads<TY>(), 0, s>>>(L.descs, L.items, L.nitems, L.base, L.k);
    return hipGetLastError();
}

template <int Q, int YBC, bool FORCED, int ST>
hipError_t launch_batch_qy(const BatchLaunch &L, hipStream_t s) {
    const bool one = L.shape.nbuf == 1;
    switch (L.shape.ty) {
    case 4:
        return one ? launch_batch_shape<Q, YBC, 4, 1, FORCED, ST>(L, s)
                   : launch_batch_shape<Q, YBC, 4, 2, FORCED, ST>(L, s);
    case 16:
        if constexpr (fits2<Q, 16>()) {
            if (!one) return launch_batch_shape<Q, YBC, 16, 2, FORCED, ST>(L, s);
        }
        return launch_batch_shape<Q, YBC, 16, 1, FORCED, ST>(L, s);
    default:
        return one ? launch_batch_shape<Q, YBC, 8, 1, FORCED, ST>(L, s)
                   : launch_batch_shape<Q, YBC, 8, 2, FORCED, ST>(L, s);
    }
}

template <int Q, int ST>
hipError_t launch_batch_qs(const BatchLaunch &L, hipStream_t s) {
    if (L.ybc == kYPeriodic) return launch_batch_qy<Q, kYPeriodic, false, ST>(L, s);
    if (L.forced) return launch_batch_qy<Q, kYJet, true, ST>(L, s);
    return launch_batch_qy<Q, kYJet, false, ST>(L, s);
}

}  // namespace march
}  // namespace cglbm
