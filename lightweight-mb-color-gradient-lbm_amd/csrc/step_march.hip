// Dispatch of the z-marching step (kernels in march_kernel.inl, instantiated per stencil in
// march_q19.hip and march_q27.hip) and its default tile shapes.
#include <cstdlib>

#include "kernels.hpp"
#include "storage.hpp"

namespace cglbm {

namespace {
int env_int(const char *name, int lo, int hi) {
    const char *e = std::getenv(name);
    const int v = e ? std::atoi(e) : 0;
    return (v >= lo && v <= hi) ? v : 0;
}
}  // namespace

MarchShape march_default_shape(const StepArgs &a, int q) {
    // Untuned default, size-aware: the first shape of the list that gives every CU a work-group
    // (256); if none does (small domains), the one with the most work-groups. 16-row tiles with
    // shorter chunks come before 4-row tiles: on domains of 0.26 to 1.6 M cells, where 16-row
    // tiles with 16-plane chunks leave CUs idle, 16 rows x 8 planes ran 20-27% faster than 4 rows
    // x 16 planes in D3Q19 (drops in 128 x 128 x 48/64/96, 128 x 64 x 128, 64 x 128 x 128, 96^3)
    // and twice as fast in D3Q27 (128 x 128 x 64), and the autotuner picks 16 rows x 2 planes on
    // the 64^3 drop. Two exchange buffers (one barrier per
    // plane; D3Q27 16-row tiles keep one, effective_shape): measured as fast or faster than one in
    // every case swept (128 x 512 x 128 Rayleigh-Taylor, 16 rows x 16 planes: 458 against 508 us).
    // CGLBM_MARCH_TY / CGLBM_MARCH_L / CGLBM_MARCH_NBUF override the choice for tuning sweeps.
    static constexpr int kCand[9][2] = {{16, 16}, {16, 8}, {16, 4}, {16, 2}, {4, 16}, {4, 8}, {4, 4}, {4, 2}, {4, 1}};
    constexpr int64_t kFill = 256;
    (void)q;
    const auto &cand = kCand;
    MarchShape m{cand[0][0], cand[0][1], 2};
    int64_t best_blocks = -1;
    for (const auto &c : cand) {
        const int L = c[1] < a.nz ? c[1] : a.nz;
        const int64_t blocks =
            int64_t((a.nx + 31) / 32) * ((a.ny + c[0] - 1) / c[0]) * ((a.nz + L - 1) / L);
        if (blocks >= kFill) {
            m = MarchShape{c[0], L, 2};
            best_blocks = -2;
            break;
        }
        if (blocks > best_blocks) {
            best_blocks = blocks;
            m = MarchShape{c[0], L, 2};
        }
    }
    if (const int ty = env_int("CGLBM_MARCH_TY", 1, 64)) m.ty = ty;
    if (const int L = env_int("CGLBM_MARCH_L", 1, 4096)) m.L = L;
    if (const int nb = env_int("CGLBM_MARCH_NBUF", 1, 2)) m.nbuf = nb;
    if (m.L > a.nz) m.L = a.nz;
    return m;
}

bool march_args(const StepArgs &a, MarchArgs &m) {
    // last addressed cell of a field, in elements; cell offsets stay 32-bit
    const int64_t nyt = a.ybc == kYPeriodic ? a.ny : a.ny + 2;
    const int64_t last = (a.nx - 1) + int64_t(a.py) * (nyt - 1) + int64_t(a.pz) * (a.nz - 1);
    if (last >= (int64_t(1) << 32)) return false;
    // Fields at one stride from one base (one allocation per state) and a state below 4 GiB: the
    // buffer form. Otherwise (fields in separate allocations, or a larger state): one 64-bit
    // pointer per field.
    const auto addr = [](const float *p) { return static_cast<int64_t>(reinterpret_cast<uintptr_t>(p)); };
    const int64_t dA = addr(a.A[1]) - addr(a.A[0]), dB = addr(a.B[1]) - addr(a.B[0]);
    bool uniform = dA > 0 && dB > 0 && dA % 4 == 0 && dB % 4 == 0 && dA / 4 < (int64_t(1) << 32) &&
                   dB / 4 < (int64_t(1) << 32);
    for (int f = 2; f < 11 && uniform; ++f)
        uniform = addr(a.A[f]) - addr(a.A[0]) == dA * f && addr(a.B[f]) - addr(a.B[0]) == dB * f;
    const int64_t fsA = uniform ? dA / 4 : 0, fsB = uniform ? dB / 4 : 0;
    const int64_t spanA = (fsA * 10 + last + 1) * 4, spanB = (fsB * 10 + last + 1) * 4;
    const bool buffer = uniform && spanA < (int64_t(1) << 32) && spanB < (int64_t(1) << 32);
    m.A0 = a.A[0];
    m.B0 = a.B[0];
    m.fsA = static_cast<uint32_t>(fsA);
    m.fsB = static_cast<uint32_t>(fsB);
    m.spanA = buffer ? static_cast<uint32_t>(spanA) : 0u;
    m.spanB = buffer ? static_cast<uint32_t>(spanB) : 0u;
    for (int f = 0; f < 11; ++f) {
        m.Af[f] = a.A[f];
        m.Bf[f] = a.B[f];
    }
    m.nx = a.nx;
    m.ny = a.ny;
    m.nz = a.nz;
    m.py = a.py;
    m.pz = a.pz;
    m.ybc = a.ybc;
    m.outlet = a.outlet;
    m.conservative = a.conservative;
    m.forced = a.forced;
    m.c = a.c;
    m.jet_x0 = a.jet_x0;
    m.jet_z0 = a.jet_z0;
    m.jet_r = a.jet_r;
    m.jet_u = a.jet_u;
    m.jet_noise = a.jet_noise;
    m.seed = a.seed;
    m.step = a.step;
    m.store = buffer ? kStoreF32 : kStoreF32Ptr;
    m.zlo = 0;
    m.zhi = a.nz;
    m.zoff = 0;
    m.zface = 0;
    m.items = nullptr;
    m.nitems = 0;
    m.wmap = nullptr;
    m.itime = nullptr;
    return true;
}

bool march_args_f16(const StepArgs &a, const float *A0, float *B0, int64_t fs, MarchArgs &m) {
    if (fs <= 0 || fs % 64 != 0 || state_bytes(kStoreF16, fs) >= (int64_t(1) << 32)) return false;
    StepArgs u = a;  // uniform-stride stand-in so march_args fills the geometry and physics
    for (int f = 0; f < 11; ++f) {
        u.A[f] = A0 + f * fs;
        u.B[f] = B0 + f * fs;
    }
    if (!march_args(u, m)) return false;
    m.store = kStoreF16;
    m.spanA = m.spanB = static_cast<uint32_t>(state_bytes(kStoreF16, fs));
    return true;
}

bool march_supports(const StepArgs &a, int q) {
    if (a.forced && a.ybc == kYPeriodic) return false;  // no periodic body-force variant
    MarchArgs m;
    return (q == 19 || q == 27) && a.nx >= 3 && a.ny >= 3 && a.nz >= 3 && march_args(a, m);
}

hipError_t launch_march_shape(const StepArgs &a, int q, MarchShape m, hipStream_t s) {
    if ((m.ty != 4 && m.ty != 8 && m.ty != 16) || m.L < 1 || (m.nbuf != 1 && m.nbuf != 2)) return hipErrorInvalidValue;
    MarchArgs ma;
    if (!march_args(a, ma)) return hipErrorInvalidValue;
    return launch_march_args(ma, q, m, s);
}

hipError_t launch_march_args(const MarchArgs &a, int q, MarchShape m, hipStream_t s) {
    if ((m.ty != 4 && m.ty != 8 && m.ty != 16) || m.L < 1 || (m.nbuf != 1 && m.nbuf != 2)) return hipErrorInvalidValue;
    if (a.store == kStoreF32Ptr) {
        if (q == 19) return launch_march_q19_f32ptr(a, m, s);
        if (q == 27) return launch_march_q27_f32ptr(a, m, s);
        return hipErrorInvalidValue;
    }
    if (q == 19) return a.store == kStoreF16 ? launch_march_q19_f16(a, m, s) : launch_march_q19_f32(a, m, s);
    if (q == 27) return a.store == kStoreF16 ? launch_march_q27_f16(a, m, s) : launch_march_q27_f32(a, m, s);
    return hipErrorInvalidValue;
}

hipError_t launch_batch(const BatchLaunch &L, int q, hipStream_t s) {
    if (L.store == kStoreF32Ptr) return hipErrorInvalidValue;  // batch members stay below 4 GiB per state
    if (q == 19) return L.store == kStoreF16 ? launch_batch_q19_f16(L, s) : launch_batch_q19_f32(L, s);
    if (q == 27) return L.store == kStoreF16 ? launch_batch_q27_f16(L, s) : launch_batch_q27_f32(L, s);
    return hipErrorInvalidValue;
}

hipError_t launch_march(const StepArgs &a, int q, hipStream_t s) {
    return launch_march_shape(a, q, march_default_shape(a, q), s);
}

}  // namespace cglbm
