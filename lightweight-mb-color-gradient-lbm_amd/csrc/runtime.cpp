// Host runtime behind include/cglbm.h: argument checking, device memory, streams, the step
// loop, initial conditions, copies, diagnostics and VTI output.
//
// Reference boundary replaced here (file:line in each entry point's comment): launch.cuh,
// memory.cuh, main.cu:38-55, io/vtiWriter.cuh, utilities/mlups_count.cuh.
#include "cglbm.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <array>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "lattice.hpp"
#include "state.hpp"
#include "storage.hpp"

using namespace cglbm;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(CGLBM_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

}  // namespace

namespace cglbm {

int set_error(int code, const char *msg) { return fail(code, "%s", msg); }

int validate_config(const cglbm_config *c) {
    if (!c) return fail(CGLBM_EINVAL, "null config");
    if (c->q != 19 && c->q != 27) return fail(CGLBM_EINVAL, "q must be 19 or 27 (got %d)", c->q);
    if (c->nx < 3 || c->ny < 3 || c->nz < 3)
        return fail(CGLBM_EINVAL, "every interior extent must be >= 3 (got %d x %d x %d)", c->nx, c->ny, c->nz);
    if (c->ybc < CGLBM_Y_PERIODIC || c->ybc > CGLBM_Y_WALLS) return fail(CGLBM_EINVAL, "unknown ybc %d", c->ybc);
    if (c->outlet != CGLBM_OUTLET_CONVECTIVE && c->outlet != CGLBM_OUTLET_ZERO_GRADIENT)
        return fail(CGLBM_EINVAL, "unknown outlet %d", c->outlet);
    if (!(c->tau > 0.5f) || !std::isfinite(c->tau)) return fail(CGLBM_EINVAL, "tau must be > 0.5 (got %g)", c->tau);
    if (!(c->sigma >= 0.0f) || !std::isfinite(c->sigma)) return fail(CGLBM_EINVAL, "sigma must be >= 0 (got %g)", c->sigma);
    if (!(c->beta >= 0.0f && c->beta <= 1.0f)) return fail(CGLBM_EINVAL, "beta must be in [0, 1] (got %g)", c->beta);
    if (!std::isfinite(c->gravity) || !std::isfinite(c->atwood)) return fail(CGLBM_EINVAL, "non-finite body force");
    if (!(c->jet_noise >= 0.0f && c->jet_noise < 1.0f)) return fail(CGLBM_EINVAL, "jet_noise must be in [0, 1)");
    const int64_t nyt = c->ybc == CGLBM_Y_PERIODIC ? c->ny : c->ny + 2;
    const int64_t cells = static_cast<int64_t>(c->nx) * nyt * c->nz;
    if (cells >= (int64_t(1) << 32)) return fail(CGLBM_EINVAL, "domain too large for 32-bit cell offsets");
    return CGLBM_OK;
}

StepArgs make_step_args(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *c,
                        const cglbm_layout *L, int64_t step) {
    StepArgs a{};
    float *const *pa = reinterpret_cast<float *const *>(A);
    float *const *pb = reinterpret_cast<float *const *>(B);
    for (int f = 0; f < 11; ++f) {
        a.A[f] = pa[f] + L->origin;
        a.B[f] = pb[f] + L->origin;
    }
    a.nx = c->nx;
    a.ny = c->ny;
    a.nz = c->nz;
    a.nyt = c->ybc == CGLBM_Y_PERIODIC ? c->ny : c->ny + 2;
    a.py = static_cast<uint32_t>(L->pitch_y);
    a.pz = static_cast<uint32_t>(L->pitch_z);
    a.ybc = c->ybc;
    a.outlet = c->outlet;
    a.conservative = c->conservative_omega2 ? 1 : 0;
    a.forced = c->gravity != 0.0f ? 1 : 0;
    a.c = make_coeffs(c->tau, c->sigma, c->beta, c->gravity, c->atwood);
    a.jet_x0 = c->jet_x0;
    a.jet_z0 = c->jet_z0;
    a.jet_r = c->jet_radius;
    a.jet_u = c->jet_velocity;
    a.jet_noise = c->jet_noise;
    a.seed = c->seed;
    a.step = step;
    return a;
}

BcArgs make_bc_args(const cglbm_moments *B, const cglbm_moments *Aold, const cglbm_config *c,
                    const cglbm_layout *L, int64_t step) {
    BcArgs a{};
    float *const *pb = reinterpret_cast<float *const *>(B);
    float *const *po = reinterpret_cast<float *const *>(Aold);
    for (int f = 0; f < 11; ++f) {
        a.B[f] = pb[f] + L->origin;
        a.Aold[f] = po[f] + L->origin;
    }
    a.nx = c->nx;
    a.ny = c->ny;
    a.nz = c->nz;
    a.nyt = c->ybc == CGLBM_Y_PERIODIC ? c->ny : c->ny + 2;
    a.py = static_cast<uint32_t>(L->pitch_y);
    a.pz = static_cast<uint32_t>(L->pitch_z);
    a.ybc = c->ybc;
    a.outlet = c->outlet;
    a.jet_x0 = c->jet_x0;
    a.jet_z0 = c->jet_z0;
    a.jet_r = c->jet_radius;
    a.jet_u = c->jet_velocity;
    a.jet_noise = c->jet_noise;
    a.seed = c->seed;
    a.step = step;
    return a;
}

int check_layout(const cglbm_config *c, const cglbm_layout *L) {
    if (!L) return fail(CGLBM_EINVAL, "null layout");
    const int64_t nyt = c->ybc == CGLBM_Y_PERIODIC ? c->ny : c->ny + 2;
    if (L->origin < 0 || L->pitch_y < c->nx || L->pitch_z < L->pitch_y * nyt)
        return fail(CGLBM_EINVAL, "layout pitches do not cover the %d x %lld x %d cells", c->nx,
                    static_cast<long long>(nyt), c->nz);
    const int64_t last = L->origin + (c->nx - 1) + L->pitch_y * (nyt - 1) + L->pitch_z * (c->nz - 1);
    if (last >= (int64_t(1) << 32)) return fail(CGLBM_EINVAL, "layout exceeds 32-bit offsets");
    return CGLBM_OK;
}

int check_moments(const cglbm_moments *m, const char *name) {
    if (!m) return fail(CGLBM_EINVAL, "null %s", name);
    float *const *p = reinterpret_cast<float *const *>(m);
    for (int f = 0; f < 11; ++f)
        if (!p[f]) return fail(CGLBM_EINVAL, "%s field %d is null", name, f);
    return CGLBM_OK;
}

namespace {
std::mutex g_tuned_mu;
std::map<std::array<int, 5>, MarchShape> g_tuned;
std::array<int, 5> tuned_key(const cglbm_config &c) { return {c.q, c.ybc, c.nx, c.ny, c.nz}; }
}  // namespace

bool tuned_shape_lookup(const cglbm_config &c, MarchShape *out) {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned.find(tuned_key(c));
    if (it == g_tuned.end()) return false;
    *out = it->second;
    return true;
}

void tuned_shape_store(const cglbm_config &c, MarchShape m) {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned[tuned_key(c)] = m;
}

int launch_step(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                const cglbm_layout *L, int64_t step, int kernel, const MarchShape *shape, hipStream_t s, int store,
                int64_t fs) {
    StepArgs a = make_step_args(A, B, cfg, L, step);
    if (store == kStoreF16) {
        // Only the z-marching kernel reads and writes the FP16-deviation format.
        MarchArgs ma;
        if (kernel == CGLBM_KERNEL_GATHER || (a.forced && a.ybc == CGLBM_Y_PERIODIC) || a.nx < 3 || a.ny < 3 ||
            a.nz < 3 || !march_args_f16(a, A->rhor, B->rhor, fs, ma))
            return fail(CGLBM_EINVAL, "FP16 storage needs the z-marching kernel and a supported case");
        MarchShape m;
        if (shape && shape->ty > 0) m = *shape;
        else if (!tuned_shape_lookup(*cfg, &m)) m = march_default_shape(a, cfg->q);
        const hipError_t e = launch_march_args(ma, cfg->q, m, s);
        if (e != hipSuccess) return hip_fail(e, "launch_march (FP16 storage)");
        return CGLBM_OK;
    }
    const bool march_ok = march_supports(a, cfg->q);
    if (kernel == CGLBM_KERNEL_MARCH && !march_ok)
        return fail(CGLBM_EINVAL, "the z-marching kernel does not support this shape/layout");
    hipError_t e;
    if (kernel != CGLBM_KERNEL_GATHER && march_ok) {
        MarchShape m;
        if (shape && shape->ty > 0) m = *shape;
        else if (!tuned_shape_lookup(*cfg, &m)) m = march_default_shape(a, cfg->q);
        e = launch_march_shape(a, cfg->q, m, s);
        if (e != hipSuccess) return hip_fail(e, "launch_march");
        return CGLBM_OK;  // boundary planes are written inside the marching step
    }
    e = launch_gather(a, cfg->q, s);
    if (e != hipSuccess) return hip_fail(e, "launch_gather");
    if (cfg->ybc != CGLBM_Y_PERIODIC) {
        BcArgs b = make_bc_args(B, A, cfg, L, step);
        e = launch_ybc(b, s);
        if (e != hipSuccess) return hip_fail(e, "launch_ybc");
    }
    return CGLBM_OK;
}

}  // namespace cglbm

// ------------------------------------------------------------------------------------------------
extern "C" {

const char *cglbm_last_error(void) { return g_err.c_str(); }

int cglbm_abi_version(void) { return CGLBM_ABI_VERSION; }

int cglbm_compact_layout(const cglbm_config *cfg, cglbm_layout *out) {
    if (int rc = validate_config(cfg)) return rc;
    if (!out) return fail(CGLBM_EINVAL, "null output");
    const int64_t nyt = cfg->ybc == CGLBM_Y_PERIODIC ? cfg->ny : cfg->ny + 2;
    out->origin = 0;
    out->pitch_y = cfg->nx;
    out->pitch_z = static_cast<int64_t>(cfg->nx) * nyt;
    return CGLBM_OK;
}

// launch_RCS (launch.cuh:89-94) + launch_JetBoundaryConditions (launch.cuh:78-83) as called by
// main.cu:41-43.
int cglbm_launch_rcs(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                     const cglbm_layout *layout, int64_t step, int32_t kernel, void *stream) {
    if (int rc = validate_config(cfg)) return rc;
    if (int rc = check_layout(cfg, layout)) return rc;
    if (int rc = check_moments(A, "A")) return rc;
    if (int rc = check_moments(B, "B")) return rc;
    if (A->rhor == B->rhor) return fail(CGLBM_EINVAL, "A and B alias");
    if (kernel < CGLBM_KERNEL_AUTO || kernel > CGLBM_KERNEL_MARCH) return fail(CGLBM_EINVAL, "unknown kernel %d", kernel);
    return launch_step(A, B, cfg, layout, step, kernel, nullptr, static_cast<hipStream_t>(stream));
}

// launch_JetBoundaryConditions (launch.cuh:78-83).
int cglbm_launch_ybc(const cglbm_moments *B, const cglbm_moments *Aold, const cglbm_config *cfg,
                     const cglbm_layout *layout, int64_t step, void *stream) {
    if (int rc = validate_config(cfg)) return rc;
    if (int rc = check_layout(cfg, layout)) return rc;
    if (int rc = check_moments(B, "B")) return rc;
    if (int rc = check_moments(Aold, "Aold")) return rc;
    BcArgs b = make_bc_args(B, Aold, cfg, layout, step);
    HIP_TRY(launch_ybc(b, static_cast<hipStream_t>(stream)));
    return CGLBM_OK;
}

}  // extern "C"
