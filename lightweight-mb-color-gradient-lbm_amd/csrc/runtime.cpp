// Host runtime behind include/cglbm.h: argument checking, device memory, streams, the step
// loop, initial conditions, copies, diagnostics and VTI output.
//
// Reference boundary replaced here (file:line in each entry point's comment): launch.cuh,
// memory.cuh, main.cu:38-55, io/vtiWriter.cuh, utilities/mlups_count.cuh.
#include "cglbm.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
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

// The update reads A while it writes B, so no field of B may share memory with any field of A, nor
// with another field of B: every pair is compared over the cells the layout addresses.
int check_disjoint(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *c, const cglbm_layout *L) {
    const int64_t nyt = c->ybc == CGLBM_Y_PERIODIC ? c->ny : c->ny + 2;
    const int64_t last = L->origin + (c->nx - 1) + L->pitch_y * (nyt - 1) + L->pitch_z * (c->nz - 1);
    const uintptr_t lo_off = static_cast<uintptr_t>(L->origin) * sizeof(float);
    const uintptr_t hi_off = static_cast<uintptr_t>(last + 1) * sizeof(float);
    float *const *pa = reinterpret_cast<float *const *>(A);
    float *const *pb = reinterpret_cast<float *const *>(B);
    for (int fa = 0; fa < 11; ++fa)
        for (int fb = 0; fb < 11; ++fb) {
            const uintptr_t a0 = reinterpret_cast<uintptr_t>(pa[fa]) + lo_off, a1 = reinterpret_cast<uintptr_t>(pa[fa]) + hi_off;
            const uintptr_t b0 = reinterpret_cast<uintptr_t>(pb[fb]) + lo_off, b1 = reinterpret_cast<uintptr_t>(pb[fb]) + hi_off;
            if (a0 < b1 && b0 < a1)
                return fail(CGLBM_EINVAL, "A and B alias or overlap (A field %d and B field %d share memory)", fa, fb);
        }
    // two fields of B at one place would take each other's stores
    for (int f = 0; f < 11; ++f)
        for (int g = f + 1; g < 11; ++g) {
            const uintptr_t f0 = reinterpret_cast<uintptr_t>(pb[f]) + lo_off, f1 = reinterpret_cast<uintptr_t>(pb[f]) + hi_off;
            const uintptr_t g0 = reinterpret_cast<uintptr_t>(pb[g]) + lo_off, g1 = reinterpret_cast<uintptr_t>(pb[g]) + hi_off;
            if (f0 < g1 && g0 < f1) return fail(CGLBM_EINVAL, "B fields %d and %d overlap", f, g);
        }
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
std::map<std::array<int, 7>, MarchShape> g_tuned;
// storage and forcing select different kernel instantiations, so they are part of the key
std::array<int, 7> tuned_key(const cglbm_config &c, int store) {
    return {c.q, c.ybc, c.nx, c.ny, c.nz, store, c.gravity != 0.0f ? 1 : 0};
}
}  // namespace

bool tuned_shape_lookup(const cglbm_config &c, int store, MarchShape *out) {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned.find(tuned_key(c, store));
    if (it == g_tuned.end()) return false;
    *out = it->second;
    return true;
}

void tuned_shape_store(const cglbm_config &c, int store, MarchShape m) {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned[tuned_key(c, store)] = m;
}

namespace {
constexpr int kTX = 32;
int ceil_div(int a, int b) { return (a + b - 1) / b; }
// Plan cost model per stencil, in bulk planes of a 16-row column tile: the extra cost of a plane
// whose SIMDs all run interface waves, and the fixed cost of a chunk (prologue and edge planes).
// A least-squares fit to per-work-group durations of MI355X steps (96^3 and 128^3 drops, chunks
// of 8 and 16 planes) gives alpha 0.7 / 1.0 and e 3.6 / 3.45 (D3Q19 / D3Q27); planned work-groups
// run longer than that model says (every CU busy to the end), and a sweep of whole-step times
// over alpha and e puts the best at the larger alpha kept here.
// CGLBM_PLAN_ALPHA / CGLBM_PLAN_E override both for tuning sweeps.
float env_float(const char *name, float dflt) {
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    char *end = nullptr;
    const float x = std::strtof(v, &end);
    return (end && *end == '\0' && x >= 0.0f && x <= 100.0f) ? x : dflt;
}
float plan_alpha(int q) { return env_float("CGLBM_PLAN_ALPHA", q == 27 ? 1.4f : 1.0f); }
float plan_overhead(int q) { return env_float("CGLBM_PLAN_E", q == 27 ? 3.45f : 3.6f); }

// Work items of a plan: one work-group per CU, or whole rounds of them when there are more
// column tiles than CUs. (More rounds with shorter chunks were measured slower: 192^3 and 256^3
// drops and a 128 x 512 x 128 RTI case lose 6-18% with two or four rounds, each chunk's fixed
// cost and its halo reads outweighing the finer balance.)
int plan_items_for(const cglbm_config &c, int ty, int ncu) {
    const int ncol = ceil_div(c.nx, kTX) * ceil_div(c.ny, ty);
    return ncu * ceil_div(ncol, ncu);
}

// Untuned shape of a library-owned domain: 16-row tiles with planned chunks (two exchange
// buffers where they fit the LDS) when the plan fits the planner, needs a single round of
// work-groups and its chunks average at least 2 planes; else the size-aware uniform default.
// Measured on drops from 64^3 to 256^3, planned chunks (with the measured correction) are within
// 2-6% of the best uniform shape at 64^3 and 1-31% faster from 96^3 up; plans of several rounds were
// slower than uniform chunks (384^3: 19.6 against 28.4 G updates/s), and at 320^3 (one round,
// with the planner given more LDS) planning gained nothing. A single-round domain with two chunks
// per column whose heavy columns are heavy along their whole length (the 128 x 512 x 128
// Rayleigh-Taylor layer) is planned 19-23% slower than uniform chunks: launch_default races the
// candidates and keeps the fastest, and the autotuner, which times both, picks uniform chunks there.
MarchShape untuned_shape(const cglbm_config &c, const MarchPlanState *plan, const MarchShape &dflt) {
    if (!plan || plan->ncu < 1) return dflt;
    // a shape forced through the environment (march_default_shape) wins over planning
    if (std::getenv("CGLBM_MARCH_TY") || std::getenv("CGLBM_MARCH_L") || std::getenv("CGLBM_MARCH_NBUF")) return dflt;
    const int ntx = ceil_div(c.nx, kTX), nty = ceil_div(c.ny, 16);
    const int nitems = plan_items_for(c, 16, plan->ncu);
    if (nitems > plan->cap || nitems > plan->ncu || !plan_fits(ntx, nty, c.nz, nitems)) return dflt;
    if (static_cast<int64_t>(ntx) * nty * c.nz < 2 * static_cast<int64_t>(nitems)) return dflt;
    return MarchShape{16, 0, c.q == 27 ? 1 : 2};
}

// The marching launch of one step, with the domain's chunk plan when the shape asks for one
// (L == 0) and the plan applies; otherwise uniform chunks of fallback_L planes.
hipError_t launch_march_planned(MarchArgs ma, const cglbm_config &c, MarchShape m, MarchPlanState *plan,
                                int fallback_L, hipStream_t s) {
    const bool want = m.L == 0;
    MarchShape mu = m;
    if (want) mu.L = fallback_L;  // chunk length of unplanned launches
    if (!want || !plan || !plan->items || !plan->wmap) return launch_march_args(ma, c.q, mu, s);
    const int ntx = ceil_div(c.nx, kTX), nty = ceil_div(c.ny, m.ty);
    const int nitems = plan_items_for(c, m.ty, plan->ncu);
    if (nitems > plan->cap || !plan_fits(ntx, nty, c.nz, nitems)) return launch_march_args(ma, c.q, mu, s);
    const bool have = plan->ty == m.ty;
    const bool record = !have || plan->since + 1 >= plan->period;
    if (have) {
        ma.items = plan->items;
        ma.nitems = nitems;
    }
    if (record) ma.wmap = plan->wmap;
    const bool timed = record && have && plan->itime;  // the items of the current plan get timed
    if (timed) ma.itime = plan->itime;
    hipError_t e = launch_march_args(ma, c.q, mu, s);
    if (e != hipSuccess) return e;
    if (record) {
        e = launch_plan(plan->wmap, timed ? plan->itime : nullptr, ntx, nty, c.nz, m.ty / 2, plan_alpha(c.q),
                        plan_overhead(c.q), plan->items, nitems, s);
        if (e != hipSuccess) {
            plan->ty = 0;
            return e;
        }
        // the first plan of a state is re-made at the next step, from its own measured durations
        plan->since = have ? 0 : plan->period - 1;
        plan->ty = m.ty;
    } else {
        ++plan->since;
    }
    return hipSuccess;
}
// The step of a domain left to the default shape (no tile set, no autotuned record). A short race
// (MarchPlanState::trial) picks among planned chunks where untuned_shape plans them, the size-aware
// uniform default, for D3Q19 4-row tiles with 32-plane chunks, and 16-row tiles with one and three
// rounds of uniform chunks: each runs kTrialSteps steps timed with events on the domain's stream,
// the host waits for the last of them (a few milliseconds, once per race), the fastest is kept,
// and the race runs again after 128, 512 and then every 2048 steps (MarchPlanState::trial_period). Planned chunks lose where the heavy column tiles are heavy along their whole
// length and one round of work-groups cannot even them out (a flat Rayleigh-Taylor layer across x
// and z); states too large for the 256 MB last-level cache run faster with the smaller 4-row
// work-groups, three of which share a CU (the 128 x 512 x 128 Rayleigh-Taylor case in FP32). The
// planned candidate's share of its re-plans (a recording step and the planner every `period`
// steps, about 2% of a step) is charged to it. Results are the same bits for every candidate.
hipError_t launch_default(const MarchArgs &ma, const cglbm_config &c, MarchPlanState *plan, const MarchShape &dflt,
                          hipStream_t s) {
    const MarchShape first = untuned_shape(c, plan, dflt);
    if (!plan || !plan->tev[0]) return launch_march_planned(ma, c, first, plan, dflt.L, s);
    MarchShape cand[kTrialMaxCand];
    int n = 0;
    cand[n++] = first;
    if (first.L == 0) cand[n++] = dflt;
    const MarchShape narrow{4, c.nz < 32 ? c.nz : 32, 2};
    if (c.q == 19 && !(dflt.ty == narrow.ty && dflt.L == narrow.L)) cand[n++] = narrow;
    // 16-row tiles with the longest uniform chunks that still give one round of work-groups (no
    // re-plans; the autotuner's pick on the 64^3 bubble) and three rounds (fewer chunk edges than
    // the default's shorter chunks, rounds enough for the dispatcher to even out heavy columns;
    // the autotuner's pick on the 128 x 512 x 128 Rayleigh-Taylor layer), at least 8 planes
    auto taken = [&](const MarchShape &m) {
        for (int k = 0; k < n; ++k)
            if (cand[k].ty == m.ty && cand[k].L == m.L) return true;
        return dflt.ty == m.ty && dflt.L == m.L;
    };
    const int64_t cols = static_cast<int64_t>(ceil_div(c.nx, kTX)) * ceil_div(c.ny, 16);
    for (const int rounds : {1, 3}) {
        const int64_t nch = (rounds * static_cast<int64_t>(plan->ncu) + cols - 1) / cols;
        const int L = ceil_div(c.nz, static_cast<int>(nch < c.nz ? nch : c.nz));
        const MarchShape m{16, L, c.q == 27 ? 1 : 2};
        if ((rounds == 1 || L >= 8) && n < kTrialMaxCand && !taken(m)) cand[n++] = m;
    }
    auto run = [&](const MarchShape &m) { return launch_march_planned(ma, c, m, plan, dflt.L, s); };
    hipError_t e = hipSuccess;
    switch (plan->trial) {
    case kTrialKeep:
        if (++plan->trial_since < plan->trial_period) return run(plan->keep);
        plan->trial_period = plan->trial_period * 4 < kTrialPeriod ? plan->trial_period * 4 : kTrialPeriod;
        plan->trial = kTrialNone;  // race again: the interface has moved
        if (plan->keep.L != 0) plan->ty = 0;  // planned chunks race on a fresh plan
        return run(plan->keep);
    case kTrialNone:
        // start once a planned first candidate has its corrected plan, with no re-plan due within
        // its timed steps
        if (first.L == 0 && !(plan->ty == first.ty && plan->since + kTrialSteps < plan->period)) return run(first);
        plan->trial = kTrialTiming;
        plan->trial_cand = 0;
        plan->trial_left = kTrialSteps;
        e = hipEventRecord(plan->tev[0], s);
        if (e != hipSuccess) return e;
        break;
    default:
        break;
    }
    // kTrialTiming: one step of the candidate being timed
    const int k = plan->trial_cand < n ? plan->trial_cand : n - 1;
    e = run(cand[k]);
    if (e != hipSuccess || --plan->trial_left > 0) return e;
    e = hipEventRecord(plan->tev[2 * k + 1], s);
    if (e != hipSuccess) return e;
    if (k + 1 < n) {
        plan->trial_cand = k + 1;
        plan->trial_left = kTrialSteps;
        return hipEventRecord(plan->tev[2 * k + 2], s);
    }
    // Every candidate timed: wait for the last event (the host would otherwise queue many steps of
    // one candidate before the result is known) and keep the fastest.
    e = hipEventSynchronize(plan->tev[2 * k + 1]);
    if (e != hipSuccess) return e;
    int best = 0;
    float best_ms = 0.0f;
    for (int j = 0; j < n; ++j) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, plan->tev[2 * j], plan->tev[2 * j + 1]);
        if (e != hipSuccess) return e;
        if (cand[j].L == 0) ms *= 1.02f;
        if (j == 0 || ms < best_ms) {
            best = j;
            best_ms = ms;
        }
    }
    plan->keep = cand[best];
    plan->trial = kTrialKeep;
    plan->trial_since = 0;
    return hipSuccess;
}
}  // namespace

int plan_items(const cglbm_config &c, int ty, int ncu) { return plan_items_for(c, ty, ncu); }

int plan_capacity(const cglbm_config &c, int ncu) {
    int best = 0;
    for (int ty : {4, 8, 16}) {
        const int n = plan_items_for(c, ty, ncu);
        best = n > best ? n : best;
    }
    return best;
}

int64_t plan_device_bytes(const cglbm_config &c) {
    // the CU count is not known without a device: bound it by 1024
    return static_cast<int64_t>(sizeof(int4) + sizeof(uint32_t)) * plan_capacity(c, 1024) +
           static_cast<int64_t>(plan_wmap_bytes(c));
}

size_t plan_wmap_bytes(const cglbm_config &c) {
    size_t best = 0;
    for (int ty : {4, 8, 16}) {
        const size_t b = static_cast<size_t>(ceil_div(c.nx, kTX)) * ceil_div(c.ny, ty) * c.nz * (ty / 2);
        best = b > best ? b : best;
    }
    return best;
}

int launch_step(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                const cglbm_layout *L, int64_t step, int kernel, const MarchShape *shape, hipStream_t s, int store,
                int64_t fs, MarchPlanState *plan) {
    StepArgs a = make_step_args(A, B, cfg, L, step);
    if (store == kStoreF16) {
        // Only the z-marching kernel reads and writes the FP16-deviation format.
        MarchArgs ma;
        if (kernel == CGLBM_KERNEL_GATHER || (a.forced && a.ybc == CGLBM_Y_PERIODIC) || a.nx < 3 || a.ny < 3 ||
            a.nz < 3 || !march_args_f16(a, A->rhor, B->rhor, fs, ma))
            return fail(CGLBM_EINVAL, "FP16 storage needs the z-marching kernel and a supported case");
        MarchShape m;
        const MarchShape dflt = march_default_shape(a, cfg->q);
        hipError_t e;
        if (shape && shape->ty > 0) e = launch_march_planned(ma, *cfg, *shape, plan, dflt.L, s);
        else if (tuned_shape_lookup(*cfg, store, &m)) e = launch_march_planned(ma, *cfg, m, plan, dflt.L, s);
        else e = launch_default(ma, *cfg, plan, dflt, s);
        if (e != hipSuccess) return hip_fail(e, "launch_march (FP16 storage)");
        return CGLBM_OK;
    }
    const bool march_ok = march_supports(a, cfg->q);
    if (kernel == CGLBM_KERNEL_MARCH && !march_ok)
        return fail(CGLBM_EINVAL, "the z-marching kernel does not support this shape/layout");
    hipError_t e;
    if (kernel != CGLBM_KERNEL_GATHER && march_ok) {
        MarchShape m{0, 0, 0};
        const MarchShape dflt = march_default_shape(a, cfg->q);
        if (shape && shape->ty > 0) m = *shape;
        else if (!tuned_shape_lookup(*cfg, store, &m)) m = MarchShape{0, 0, 0};
        if (m.ty != 0 && ((m.ty != 4 && m.ty != 8 && m.ty != 16) || m.L < 0 || (m.nbuf != 1 && m.nbuf != 2)))
            return fail(CGLBM_EINVAL, "bad marching shape");
        MarchArgs ma;
        if (!march_args(a, ma)) return fail(CGLBM_EINVAL, "the z-marching kernel does not support this shape/layout");
        e = m.ty != 0 ? launch_march_planned(ma, *cfg, m, plan, dflt.L, s) : launch_default(ma, *cfg, plan, dflt, s);
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

// A caller-held chunk plan (cglbm_plan_create): the plan state of one pair of caller-owned states.
struct cglbm_plan {
    cglbm_config cfg{};
    int device = 0;
    cglbm::MarchPlanState st;
};

// ---- Scratch of the no-allocation planned launch ---------------------------------------------------
// cglbm_launch_rcs_planned_noalloc keeps its chunk plan (work items, their measured durations and
// the interface map) in this file-scope device array instead of a cglbm_plan: each GPU the library
// runs on holds its own copy until the process ends. Host code never passes the array's name to a
// kernel or a copy: it takes the device address once per device with hipGetSymbolAddress and uses
// that. Nothing in it needs to start at zero: after a change of domain shape the first call records
// a complete interface map (every column tile and plane) before any plan is read.
constexpr size_t kNoAllocScratchBytes = 384 * 1024;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct NoAllocSlot {
    bool ready = false;
    unsigned char *ptr = nullptr;  // this device's copy of `scratch`
    cglbm_config key{};            // extents and stencil the current plan was made for
    MarchPlanState st;             // host side of the plan, its arrays inside `ptr`
};
std::mutex g_noalloc_mu;
NoAllocSlot g_noalloc[kMaxDevices];

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// The slot of the current device, set up on its first use (symbol address, CU count).
int noalloc_slot(NoAllocSlot **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(CGLBM_EINVAL, "device ordinal %d beyond %d", dev, kMaxDevices);
    NoAllocSlot &sl = g_noalloc[dev];
    if (!sl.ready) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        int ncu = 0;
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
        sl.ptr = static_cast<unsigned char *>(p);
        sl.st.ncu = ncu;
        sl.ready = true;
    }
    *out = &sl;
    return CGLBM_OK;
}
}  // namespace

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
    if (int rc = check_disjoint(A, B, cfg, layout)) return rc;
    if (kernel < CGLBM_KERNEL_AUTO || kernel > CGLBM_KERNEL_MARCH) return fail(CGLBM_EINVAL, "unknown kernel %d", kernel);
    return launch_step(A, B, cfg, layout, step, kernel, nullptr, static_cast<hipStream_t>(stream));
}

int cglbm_launch_rcs_zslab(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                           const cglbm_layout *layout, int32_t z_lo, int32_t z_hi, int32_t z_offset, int64_t step,
                           void *stream) {
    if (int rc = validate_config(cfg)) return rc;
    if (int rc = check_layout(cfg, layout)) return rc;
    if (int rc = check_moments(A, "A")) return rc;
    if (int rc = check_moments(B, "B")) return rc;
    if (int rc = check_disjoint(A, B, cfg, layout)) return rc;
    if (!(z_lo >= 2 && z_lo < z_hi && z_hi <= cfg->nz - 2))
        return fail(CGLBM_EINVAL, "z-slab range [%d, %d) must satisfy 2 <= z_lo < z_hi <= nz - 2 (two halo planes "
                                  "on each side; nz = %d)", z_lo, z_hi, cfg->nz);
    const StepArgs a = make_step_args(A, B, cfg, layout, step);
    MarchArgs ma;
    if (!march_supports(a, cfg->q) || !march_args(a, ma))
        return fail(CGLBM_EINVAL, "a z-slab update needs the z-marching kernel (a periodic domain with a body force "
                                  "takes the gather kernel)");
    ma.zlo = z_lo;
    ma.zhi = z_hi;
    ma.zoff = z_offset;
    // the size-aware uniform default of a domain of the slab's own planes
    StepArgs own = a;
    own.nz = z_hi - z_lo;
    const MarchShape m = march_default_shape(own, cfg->q);
    const hipError_t e = launch_march_args(ma, cfg->q, effective_shape(cfg->q, m), static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "launch_march (z-slab)");
    return CGLBM_OK;
}

int cglbm_launch_rcs_zslab_faces(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                                 const cglbm_layout *layout, int32_t z_lo, int32_t z_hi, int32_t width,
                                 int32_t z_offset, int64_t step, void *stream) {
    if (int rc = validate_config(cfg)) return rc;
    if (int rc = check_layout(cfg, layout)) return rc;
    if (int rc = check_moments(A, "A")) return rc;
    if (int rc = check_moments(B, "B")) return rc;
    if (int rc = check_disjoint(A, B, cfg, layout)) return rc;
    if (!(z_lo >= 2 && z_lo < z_hi && z_hi <= cfg->nz - 2))
        return fail(CGLBM_EINVAL, "z-slab range [%d, %d) must satisfy 2 <= z_lo < z_hi <= nz - 2 (two halo planes "
                                  "on each side; nz = %d)", z_lo, z_hi, cfg->nz);
    if (width < 1 || 2 * static_cast<int64_t>(width) > z_hi - z_lo)
        return fail(CGLBM_EINVAL, "face width %d must be at least 1 and at most half the slab (%d planes)", width,
                    z_hi - z_lo);
    const StepArgs a = make_step_args(A, B, cfg, layout, step);
    MarchArgs ma;
    if (!march_supports(a, cfg->q) || !march_args(a, ma))
        return fail(CGLBM_EINVAL, "a z-slab update needs the z-marching kernel (a periodic domain with a body force "
                                  "takes the gather kernel)");
    ma.zlo = z_lo;
    ma.zhi = z_hi;
    ma.zoff = z_offset;
    ma.zface = width;
    // the size-aware uniform default of a domain of both faces' planes
    StepArgs own = a;
    own.nz = 2 * width;
    const MarchShape m = march_default_shape(own, cfg->q);
    const hipError_t e = launch_march_args(ma, cfg->q, effective_shape(cfg->q, m), static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "launch_march (z-slab faces)");
    return CGLBM_OK;
}

static int zslab_halo(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, int32_t z_lo,
                      int32_t z_hi, float *lo, float *hi, bool pack, void *stream) {
    if (int rc = validate_config(cfg)) return rc;
    if (int rc = check_layout(cfg, layout)) return rc;
    if (int rc = check_moments(A, "A")) return rc;
    if (!lo || !hi) return fail(CGLBM_EINVAL, "null halo buffer");
    if (!(z_lo >= 2 && z_lo < z_hi && z_hi <= cfg->nz - 2))
        return fail(CGLBM_EINVAL, "z-slab range [%d, %d) must satisfy 2 <= z_lo < z_hi <= nz - 2 (two halo planes "
                                  "on each side; nz = %d)", z_lo, z_hi, cfg->nz);
    if (pack && z_hi - z_lo < 2) return fail(CGLBM_EINVAL, "packing needs a slab of at least 2 planes");
    const int64_t nyt = cfg->ybc == CGLBM_Y_PERIODIC ? cfg->ny : cfg->ny + 2;
    if (26 * static_cast<int64_t>(cfg->nx) * nyt >= (int64_t(1) << 31)) return fail(CGLBM_EINVAL, "halo planes too large");
    const StepArgs a = make_step_args(A, A, cfg, layout, 0);
    const hipError_t e = launch_zslab_halo(a, z_lo, z_hi, lo, hi, pack, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, pack ? "zslab pack" : "zslab unpack");
    return CGLBM_OK;
}

int cglbm_zslab_pack(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, int32_t z_lo,
                     int32_t z_hi, float *send_lo, float *send_hi, void *stream) {
    return zslab_halo(A, cfg, layout, z_lo, z_hi, send_lo, send_hi, true, stream);
}

int cglbm_zslab_unpack(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, int32_t z_lo,
                       int32_t z_hi, const float *recv_lo, const float *recv_hi, void *stream) {
    return zslab_halo(A, cfg, layout, z_lo, z_hi, const_cast<float *>(recv_lo), const_cast<float *>(recv_hi), false,
                      stream);
}

int cglbm_plan_create(const cglbm_config *cfg, cglbm_plan **out) {
    if (int rc = validate_config(cfg)) return rc;
    if (!out) return fail(CGLBM_EINVAL, "null output");
    *out = nullptr;
    cglbm_plan *p = new (std::nothrow) cglbm_plan;
    if (!p) return fail(CGLBM_ENOMEM, "host allocation failed");
    p->cfg = *cfg;
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&p->st.ncu, hipDeviceAttributeMultiprocessorCount, p->device);
    if (e == hipSuccess) {
        p->st.cap = plan_capacity(*cfg, p->st.ncu);
        e = hipMalloc(&p->st.items, sizeof(int4) * p->st.cap);
    }
    if (e == hipSuccess) e = hipMalloc(&p->st.wmap, plan_wmap_bytes(*cfg));
    if (e == hipSuccess) e = hipMalloc(&p->st.itime, sizeof(uint32_t) * p->st.cap);
    if (e != hipSuccess) {
        (void)cglbm_plan_destroy(p);
        return hip_fail(e, "plan setup");
    }
    *out = p;
    return CGLBM_OK;
}

int cglbm_plan_destroy(cglbm_plan *plan) {
    if (!plan) return CGLBM_OK;
    if (plan->st.items) (void)hipFree(plan->st.items);
    if (plan->st.wmap) (void)hipFree(plan->st.wmap);
    if (plan->st.itime) (void)hipFree(plan->st.itime);
    delete plan;
    return CGLBM_OK;
}

int cglbm_plan_reset(cglbm_plan *plan) {
    if (!plan) return fail(CGLBM_EINVAL, "null plan");
    plan->st.ty = 0;
    return CGLBM_OK;
}

int cglbm_launch_rcs_planned(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                             const cglbm_layout *layout, int64_t step, cglbm_plan *plan, void *stream) {
    if (!plan) return fail(CGLBM_EINVAL, "null plan");
    if (int rc = validate_config(cfg)) return rc;
    if (cfg->nx != plan->cfg.nx || cfg->ny != plan->cfg.ny || cfg->nz != plan->cfg.nz || cfg->q != plan->cfg.q)
        return fail(CGLBM_EINVAL, "the plan was made for another domain shape or stencil");
    if (int rc = check_layout(cfg, layout)) return rc;
    if (int rc = check_moments(A, "A")) return rc;
    if (int rc = check_moments(B, "B")) return rc;
    if (int rc = check_disjoint(A, B, cfg, layout)) return rc;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // plans of several rounds of work-groups are slower than uniform chunks: plan one round only
    if (plan_items(*cfg, 16, plan->st.ncu) > plan->st.ncu)
        return launch_step(A, B, cfg, layout, step, CGLBM_KERNEL_AUTO, nullptr, s);
    const MarchShape m{16, 0, cfg->q == 27 ? 1 : 2};
    return launch_step(A, B, cfg, layout, step, CGLBM_KERNEL_AUTO, &m, s, 0, 0, &plan->st);
}

// cglbm_launch_rcs_planned with the plan in the per-device `scratch` array: no allocation, no plan
// object. Calls that overlap on one GPU share that array and race.
int cglbm_launch_rcs_planned_noalloc(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                                     const cglbm_layout *layout, int64_t step, void *stream) {
    if (int rc = validate_config(cfg)) return rc;
    if (int rc = check_layout(cfg, layout)) return rc;
    if (int rc = check_moments(A, "A")) return rc;
    if (int rc = check_moments(B, "B")) return rc;
    if (int rc = check_disjoint(A, B, cfg, layout)) return rc;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    NoAllocSlot *sl = nullptr;
    if (int rc = noalloc_slot(&sl)) return rc;
    const int ncu = sl->st.ncu;
    // as cglbm_launch_rcs_planned: plans of one round of work-groups only
    if (plan_items(*cfg, 16, ncu) > ncu) return launch_step(A, B, cfg, layout, step, CGLBM_KERNEL_AUTO, nullptr, s);
    const size_t items_b = align256(sizeof(int4) * static_cast<size_t>(ncu));
    const size_t itime_b = align256(sizeof(uint32_t) * static_cast<size_t>(ncu));
    const size_t wmap_b = align256(static_cast<size_t>(ceil_div(cfg->nx, kTX)) * ceil_div(cfg->ny, 16) *
                                   static_cast<size_t>(cfg->nz) * 8);
    if (items_b + itime_b + wmap_b > kNoAllocScratchBytes)
        return fail(CGLBM_EINVAL, "the no-allocation planned launch needs %zu bytes of scratch for a %d x %d x %d domain; "
                    "it has %zu (use cglbm_launch_rcs_planned with a cglbm_plan)", items_b + itime_b + wmap_b, cfg->nx,
                    cfg->ny, cfg->nz, kNoAllocScratchBytes);
    const cglbm_config &k = sl->key;
    if (k.nx != cfg->nx || k.ny != cfg->ny || k.nz != cfg->nz || k.q != cfg->q || k.ybc != cfg->ybc) {
        sl->key = *cfg;
        sl->st.ty = 0;  // another domain shape: record a fresh interface map first
    }
    sl->st.items = reinterpret_cast<int4 *>(sl->ptr);
    sl->st.itime = reinterpret_cast<uint32_t *>(sl->ptr + items_b);
    sl->st.wmap = sl->ptr + items_b + itime_b;
    sl->st.cap = ncu;
    const MarchShape m{16, 0, cfg->q == 27 ? 1 : 2};
    return launch_step(A, B, cfg, layout, step, CGLBM_KERNEL_AUTO, &m, s, 0, 0, &sl->st);
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
