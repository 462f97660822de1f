// Library-owned domains: the LbmDevice/LbmHost ownership of memory.cuh:14-219, the step loop
// of main.cu:38-55, the case initialisers of bubble/bubbleIn.cuh, jet/jetIn.cuh and
// rti/rtiIn.cuh, the D2H copy of memory.cuh:209-217, the VTI writer of io/vtiWriter.cuh and
// the event timing of utilities/mlups_count.cuh.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <new>
#include <string>
#include <vector>

#include "cglbm.h"
#include "kernels.hpp"
#include "lattice.hpp"
#include "state.hpp"
#include "storage.hpp"
#include "vti.hpp"
#include "domain_internal.hpp"

using namespace cglbm;

namespace {

int herr(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(CGLBM_EHIP, m.c_str());
}

#define HTRY(expr)                                  \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return herr(e_, #expr); \
    } while (0)

// FP32: the 11 field pointers of each state. FP16: only rhor/rhob are FP32 arrays; the other
// pointers stay null (the step takes the state base, the copies go through an FP32 staging area).
void make_views(cglbm_domain *d) {
    for (int b = 0; b < 2; ++b) {
        float **v = reinterpret_cast<float **>(&d->view[b]);
        for (int f = 0; f < 11; ++f)
            v[f] = (d->storage == CGLBM_STORE_F32 || f < 2) ? d->buf[b] + f * d->fstride : nullptr;
    }
}

// Transient FP32 area for an FP16 domain: nf fields of fstride floats (11 for the initialisers,
// which assemble a whole FP32 state; 1 for the host copies, which convert one field at a time),
// with field views when nf == 11; freed by the destructor.
struct Staging {
    float *p = nullptr;
    cglbm_moments view{};
    ~Staging() {
        if (p) (void)hipFree(p);
    }
};

int make_staging(const cglbm_domain *d, Staging &st, int nf) {
    HTRY(hipMalloc(&st.p, sizeof(float) * nf * static_cast<size_t>(d->fstride)));
    if (nf == 11) {
        float **v = reinterpret_cast<float **>(&st.view);
        for (int f = 0; f < 11; ++f) v[f] = st.p + f * d->fstride;
    }
    return CGLBM_OK;
}

void release(cglbm_domain *d) {
    if (!d) return;
    for (auto &b : d->buf)
        if (b) (void)hipFree(b);
    if (d->d_partials) (void)hipFree(d->d_partials);
    if (d->d_result) (void)hipFree(d->d_result);
    if (d->plan.items) (void)hipFree(d->plan.items);
    if (d->plan.wmap) (void)hipFree(d->plan.wmap);
    if (d->plan.itime) (void)hipFree(d->plan.itime);
    for (auto &ev : d->plan.tev)
        if (ev) (void)hipEventDestroy(ev);
    if (d->ev0) (void)hipEventDestroy(d->ev0);
    if (d->ev1) (void)hipEventDestroy(d->ev1);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

int apply_ybc_views(cglbm_domain *d, const cglbm_moments &m) {
    BcArgs b = make_bc_args(&m, &m, &d->cfg, &d->layout, 0);
    HTRY(launch_ybc(b, d->stream));
    return CGLBM_OK;
}

// Initial state from the host (11 fields, compact layout), then, if `bc`, the boundary planes
// computed from it in place (main.cu:26 calls the boundary kernel with B == Aold). FP16 domains
// assemble the FP32 state in staging and encode it once at the end.
int upload_init(cglbm_domain *d, const std::vector<float> &host, bool bc) {
    const int64_t n = stored_cells(d->cfg);
    d->plan.ty = 0;  // a new state: the next planned step records its interface map afresh
    d->plan.trial = kTrialNone;  // and the default shapes race again, soon again after that
    d->plan.trial_period = kTrialFirstPeriod;
    Staging st;
    const cglbm_moments *m = &d->view[d->cur];
    if (d->storage == CGLBM_STORE_F16) {
        if (int rc = make_staging(d, st, 11)) return rc;
        m = &st.view;
    }
    float *const *v = reinterpret_cast<float *const *>(m);
    for (int f = 0; f < 11; ++f)
        HTRY(hipMemcpyAsync(v[f], host.data() + f * n, sizeof(float) * n, hipMemcpyHostToDevice, d->stream));
    if (bc && d->cfg.ybc != CGLBM_Y_PERIODIC)
        if (int rc = apply_ybc_views(d, *m)) return rc;
    if (d->storage == CGLBM_STORE_F16)
        HTRY(launch_pack_f16(st.p, d->fstride, d->buf[d->cur], d->fstride, n, 0, 11, d->stream));
    HTRY(hipStreamSynchronize(d->stream));
    return CGLBM_OK;
}

int check_dom(const cglbm_domain *d) {
    if (!d) return set_error(CGLBM_EINVAL, "null domain");
    return CGLBM_OK;
}

// A batch member is stepped on the batch's stream; synchronous member operations wait for it.
int wait_batch(cglbm_domain *d) {
    if (d->batch_stream) HTRY(hipStreamSynchronize(d->batch_stream));
    return CGLBM_OK;
}


// Host-built initial states (11 fields, compact layout, computed in double and rounded once).
int host_bubble(const cglbm_config &c, double x0, double y0, double z0, double radius, double width, int32_t sharp,
                std::vector<float> &h) {
    if (!(radius > 0) || (!sharp && !(width > 0))) return set_error(CGLBM_EINVAL, "radius/width must be positive");
    const int nyt = stored_ny(c);
    const int yoff = c.ybc == CGLBM_Y_PERIODIC ? 0 : 1;
    const int64_t n = stored_cells(c);
    h.assign(11 * n, 0.0f);
    for (int z = 0; z < c.nz; ++z)
        for (int yy = 0; yy < nyt; ++yy)
            for (int x = 0; x < c.nx; ++x) {
                const int64_t id = x + static_cast<int64_t>(c.nx) * (yy + static_cast<int64_t>(nyt) * z);
                const double dx = x - x0, dy = (yy - yoff) - y0, dz = z - z0;
                if (sharp) {
                    const float fx = static_cast<float>(dx), fy = static_cast<float>(dy), fz = static_cast<float>(dz);
                    const float r2 = fx * fx + fy * fy + fz * fz;
                    const float R = static_cast<float>(radius);
                    const float inside = static_cast<float>(r2 <= R * R);
                    h[id] = (1.0f - inside) * 1.0f;
                    h[n + id] = inside * 1.0f;
                } else {
                    const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                    const double phi = std::tanh(2.0 * (radius - r) / width);
                    h[id] = static_cast<float>(0.5 * (1.0 - phi));
                    h[n + id] = static_cast<float>(0.5 * (1.0 + phi));
                }
            }
    return CGLBM_OK;
}

// Red fluid at rest everywhere (jetIn.cuh:16-32); the inlet disk comes from the boundary planes.
int host_jet(const cglbm_config &c, std::vector<float> &h) {
    if (c.ybc != CGLBM_Y_JET) return set_error(CGLBM_EINVAL, "jet initialisation needs ybc = CGLBM_Y_JET");
    const int64_t n = stored_cells(c);
    h.assign(11 * n, 0.0f);
    for (int64_t i = 0; i < n; ++i) h[i] = 1.0f;
    return CGLBM_OK;
}

int host_rti(const cglbm_config &c, double amplitude, double width, int32_t n_random_modes, double random_amplitude,
             uint32_t seed, std::vector<float> &h) {
    if (!(width > 0) || n_random_modes < 0 || n_random_modes > 64) return set_error(CGLBM_EINVAL, "bad RTI parameters");
    const int nyt = stored_ny(c);
    const bool per = c.ybc == CGLBM_Y_PERIODIC;
    const double PI = 3.14159265358979323846;
    const double Lx = c.nx, Lz = c.nz;
    const double y0 = (nyt - 1) / 2.0 + (per ? 1.0 : 0.0);
    std::vector<double> ma(n_random_modes), mb(n_random_modes), ph(n_random_modes);
    uint64_t s = seed;
    for (int m = 0; m < n_random_modes; ++m) {
        s += 0x9E3779B97F4A7C15ull;
        ma[m] = 1 + (mix64(s) % 8);
        s += 0x9E3779B97F4A7C15ull;
        mb[m] = 1 + (mix64(s) % 8);
        s += 0x9E3779B97F4A7C15ull;
        ph[m] = 2.0 * PI * (static_cast<double>(mix64(s) >> 11) * (1.0 / 9007199254740992.0));
    }
    const int64_t n = stored_cells(c);
    h.assign(11 * n, 0.0f);
    for (int z = 0; z < c.nz; ++z)
        for (int x = 0; x < c.nx; ++x) {
            const double xr = x + 1.0, zr = z + 1.0;
            double yi = y0 + amplitude * std::cos(2.0 * PI * xr / Lx) * std::cos(2.0 * PI * zr / Lz);
            for (int m = 0; m < n_random_modes; ++m)
                yi += random_amplitude * std::cos(2.0 * PI * (ma[m] * xr / Lx + mb[m] * zr / Lz) + ph[m]);
            for (int yy = 0; yy < nyt; ++yy) {
                const double yr = yy + (per ? 1.0 : 0.0);
                const double phi = std::tanh((yr - yi) / width);
                const int64_t id = x + static_cast<int64_t>(c.nx) * (yy + static_cast<int64_t>(nyt) * z);
                h[id] = static_cast<float>(0.5 * (1.0 - phi));
                h[n + id] = static_cast<float>(0.5 * (1.0 + phi));
            }
        }
    return CGLBM_OK;
}

int check_launch_init(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout) {
    if (int rc = validate_config(cfg)) return rc;
    if (int rc = check_layout(cfg, layout)) return rc;
    if (!A) return set_error(CGLBM_EINVAL, "null moments");
    float *const *p = reinterpret_cast<float *const *>(A);
    for (int f = 0; f < 11; ++f)
        if (!p[f]) return set_error(CGLBM_EINVAL, "null field pointer");
    return CGLBM_OK;
}

// A host-built compact state into caller arrays at `layout` (staging upload + scatter), then, if
// `bc` and y is bounded, the boundary planes computed from it in place (main.cu:26).
int upload_to_layout(const cglbm_moments *A, const cglbm_config &c, const cglbm_layout &L, const std::vector<float> &host,
                     bool bc, hipStream_t s) {
    Staging st;
    const int64_t n = stored_cells(c);
    HTRY(hipMalloc(&st.p, sizeof(float) * 11 * static_cast<size_t>(n)));
    HTRY(hipMemcpyAsync(st.p, host.data(), sizeof(float) * 11 * static_cast<size_t>(n), hipMemcpyHostToDevice, s));
    FieldPtrs dst;
    float *const *p = reinterpret_cast<float *const *>(A);
    for (int f = 0; f < 11; ++f) dst.p[f] = p[f] + L.origin;
    HTRY(launch_scatter_fields(st.p, dst, c.nx, stored_ny(c), c.nz, static_cast<uint32_t>(L.pitch_y),
                               static_cast<uint32_t>(L.pitch_z), s));
    if (bc && c.ybc != CGLBM_Y_PERIODIC) {
        BcArgs b = make_bc_args(A, A, &c, &L, 0);
        HTRY(launch_ybc(b, s));
    }
    HTRY(hipStreamSynchronize(s));
    return CGLBM_OK;
}

}  // namespace

extern "C" {

int64_t cglbm_domain_bytes(const cglbm_config *cfg, int32_t storage) {
    if (validate_config(cfg)) return -1;
    if (storage != CGLBM_STORE_F32 && storage != CGLBM_STORE_F16) return -1;
    return 2 * state_bytes(storage, field_stride(*cfg)) + static_cast<int64_t>(sizeof(double)) * (kReduceBlocks * 6 + 8) +
           plan_device_bytes(*cfg);
}

// allocate_device_memory (memory.cuh:153-164): two zeroed 11-field states.
int cglbm_domain_create(const cglbm_config *cfg, int32_t storage, cglbm_domain **out) {
    if (!out) return set_error(CGLBM_EINVAL, "null output");
    *out = nullptr;
    if (int rc = validate_config(cfg)) return rc;
    if (storage != CGLBM_STORE_F32 && storage != CGLBM_STORE_F16) return set_error(CGLBM_EINVAL, "unknown storage mode");
    if (storage == CGLBM_STORE_F16) {
        if (cfg->gravity != 0.0f && cfg->ybc == CGLBM_Y_PERIODIC)
            return set_error(CGLBM_EINVAL, "FP16 storage: a body force needs bounded y (walls or jet planes)");
        if (cfg->nx < 3 || cfg->ny < 3 || cfg->nz < 3)
            return set_error(CGLBM_EINVAL, "FP16 storage needs at least 3 cells along each axis");
        if (state_bytes(kStoreF16, field_stride(*cfg)) >= (int64_t(1) << 32))
            return set_error(CGLBM_EINVAL, "FP16 storage: a state must stay below 4 GiB");
    }
    cglbm_domain *d = new (std::nothrow) cglbm_domain;
    if (!d) return set_error(CGLBM_ENOMEM, "host allocation failed");
    d->cfg = *cfg;
    d->storage = storage;
    d->fstride = field_stride(*cfg);
    if (hipGetDevice(&d->device) != hipSuccess) {
        release(d);
        return set_error(CGLBM_EHIP, "no HIP device");
    }
    const size_t bytes = static_cast<size_t>(state_bytes(storage, d->fstride));
    for (int b = 0; b < 2; ++b) {
        if (hipMalloc(&d->buf[b], bytes) != hipSuccess) {
            release(d);
            return set_error(CGLBM_ENOMEM, "device allocation of the moment state failed");
        }
    }
    hipError_t e = hipStreamCreate(&d->stream);
    if (e == hipSuccess) e = hipEventCreate(&d->ev0);
    if (e == hipSuccess) e = hipEventCreate(&d->ev1);
    if (e == hipSuccess) e = hipMalloc(&d->d_partials, sizeof(double) * kReduceBlocks * 6);
    if (e == hipSuccess) e = hipMalloc(&d->d_result, sizeof(double) * 8);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&d->plan.ncu, hipDeviceAttributeMultiprocessorCount, d->device);
    if (e == hipSuccess) {
        // steps between re-plans; CGLBM_PLAN_PERIOD (1..100000) overrides for tuning sweeps
        if (const char *pp = std::getenv("CGLBM_PLAN_PERIOD")) {
            const long v = std::strtol(pp, nullptr, 10);
            if (v >= 1 && v <= 100000) d->plan.period = static_cast<int>(v);
        }
        d->plan.cap = plan_capacity(*cfg, d->plan.ncu);
        e = hipMalloc(&d->plan.items, sizeof(int4) * d->plan.cap);
    }
    if (e == hipSuccess) e = hipMalloc(&d->plan.wmap, plan_wmap_bytes(*cfg));
    if (e == hipSuccess) e = hipMalloc(&d->plan.itime, sizeof(uint32_t) * d->plan.cap);
    for (auto &ev : d->plan.tev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipMemsetAsync(d->buf[0], 0, bytes, d->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d->buf[1], 0, bytes, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) {
        release(d);
        return herr(e, "domain setup");
    }
    make_views(d);
    cglbm_compact_layout(cfg, &d->layout);
    *out = d;
    return CGLBM_OK;
}

int cglbm_domain_destroy(cglbm_domain *dom) {
    if (!dom) return CGLBM_OK;
    if (dom->batch) return set_error(CGLBM_ESTATE, "remove the domain from its batch first");
    if (dom->stream) (void)hipStreamSynchronize(dom->stream);
    release(dom);
    return CGLBM_OK;
}

int cglbm_domain_moments(cglbm_domain *dom, int32_t which, cglbm_moments *out) {
    if (int rc = check_dom(dom)) return rc;
    if (!out || (which != 0 && which != 1)) return set_error(CGLBM_EINVAL, "bad arguments");
    if (dom->storage != CGLBM_STORE_F32)
        return set_error(CGLBM_EINVAL, "FP16-deviation states have no FP32 field views (use cglbm_copy_out)");
    *out = dom->view[which == 0 ? dom->cur : 1 - dom->cur];
    return CGLBM_OK;
}

int cglbm_domain_config(const cglbm_domain *dom, cglbm_config *out) {
    if (int rc = check_dom(dom)) return rc;
    if (!out) return set_error(CGLBM_EINVAL, "null output");
    *out = dom->cfg;
    return CGLBM_OK;
}

int cglbm_domain_set_kernel(cglbm_domain *dom, int32_t kernel) {
    if (int rc = check_dom(dom)) return rc;
    if (kernel < CGLBM_KERNEL_AUTO || kernel > CGLBM_KERNEL_MARCH) return set_error(CGLBM_EINVAL, "unknown kernel");
    dom->kernel = kernel;
    return CGLBM_OK;
}

int64_t cglbm_domain_step_count(const cglbm_domain *dom) { return dom ? dom->steps : -1; }

int cglbm_domain_set_tile(cglbm_domain *dom, int32_t ty, int32_t L, int32_t nbuf) {
    if (int rc = check_dom(dom)) return rc;
    if (ty == 0) {
        dom->shape = MarchShape{0, 0, 0};
        return CGLBM_OK;
    }
    if ((ty != 4 && ty != 8 && ty != 16) || L < 0 || L > dom->cfg.nz || (nbuf != 1 && nbuf != 2))
        return set_error(CGLBM_EINVAL, "tile rows must be 4, 8 or 16, 0 <= L <= nz, nbuf 1 or 2");
    dom->shape = effective_shape(dom->cfg.q, MarchShape{ty, L, nbuf});
    return CGLBM_OK;
}

int cglbm_domain_plan(cglbm_domain *dom, int32_t *items, int32_t cap, int32_t *n_items, int32_t *ty) {
    if (int rc = check_dom(dom)) return rc;
    if (!n_items || !ty || cap < 0 || (cap > 0 && !items)) return set_error(CGLBM_EINVAL, "bad arguments");
    HTRY(hipStreamSynchronize(dom->stream));
    const MarchPlanState &p = dom->plan;
    *ty = p.ty;
    *n_items = p.ty > 0 ? plan_items(dom->cfg, p.ty, p.ncu) : 0;
    if (*n_items > 0 && cap >= *n_items)
        HTRY(hipMemcpy(items, p.items, sizeof(int4) * static_cast<size_t>(*n_items), hipMemcpyDeviceToHost));
    return CGLBM_OK;
}

// Project autotuner: every (tile rows, chunk length) candidate runs a few updates from the current
// state into the other buffer (the state itself does not advance), timed with HIP events on the
// domain's stream; the fastest is kept for this domain and recorded for its configuration.
int cglbm_domain_autotune(cglbm_domain *dom, int32_t steps_per_candidate, int32_t *ty_out, int32_t *L_out,
                          int32_t *nbuf_out, double *ms_out) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    if (steps_per_candidate < 1 || steps_per_candidate > 1000) return set_error(CGLBM_EINVAL, "bad step count");
    const int tys[3] = {4, 8, 16};
    const int Ls[11] = {0, 2, 4, 6, 8, 12, 16, 24, 32, 48, 64};  // 0: planned chunks
    MarchShape best{0, 0, 0};
    double best_ms = 1e30;
    const int nxt = 1 - dom->cur;
    for (int nbuf = 1; nbuf <= 2; ++nbuf)
    for (int ty : tys)
        for (int L : Ls) {
            if (L > dom->cfg.nz) continue;
            const MarchShape m{ty, L, nbuf};
            if (effective_shape(dom->cfg.q, m).nbuf != nbuf) continue;  // runs as another candidate
            // warm-up launches (the first launch of a kernel pays a one-time cost), then the
            // timed ones; a planned candidate starts from a fresh plan and warms up through its
            // two planning steps (the map, then the measured correction) and one planned step,
            // so its timed steps are the steady state between re-plans
            if (L == 0) dom->plan.ty = 0;
            for (int w = 0; w < (L == 0 ? 3 : 1); ++w)
                if (int rc = launch_step(&dom->view[dom->cur], &dom->view[nxt], &dom->cfg, &dom->layout,
                                         dom->steps + 1, CGLBM_KERNEL_MARCH, &m, dom->stream, dom->storage,
                                         dom->fstride, &dom->plan))
                    return rc;
            HTRY(hipEventRecord(dom->ev0, dom->stream));
            for (int k = 0; k < steps_per_candidate; ++k)
                if (int rc = launch_step(&dom->view[dom->cur], &dom->view[nxt], &dom->cfg, &dom->layout,
                                         dom->steps + 1, CGLBM_KERNEL_MARCH, &m, dom->stream, dom->storage,
                                         dom->fstride, &dom->plan))
                    return rc;
            HTRY(hipEventRecord(dom->ev1, dom->stream));
            HTRY(hipEventSynchronize(dom->ev1));
            float ms = 0.0f;
            HTRY(hipEventElapsedTime(&ms, dom->ev0, dom->ev1));
            const double per = static_cast<double>(ms) / steps_per_candidate;
            if (per < best_ms) {
                best_ms = per;
                best = m;
            }
        }
    dom->shape = best;
    tuned_shape_store(dom->cfg, dom->storage, best);
    if (ty_out) *ty_out = best.ty;
    if (L_out) *L_out = best.L;
    if (nbuf_out) *nbuf_out = best.nbuf;
    if (ms_out) *ms_out = best_ms;
    return CGLBM_OK;
}

void *cglbm_domain_stream(cglbm_domain *dom) { return dom ? static_cast<void *>(dom->stream) : nullptr; }

// launch_InitStaticBubble (launch.cuh:14-20) with bubbleIn.cuh:15-52 (sharp) or the diffuse
// tanh profile phi = tanh(2(R - r)/w), blue inside. Centre in interior coordinates.
int cglbm_init_bubble(cglbm_domain *dom, double x0, double y0, double z0, double radius, double width,
                      int32_t sharp) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    std::vector<float> h;
    if (int rc = host_bubble(dom->cfg, x0, y0, z0, radius, width, sharp, h)) return rc;
    dom->steps = 0;
    if (dom->batch) batch_member_changed(dom->batch);
    return upload_init(dom, h, true);
}

// launch_InitJet (launch.cuh:52-72) followed by main.cu:26's boundary call with B == Aold.
int cglbm_init_jet(cglbm_domain *dom) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    std::vector<float> h;
    if (int rc = host_jet(dom->cfg, h)) return rc;
    dom->steps = 0;
    if (dom->batch) batch_member_changed(dom->batch);
    return upload_init(dom, h, true);
}

// launch_InitRTI (launch.cuh:42-46): rtiIn.cuh:18-71 evaluated in double (plus optional seeded
// random modes), then the wall planes (rtiIn.cuh:77-119).
int cglbm_init_rti(cglbm_domain *dom, double amplitude, double width, int32_t n_random_modes,
                   double random_amplitude, uint32_t seed) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    std::vector<float> h;
    if (int rc = host_rti(dom->cfg, amplitude, width, n_random_modes, random_amplitude, seed, h)) return rc;
    dom->steps = 0;
    if (dom->batch) batch_member_changed(dom->batch);
    return upload_init(dom, h, dom->cfg.ybc == CGLBM_Y_WALLS);
}

// The same initial states on caller-owned arrays (launch_InitStaticBubble, launch_InitJet,
// launch_InitRTI of launch.cuh:14-72, which take a MomentsDevice): the host-built state is
// uploaded to a staging area and scattered into the arrays at `layout`; cells the layout does not
// address (the reference's ghost columns) are left as they are. Synchronous on `stream`, as the
// reference's launchers synchronise the device.
int cglbm_launch_init_bubble(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, double x0,
                             double y0, double z0, double radius, double width, int32_t sharp, void *stream) {
    if (int rc = check_launch_init(A, cfg, layout)) return rc;
    std::vector<float> h;
    if (int rc = host_bubble(*cfg, x0, y0, z0, radius, width, sharp, h)) return rc;
    return upload_to_layout(A, *cfg, *layout, h, true, static_cast<hipStream_t>(stream));
}

int cglbm_launch_init_jet(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, void *stream) {
    if (int rc = check_launch_init(A, cfg, layout)) return rc;
    std::vector<float> h;
    if (int rc = host_jet(*cfg, h)) return rc;
    return upload_to_layout(A, *cfg, *layout, h, true, static_cast<hipStream_t>(stream));
}

int cglbm_launch_init_rti(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout,
                          double amplitude, double width, int32_t n_random_modes, double random_amplitude,
                          uint32_t seed, void *stream) {
    if (int rc = check_launch_init(A, cfg, layout)) return rc;
    std::vector<float> h;
    if (int rc = host_rti(*cfg, amplitude, width, n_random_modes, random_amplitude, seed, h)) return rc;
    return upload_to_layout(A, *cfg, *layout, h, cfg->ybc == CGLBM_Y_WALLS, static_cast<hipStream_t>(stream));
}

// main.cu:38-55: update, boundary planes, swap.
int cglbm_step(cglbm_domain *dom, int32_t nsteps) {
    if (int rc = check_dom(dom)) return rc;
    if (dom->batch) return set_error(CGLBM_ESTATE, "domain is a batch member; step it through its batch");
    if (nsteps < 0) return set_error(CGLBM_EINVAL, "negative step count");
    for (int k = 0; k < nsteps; ++k) {
        const int nxt = 1 - dom->cur;
        if (int rc = launch_step(&dom->view[dom->cur], &dom->view[nxt], &dom->cfg, &dom->layout, dom->steps + 1,
                                 dom->kernel, &dom->shape, dom->stream, dom->storage, dom->fstride, &dom->plan))
            return rc;
        dom->cur = nxt;
        ++dom->steps;
    }
    return CGLBM_OK;
}

int cglbm_time_steps(cglbm_domain *dom, int32_t nsteps, double *ms_total) {
    if (int rc = check_dom(dom)) return rc;
    if (!ms_total) return set_error(CGLBM_EINVAL, "null output");
    HTRY(hipEventRecord(dom->ev0, dom->stream));
    if (int rc = cglbm_step(dom, nsteps)) return rc;
    HTRY(hipEventRecord(dom->ev1, dom->stream));
    HTRY(hipEventSynchronize(dom->ev1));
    float ms = 0.0f;
    HTRY(hipEventElapsedTime(&ms, dom->ev0, dom->ev1));
    *ms_total = ms;
    return CGLBM_OK;
}

int cglbm_sync(cglbm_domain *dom) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    HTRY(hipStreamSynchronize(dom->stream));
    return CGLBM_OK;
}

// copy_out_D2H (memory.cuh:209-217), any subset of the 11 fields. FP16 states decode one field at
// a time through a one-field staging area (stream order keeps each copy ahead of the next decode).
int cglbm_copy_out(cglbm_domain *dom, uint32_t field_mask, float *host) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    if (!host || (field_mask & ~0x7FFu)) return set_error(CGLBM_EINVAL, "bad arguments");
    const int64_t n = stored_cells(dom->cfg);
    const bool f16 = dom->storage == CGLBM_STORE_F16;
    Staging st;
    if (f16 && (field_mask & ~0x3u))
        if (int rc = make_staging(dom, st, 1)) return rc;
    int k = 0;
    for (int f = 0; f < 11; ++f) {
        if (!(field_mask & (1u << f))) continue;
        const float *src = dom->buf[dom->cur] + f * dom->fstride;
        if (f16 && f >= 2) {
            HTRY(launch_unpack_f16(dom->buf[dom->cur], dom->fstride, st.p, dom->fstride, n, f, f + 1, dom->stream));
            src = st.p;
        }
        HTRY(hipMemcpyAsync(host + (k++) * n, src, sizeof(float) * n, hipMemcpyDeviceToHost, dom->stream));
    }
    HTRY(hipStreamSynchronize(dom->stream));
    return CGLBM_OK;
}

int cglbm_copy_in(cglbm_domain *dom, uint32_t field_mask, const float *host) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    if (!host || (field_mask & ~0x7FFu)) return set_error(CGLBM_EINVAL, "bad arguments");
    const int64_t n = stored_cells(dom->cfg);
    const bool f16 = dom->storage == CGLBM_STORE_F16;
    dom->plan.ty = 0;
    dom->plan.trial = kTrialNone;
    dom->plan.trial_period = kTrialFirstPeriod;
    Staging st;
    if (f16 && (field_mask & ~0x3u))
        if (int rc = make_staging(dom, st, 1)) return rc;
    int k = 0;
    for (int f = 0; f < 11; ++f) {
        if (!(field_mask & (1u << f))) continue;
        const float *src = host + (k++) * n;
        if (f16 && f >= 2) {
            HTRY(hipMemcpyAsync(st.p, src, sizeof(float) * n, hipMemcpyHostToDevice, dom->stream));
            HTRY(launch_pack_f16(st.p, dom->fstride, dom->buf[dom->cur], dom->fstride, n, f, f + 1, dom->stream));
        } else {
            HTRY(hipMemcpyAsync(dom->buf[dom->cur] + f * dom->fstride, src, sizeof(float) * n, hipMemcpyHostToDevice,
                                dom->stream));
        }
    }
    HTRY(hipStreamSynchronize(dom->stream));
    return CGLBM_OK;
}

int cglbm_reduce(cglbm_domain *dom, double out[6]) {
    if (int rc = check_dom(dom)) return rc;
    if (int rc = wait_batch(dom)) return rc;
    if (!out) return set_error(CGLBM_EINVAL, "null output");
    const cglbm_config &c = dom->cfg;
    const int y0 = c.ybc == CGLBM_Y_PERIODIC ? 0 : 1;
    const uint32_t py = static_cast<uint32_t>(dom->layout.pitch_y), pz = static_cast<uint32_t>(dom->layout.pitch_z);
    if (dom->storage == CGLBM_STORE_F16) {
        HTRY(launch_reduce_f16(dom->buf[dom->cur], dom->fstride, c.nx, c.ny, c.nz, y0, py, pz, dom->d_partials,
                               dom->d_result, dom->stream));
    } else {
        float *const *v = reinterpret_cast<float *const *>(&dom->view[dom->cur]);
        HTRY(launch_reduce(v, c.nx, c.ny, c.nz, y0, py, pz, dom->d_partials, dom->d_result, dom->stream));
    }
    HTRY(hipMemcpyAsync(out, dom->d_result, sizeof(double) * 6, hipMemcpyDeviceToHost, dom->stream));
    HTRY(hipStreamSynchronize(dom->stream));
    return CGLBM_OK;
}

// write_vti_step_device (io/vtiWriter.cuh:151-164): copy rhor rhob ux uy uz of the current state
// to the host and print them over the stored extent with the reference's text layout.
int cglbm_write_vti(cglbm_domain *dom, const char *path) {
    if (int rc = check_dom(dom)) return rc;
    if (!path) return set_error(CGLBM_EINVAL, "null path");
    const cglbm_config &c = dom->cfg;
    const int64_t n = stored_cells(c);
    std::vector<float> h(5 * n);
    if (int rc = cglbm_copy_out(dom, 0x1Fu, h.data())) return rc;
    const float *const f[5] = {h.data(), h.data() + n, h.data() + 2 * n, h.data() + 3 * n, h.data() + 4 * n};
    return write_vti_arrays(path, c.nx, stored_ny(c), c.nz, f);
}

}  // extern "C"
