// Batches: many domains stepped together, one launch per kernel class (stencil, y boundary,
// body force) per step, over a flattened list of {domain, tile, chunk} work items. Steps can be
// issued as replays of a captured multi-step HIP graph. Each work item is a pure function of its
// domain's state, so a member's results do not depend on what else is in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <new>
#include <string>
#include <vector>

#include "cglbm.h"
#include "domain_internal.hpp"
#include "kernels.hpp"
#include "state.hpp"
#include "storage.hpp"

using namespace cglbm;

namespace {

struct Class {
    int q, ybc, forced, store;
    std::vector<cglbm_domain *> members;
    MarchShape shape{8, 16, 2};
    bool shape_set = false;
    int desc_off = 0, item_off = 0, nitems = 0;
};

}  // namespace

struct cglbm_batch {
    std::vector<cglbm_domain *> members;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    BatchDesc *d_descs = nullptr;
    int4 *d_items = nullptr;
    int64_t *d_base = nullptr;
    size_t cap_descs = 0, cap_items = 0;
    std::vector<Class> classes;
    bool dirty = true;
    MarchShape forced_shape{0, 0, 0};  // ty == 0: per-class autotuned/default
    // graph bundles
    int bundle = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // side streams: the kernel classes of one step run concurrently (fork from `stream`, join back)
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kSide] = {nullptr, nullptr, nullptr};
};

namespace {

int herr(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(CGLBM_EHIP, m.c_str());
}

#define HTRY(expr)                                    \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return herr(e_, #expr); \
    } while (0)

void drop_graph(cglbm_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

MarchShape class_shape(const cglbm_batch *b, const Class &c) {
    if (b->forced_shape.ty > 0) return effective_shape(c.q, b->forced_shape);
    if (c.shape_set) return c.shape;
    // Untuned class defaults, from class-by-class tile sweeps of batched members on MI355X:
    // 16-row tiles for D3Q27 and periodic D3Q19, 4-row tiles for bounded-y D3Q19.
    if (c.q == 27) return MarchShape{16, 16, 1};
    if (c.ybc != 0) return MarchShape{4, 16, 1};
    return MarchShape{16, 32, 1};
}

// Rebuild descriptors and work items from the current member states (synchronous).
int rebuild(cglbm_batch *b) {
    HTRY(hipStreamSynchronize(b->stream));
    drop_graph(b);
    std::vector<Class> old = std::move(b->classes);
    b->classes.clear();
    for (cglbm_domain *d : b->members) {
        const int forced = d->cfg.gravity != 0.0f ? 1 : 0;
        auto it = std::find_if(b->classes.begin(), b->classes.end(), [&](const Class &c) {
            return c.q == d->cfg.q && c.ybc == (d->cfg.ybc == CGLBM_Y_PERIODIC ? 0 : 1) && c.forced == forced &&
                   c.store == d->storage;
        });
        if (it == b->classes.end()) {
            Class c;
            c.q = d->cfg.q;
            c.ybc = d->cfg.ybc == CGLBM_Y_PERIODIC ? 0 : 1;
            c.forced = forced;
            c.store = d->storage;
            for (const Class &o : old)
                if (o.q == c.q && o.ybc == c.ybc && o.forced == c.forced && o.store == c.store && o.shape_set) {
                    c.shape = o.shape;
                    c.shape_set = true;
                }
            b->classes.push_back(c);
            it = b->classes.end() - 1;
        }
        it->members.push_back(d);
    }
    std::vector<BatchDesc> descs;
    std::vector<int4> items;
    for (Class &c : b->classes) {
        const MarchShape sh = class_shape(b, c);
        c.desc_off = static_cast<int>(descs.size());
        c.item_off = static_cast<int>(items.size());
        for (cglbm_domain *d : c.members) {
            StepArgs sa = make_step_args(&d->view[0], &d->view[1], &d->cfg, &d->layout, 0);
            BatchDesc bd{};
            const bool ok = d->storage == CGLBM_STORE_F16 ? march_args_f16(sa, d->buf[0], d->buf[1], d->fstride, bd.args)
                                                          : march_args(sa, bd.args);
            if (!ok) return set_error(CGLBM_EINVAL, "batch member not addressable by the fast path");
            if (bd.args.store == kStoreF32Ptr)
                return set_error(CGLBM_EINVAL, "batch members must stay below 4 GiB per state");
            bd.buf0 = d->buf[0];
            bd.buf1 = d->buf[1];
            bd.cur0 = d->cur;
            bd.step0 = d->steps;
            const int L = std::min(sh.L, d->cfg.nz);
            bd.nchunk = (d->cfg.nz + L - 1) / L;
            // descriptor index relative to the class: the launch passes descs + desc_off
            const int di = static_cast<int>(descs.size()) - c.desc_off;
            const int ntx = (d->cfg.nx + 31) / 32, nty = (d->cfg.ny + sh.ty - 1) / sh.ty;
            for (int ch = 0; ch < bd.nchunk; ++ch)
                for (int ty = 0; ty < nty; ++ty)
                    for (int tx = 0; tx < ntx; ++tx) items.push_back(make_int4(di, tx, ty, ch));
            descs.push_back(bd);
        }
        c.nitems = static_cast<int>(items.size()) - c.item_off;
        const int ndesc = static_cast<int>(descs.size()) - c.desc_off;
        for (int i = c.item_off; i < c.item_off + c.nitems; ++i)
            if (items[i].x < 0 || items[i].x >= ndesc) return set_error(CGLBM_EINVAL, "batch work item out of range");
    }
    if (descs.size() > b->cap_descs) {
        if (b->d_descs) HTRY(hipFree(b->d_descs));
        b->d_descs = nullptr;
        HTRY(hipMalloc(&b->d_descs, sizeof(BatchDesc) * descs.size()));
        b->cap_descs = descs.size();
    }
    if (items.size() > b->cap_items) {
        if (b->d_items) HTRY(hipFree(b->d_items));
        b->d_items = nullptr;
        HTRY(hipMalloc(&b->d_items, sizeof(int4) * items.size()));
        b->cap_items = items.size();
    }
    if (!descs.empty())
        HTRY(hipMemcpyAsync(b->d_descs, descs.data(), sizeof(BatchDesc) * descs.size(), hipMemcpyHostToDevice,
                            b->stream));
    if (!items.empty())
        HTRY(hipMemcpyAsync(b->d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice,
                            b->stream));
    const int64_t zero = 0;
    HTRY(hipMemcpyAsync(b->d_base, &zero, sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
    HTRY(hipStreamSynchronize(b->stream));
    b->dirty = false;
    return CGLBM_OK;
}

// Enqueue step k of a bundle for every class (no host synchronisation).
int ensure_side(cglbm_batch *b) {
    if (b->fork) return CGLBM_OK;
    for (int i = 0; i < cglbm_batch::kSide; ++i) {
        HTRY(hipStreamCreate(&b->side[i]));
        HTRY(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
    }
    HTRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    return CGLBM_OK;
}

// Enqueue step k of a bundle for every class (no host synchronisation). With several classes the
// first runs on the batch stream and the others on side streams forked from it and joined back, so
// small classes do not leave the GPU idle while they run; graph capture records the fork and join
// as dependencies.
int enqueue_step(cglbm_batch *b, int k) {
    int nclass = 0;
    for (const Class &c : b->classes) nclass += c.nitems > 0 ? 1 : 0;
    const bool fan = nclass > 1;
    bool used[cglbm_batch::kSide] = {false, false, false};
    if (fan) {
        if (int rc = ensure_side(b)) return rc;
        HTRY(hipEventRecord(b->fork, b->stream));
    }
    int slot = 0;
    for (const Class &c : b->classes) {
        if (c.nitems == 0) continue;
        hipStream_t s = b->stream;
        if (fan && slot > 0) {
            const int j = (slot - 1) % cglbm_batch::kSide;
            s = b->side[j];
            if (!used[j]) HTRY(hipStreamWaitEvent(s, b->fork, 0));
            used[j] = true;
        }
        ++slot;
        BatchLaunch L{};
        L.descs = b->d_descs + c.desc_off;
        L.items = b->d_items + c.item_off;
        L.nitems = c.nitems;
        L.base = b->d_base;
        L.k = k;
        L.grid = (c.nitems + 7) / 8 * 8;  // one block per work item
        L.ybc = c.ybc;
        L.forced = c.forced;
        L.store = c.store;
        L.shape = class_shape(b, c);
        const hipError_t e = launch_batch(L, c.q, s);
        if (e != hipSuccess) return herr(e, "batch launch");
    }
    for (int j = 0; j < cglbm_batch::kSide; ++j) {
        if (!used[j]) continue;
        HTRY(hipEventRecord(b->join[j], b->side[j]));
        HTRY(hipStreamWaitEvent(b->stream, b->join[j], 0));
    }
    return CGLBM_OK;
}

void advance_members(cglbm_batch *b, int n) {
    for (cglbm_domain *d : b->members) {
        d->steps += n;
        if (n & 1) d->cur = 1 - d->cur;
    }
}

int capture(cglbm_batch *b) {
    drop_graph(b);
    HTRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = CGLBM_OK;
    for (int k = 0; k < b->bundle && rc == CGLBM_OK; ++k) rc = enqueue_step(b, k);
    hipError_t e = hipSuccess;
    if (rc == CGLBM_OK) e = launch_advance(b->d_base, b->bundle, b->stream);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (rc != CGLBM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return herr(e, "launch_advance");
    if (e2 != hipSuccess) return herr(e2, "hipStreamEndCapture");
    b->graph = g;
    HTRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return CGLBM_OK;
}

int check_batch(const cglbm_batch *b) {
    if (!b) return set_error(CGLBM_EINVAL, "null batch");
    return CGLBM_OK;
}

}  // namespace

void batch_member_changed(cglbm_batch *b) {
    if (b) b->dirty = true;
}

hipStream_t batch_stream(cglbm_batch *b) { return b ? b->stream : nullptr; }

extern "C" {

int cglbm_batch_create(cglbm_batch **out) {
    if (!out) return set_error(CGLBM_EINVAL, "null output");
    *out = nullptr;
    cglbm_batch *b = new (std::nothrow) cglbm_batch;
    if (!b) return set_error(CGLBM_ENOMEM, "host allocation failed");
    hipError_t e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipMalloc(&b->d_base, sizeof(int64_t) * 2);
    if (e != hipSuccess) {
        if (b->d_base) (void)hipFree(b->d_base);
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        if (b->ev1) (void)hipEventDestroy(b->ev1);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return herr(e, "batch setup");
    }
    *out = b;
    return CGLBM_OK;
}

int cglbm_batch_destroy(cglbm_batch *b) {
    if (!b) return CGLBM_OK;
    (void)hipStreamSynchronize(b->stream);
    for (cglbm_domain *d : b->members) {
        d->batch = nullptr;
        d->batch_stream = nullptr;
    }
    drop_graph(b);
    if (b->d_descs) (void)hipFree(b->d_descs);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_base) (void)hipFree(b->d_base);
    (void)hipEventDestroy(b->ev0);
    (void)hipEventDestroy(b->ev1);
    for (int i = 0; i < cglbm_batch::kSide; ++i) {
        if (b->side[i]) (void)hipStreamSynchronize(b->side[i]);
        if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
        if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return CGLBM_OK;
}

int cglbm_batch_add(cglbm_batch *b, cglbm_domain *d) {
    if (int rc = check_batch(b)) return rc;
    if (!d) return set_error(CGLBM_EINVAL, "null domain");
    if (d->batch) return set_error(CGLBM_ESTATE, "domain already belongs to a batch");
    if (d->cfg.gravity != 0.0f && d->cfg.ybc == CGLBM_Y_PERIODIC)
        return set_error(CGLBM_EINVAL, "a forced periodic domain has no batched kernel");
    HTRY(hipStreamSynchronize(d->stream));
    HTRY(hipStreamSynchronize(b->stream));
    d->batch = b;
    d->batch_stream = b->stream;
    b->members.push_back(d);
    b->dirty = true;
    return CGLBM_OK;
}

int cglbm_batch_remove(cglbm_batch *b, cglbm_domain *d) {
    if (int rc = check_batch(b)) return rc;
    auto it = std::find(b->members.begin(), b->members.end(), d);
    if (it == b->members.end()) return set_error(CGLBM_EINVAL, "domain is not a member of this batch");
    HTRY(hipStreamSynchronize(b->stream));
    b->members.erase(it);
    d->batch = nullptr;
    d->batch_stream = nullptr;
    d->plan.ty = 0;  // its state moved on in the batch: a planned shape records a fresh map
    b->dirty = true;
    return CGLBM_OK;
}

int32_t cglbm_batch_size(const cglbm_batch *b) { return b ? static_cast<int32_t>(b->members.size()) : -1; }

int cglbm_batch_set_tile(cglbm_batch *b, int32_t ty, int32_t L, int32_t nbuf) {
    if (int rc = check_batch(b)) return rc;
    if (ty != 0 && ((ty != 4 && ty != 8 && ty != 16) || L < 1 || (nbuf != 1 && nbuf != 2)))
        return set_error(CGLBM_EINVAL, "tile rows must be 4, 8 or 16, L >= 1, nbuf 1 or 2");
    b->forced_shape = ty == 0 ? MarchShape{0, 0, 0} : MarchShape{ty, L, nbuf};
    b->dirty = true;
    return CGLBM_OK;
}

int cglbm_batch_set_graph(cglbm_batch *b, int32_t bundle) {
    if (int rc = check_batch(b)) return rc;
    if (bundle < 0 || bundle > 4096) return set_error(CGLBM_EINVAL, "bundle must be in [0, 4096]");
    HTRY(hipStreamSynchronize(b->stream));
    drop_graph(b);
    b->bundle = bundle;
    return CGLBM_OK;
}

int cglbm_batch_step(cglbm_batch *b, int32_t nsteps) {
    if (int rc = check_batch(b)) return rc;
    if (nsteps < 0) return set_error(CGLBM_EINVAL, "negative step count");
    if (b->members.empty() || nsteps == 0) return CGLBM_OK;
    if (b->dirty)
        if (int rc = rebuild(b)) return rc;
    int left = nsteps;
    if (b->bundle > 0) {
        if (!b->exec)
            if (int rc = capture(b)) return rc;
        while (left >= b->bundle) {
            HTRY(hipGraphLaunch(b->exec, b->stream));
            advance_members(b, b->bundle);
            left -= b->bundle;
        }
    }
    if (left > 0) {
        for (int k = 0; k < left; ++k)
            if (int rc = enqueue_step(b, k)) return rc;
        HTRY(launch_advance(b->d_base, left, b->stream));
        advance_members(b, left);
    }
    return CGLBM_OK;
}

int cglbm_batch_time_steps(cglbm_batch *b, int32_t nsteps, double *ms_total) {
    if (int rc = check_batch(b)) return rc;
    if (!ms_total) return set_error(CGLBM_EINVAL, "null output");
    if (b->dirty)
        if (int rc = rebuild(b)) return rc;
    if (b->bundle > 0 && !b->exec)
        if (int rc = capture(b)) return rc;
    HTRY(hipEventRecord(b->ev0, b->stream));
    if (int rc = cglbm_batch_step(b, nsteps)) return rc;
    HTRY(hipEventRecord(b->ev1, b->stream));
    HTRY(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    HTRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    *ms_total = ms;
    return CGLBM_OK;
}

// Baseline for the batch: every member stepped by its own plain launches, one after another on
// the batch stream (the same single-domain kernels, no batching).
int cglbm_batch_time_sequential(cglbm_batch *b, int32_t nsteps, double *ms_total) {
    if (int rc = check_batch(b)) return rc;
    if (!ms_total || nsteps < 0) return set_error(CGLBM_EINVAL, "bad arguments");
    HTRY(hipStreamSynchronize(b->stream));
    HTRY(hipEventRecord(b->ev0, b->stream));
    for (int k = 0; k < nsteps; ++k)
        for (cglbm_domain *d : b->members) {
            const int nxt = 1 - d->cur;
            if (int rc = launch_step(&d->view[d->cur], &d->view[nxt], &d->cfg, &d->layout, d->steps + 1, d->kernel,
                                     &d->shape, b->stream, d->storage, d->fstride))
                return rc;
            d->cur = nxt;
            ++d->steps;
        }
    HTRY(hipEventRecord(b->ev1, b->stream));
    HTRY(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    HTRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    *ms_total = ms;
    b->dirty = true;  // members moved outside the batch's descriptors
    return CGLBM_OK;
}

int cglbm_batch_sync(cglbm_batch *b) {
    if (int rc = check_batch(b)) return rc;
    HTRY(hipStreamSynchronize(b->stream));
    return CGLBM_OK;
}

// Batch autotuner: per kernel class, every (tile rows, chunk length, exchange buffers) candidate
// is timed over a few batched steps from the current states into the spare buffers (the members
// do not advance); the fastest is kept for the class.
int cglbm_batch_autotune(cglbm_batch *b, int32_t steps_per_candidate) {
    if (int rc = check_batch(b)) return rc;
    if (steps_per_candidate < 1 || steps_per_candidate > 100) return set_error(CGLBM_EINVAL, "bad step count");
    if (b->members.empty()) return CGLBM_OK;
    b->forced_shape = MarchShape{0, 0, 0};
    if (int rc = rebuild(b)) return rc;
    const int tys[3] = {4, 8, 16};
    const int Ls[8] = {4, 8, 12, 16, 24, 32, 48, 64};
    for (size_t ci = 0; ci < b->classes.size(); ++ci) {
        MarchShape best{0, 0, 0};
        double best_ms = 1e30;
        for (int nbuf = 1; nbuf <= 2; ++nbuf)
            for (int ty : tys)
                for (int L : Ls) {
                    if (effective_shape(b->classes[ci].q, MarchShape{ty, L, nbuf}).nbuf != nbuf) continue;
                    b->classes[ci].shape = MarchShape{ty, L, nbuf};
                    b->classes[ci].shape_set = true;
                    if (int rc = rebuild(b)) return rc;
                    // Timing launches read each member's current state and write its spare one
                    // (k fixed at 0, device base untouched): states do not advance.
                    const Class &c = b->classes[ci];
                    BatchLaunch Lh{};
                    Lh.descs = b->d_descs + c.desc_off;
                    Lh.items = b->d_items + c.item_off;
                    Lh.nitems = c.nitems;
                    Lh.base = b->d_base;
                    Lh.k = 0;
                    Lh.grid = (c.nitems + 7) / 8 * 8;
                    Lh.ybc = c.ybc;
                    Lh.forced = c.forced;
                    Lh.store = c.store;
                    Lh.shape = c.shape;
                    auto go = [&]() { return launch_batch(Lh, c.q, b->stream); };
                    HTRY(go());
                    HTRY(hipEventRecord(b->ev0, b->stream));
                    for (int k = 0; k < steps_per_candidate; ++k) HTRY(go());
                    HTRY(hipEventRecord(b->ev1, b->stream));
                    HTRY(hipEventSynchronize(b->ev1));
                    float ms = 0.0f;
                    HTRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c.shape;
                    }
                }
        b->classes[ci].shape = best;
        b->classes[ci].shape_set = true;
    }
    return rebuild(b);
}

}  // extern "C"
