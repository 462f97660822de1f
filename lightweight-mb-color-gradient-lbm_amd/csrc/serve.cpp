// Continuous-batching server over library domains and one batch. The reference has no request
// loop (main.cu runs one case per process); this is the serving layer around its update.
//
// Threads:
//   admission workers  pop the FIFO queue when the front case fits the HBM budget and a slot is
//                      free, create and initialise the domain on its own stream (overlapping the
//                      batch's stepping), and hand it to the engine;
//   engine             owns the batch: adds ready cases, steps every resident case together in
//                      bundles of min(bundle, smallest remaining step count) with at most two
//                      bundles in flight, and removes cases that reached their step count;
//   completion         copies results to the caller, computes the FP64 diagnostics, destroys the
//                      domain and releases its budget.
// Budget: a case reserves its two states and reduce scratch for its life, plus during
// initialisation of an FP16 case the FP32 staging state, plus one FP32 field for the result
// copy of an FP16 case.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "cglbm.h"
#include "domain_internal.hpp"
#include "state.hpp"
#include "storage.hpp"

using namespace cglbm;

namespace {

struct Job {
    cglbm_request req{};
    std::vector<float> initial;  // CGLBM_CASE_STATE
    cglbm_domain *dom = nullptr;
    int64_t remaining = 0;
    int64_t bytes_life = 0;  // states + scratch (+ result staging for FP16)
    int64_t bytes_init = 0;  // FP32 staging while an FP16 case is initialised
    cglbm_result res{};
};

void fill_message(cglbm_result &r, int rc, const char *what) {
    r.status = rc;
    std::snprintf(r.message, sizeof(r.message), "%s: %s", what, cglbm_last_error());
}

}  // namespace

struct cglbm_server {
    cglbm_server_opts opts{};
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    cglbm_batch *batch = nullptr;

    std::mutex mu;
    std::condition_variable cv_admit, cv_engine, cv_complete, cv_done;
    std::deque<Job *> pending, ready, finishing;
    std::deque<cglbm_result> done;
    cglbm_server_stats st{};
    uint64_t next_id = 1;
    bool stop = false;         // no new admissions; the engine finishes resident cases
    bool admit_done = false;   // admission workers have exited (set after joining them)
    bool engine_done = false;  // the engine has exited; completion drains and exits
    int engine_rc = CGLBM_OK;

    std::vector<std::thread> admitters;
    std::thread engine, completer;

    double clock() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    bool fits(const Job *j) const {
        return st.bytes_reserved + j->bytes_life + j->bytes_init <= st.budget_bytes && st.resident < opts.max_members;
    }
    // Caller holds mu.
    void finish_locked(Job *j) {
        st.bytes_reserved -= j->bytes_life;
        --st.resident;
        if (j->res.status == CGLBM_OK) {
            ++st.completed;
            st.updates += j->res.cells * j->res.steps;
        } else {
            ++st.failed;
        }
        done.push_back(j->res);
        delete j;
        cv_done.notify_all();
        cv_admit.notify_all();
    }

    void admit_loop();
    void engine_loop();
    void complete_loop();
};

// ---- admission ------------------------------------------------------------------------------
void cglbm_server::admit_loop() {
    (void)hipSetDevice(device);
    for (;;) {
        std::unique_lock<std::mutex> lk(mu);
        cv_admit.wait(lk, [&] { return stop || (!pending.empty() && fits(pending.front())); });
        if (stop) return;
        Job *j = pending.front();
        pending.pop_front();
        --st.queued;
        ++st.admitted;
        ++st.resident;
        st.peak_resident = std::max(st.peak_resident, st.resident);
        st.bytes_reserved += j->bytes_life + j->bytes_init;
        st.peak_bytes_reserved = std::max(st.peak_bytes_reserved, st.bytes_reserved);
        lk.unlock();

        j->res.t_admit = clock();
        const cglbm_request &q = j->req;
        int rc = cglbm_domain_create(&q.cfg, q.storage, &j->dom);
        if (rc == CGLBM_OK) {
            const cglbm_config &c = q.cfg;
            switch (q.case_kind) {
            case CGLBM_CASE_BUBBLE:
                rc = cglbm_init_bubble(j->dom, (c.nx - 1) / 2.0, (c.ny - 1) / 2.0, (c.nz - 1) / 2.0, q.radius, q.width,
                                       q.sharp);
                break;
            case CGLBM_CASE_JET: rc = cglbm_init_jet(j->dom); break;
            case CGLBM_CASE_RTI:
                rc = cglbm_init_rti(j->dom, q.amplitude, q.width, q.n_random_modes, q.random_amplitude, q.init_seed);
                break;
            default:
                rc = cglbm_copy_in(j->dom, 0x7FFu, j->initial.data());
                j->initial = std::vector<float>();
                break;
            }
            if (rc != CGLBM_OK) fill_message(j->res, rc, "case initialisation");
        } else {
            fill_message(j->res, rc, "domain creation");
        }
        if (rc != CGLBM_OK && j->dom) {
            (void)cglbm_domain_destroy(j->dom);
            j->dom = nullptr;
        }
        lk.lock();
        st.bytes_reserved -= j->bytes_init;
        j->bytes_init = 0;
        if (rc != CGLBM_OK) {
            j->res.t_done = clock();
            finish_locked(j);
        } else {
            ready.push_back(j);
            cv_engine.notify_one();
        }
        cv_admit.notify_all();
    }
}

// ---- engine -----------------------------------------------------------------------------------
void cglbm_server::engine_loop() {
    (void)hipSetDevice(device);
    std::vector<Job *> run;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t bs = nullptr;
    int rc = cglbm_batch_create(&batch);
    if (rc == CGLBM_OK) rc = cglbm_batch_set_graph(batch, opts.bundle);
    if (rc == CGLBM_OK && opts.tile_ty > 0) rc = cglbm_batch_set_tile(batch, opts.tile_ty, opts.tile_L, opts.tile_nbuf);
    if (rc == CGLBM_OK && (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
                           hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess))
        rc = CGLBM_EHIP;
    if (rc == CGLBM_OK) bs = batch_stream(batch);
    int ei = 0;
    bool in_flight = false;
    double t_last = clock();

    auto to_completion = [&](Job *j) {
        std::lock_guard<std::mutex> g(mu);
        finishing.push_back(j);
        cv_complete.notify_one();
    };
    auto fail_job = [&](Job *j, int code, const char *what) {
        if (j->res.status == CGLBM_OK) fill_message(j->res, code, what);
        to_completion(j);
    };

    for (;;) {
        std::vector<Job *> add;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (run.empty()) {
                cv_engine.wait(lk, [&] { return admit_done || !ready.empty(); });
                t_last = clock();
            }
            if (run.empty() && ready.empty() && admit_done) break;
            add.assign(ready.begin(), ready.end());
            ready.clear();
        }
        for (Job *j : add) {
            if (rc != CGLBM_OK) {
                fail_job(j, rc, "server engine");
                continue;
            }
            j->res.t_start = clock();
            if (j->remaining == 0) {
                to_completion(j);
                continue;
            }
            const int arc = cglbm_batch_add(batch, j->dom);
            if (arc != CGLBM_OK) {
                fail_job(j, arc, "batch admission");
                continue;
            }
            run.push_back(j);
        }
        if (run.empty()) continue;

        int64_t n = opts.bundle;
        for (Job *j : run) n = std::min(n, j->remaining);
        int src = cglbm_batch_step(batch, static_cast<int32_t>(n));
        if (src == CGLBM_OK && hipEventRecord(ev[ei], bs) != hipSuccess) src = CGLBM_EHIP;
        // at most two step calls in flight: wait for the previous one
        if (src == CGLBM_OK && in_flight && hipEventSynchronize(ev[ei ^ 1]) != hipSuccess) src = CGLBM_EHIP;
        ei ^= 1;
        in_flight = true;
        const double t = clock();
        {
            std::lock_guard<std::mutex> g(mu);
            ++st.bundles;
            st.busy_seconds += t - t_last;
        }
        t_last = t;
        std::vector<Job *> keep;
        for (Job *j : run) {
            if (src != CGLBM_OK) {
                (void)cglbm_batch_remove(batch, j->dom);
                fail_job(j, src, "batched step");
                continue;
            }
            j->remaining -= n;
            j->res.steps += n;
            if (j->remaining > 0) {
                keep.push_back(j);
                continue;
            }
            const int rrc = cglbm_batch_remove(batch, j->dom);  // waits for the batch stream
            if (rrc != CGLBM_OK) fill_message(j->res, rrc, "batch removal");
            to_completion(j);
        }
        run.swap(keep);
        if (run.empty()) in_flight = false;
        if (src != CGLBM_OK) rc = src;
    }
    if (batch) (void)cglbm_batch_sync(batch);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    std::lock_guard<std::mutex> g(mu);
    engine_rc = rc;
    engine_done = true;
    cv_complete.notify_all();
}

// ---- completion ---------------------------------------------------------------------------------
void cglbm_server::complete_loop() {
    (void)hipSetDevice(device);
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_complete.wait(lk, [&] { return !finishing.empty() || engine_done; });
            if (finishing.empty()) return;
            j = finishing.front();
            finishing.pop_front();
        }
        if (j->res.status == CGLBM_OK && j->req.result_mask) {
            const int rc = cglbm_copy_out(j->dom, j->req.result_mask, j->req.result);
            if (rc != CGLBM_OK) fill_message(j->res, rc, "result copy");
        }
        if (j->res.status == CGLBM_OK) {
            const int rc = cglbm_reduce(j->dom, j->res.diag);
            if (rc != CGLBM_OK) fill_message(j->res, rc, "diagnostics");
        }
        if (j->dom) (void)cglbm_domain_destroy(j->dom);
        j->dom = nullptr;
        j->res.t_done = clock();
        std::lock_guard<std::mutex> g(mu);
        finish_locked(j);
    }
}

// ---- C ABI ------------------------------------------------------------------------------------------
extern "C" {

int cglbm_server_create(const cglbm_server_opts *opts, cglbm_server **out) {
    if (!out) return set_error(CGLBM_EINVAL, "null output");
    *out = nullptr;
    cglbm_server_opts o{};
    if (opts) o = *opts;
    if (o.max_members <= 0) o.max_members = 1024;
    if (o.bundle <= 0) o.bundle = 8;
    if (o.admit_threads <= 0) o.admit_threads = 2;
    if (o.bundle > 4096 || o.admit_threads > 16 || o.hbm_budget_bytes < 0)
        return set_error(CGLBM_EINVAL, "bundle <= 4096, admit_threads <= 16, budget >= 0");
    if (o.tile_ty != 0 && ((o.tile_ty != 4 && o.tile_ty != 8 && o.tile_ty != 16) || o.tile_L < 1 ||
                           (o.tile_nbuf != 1 && o.tile_nbuf != 2)))
        return set_error(CGLBM_EINVAL, "tile rows 4, 8 or 16, L >= 1, nbuf 1 or 2");
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return set_error(CGLBM_EHIP, "no HIP device");
    if (o.hbm_budget_bytes == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return set_error(CGLBM_EHIP, "hipMemGetInfo failed");
        o.hbm_budget_bytes = static_cast<int64_t>(0.8 * static_cast<double>(free_b));
    }
    cglbm_server *s = new (std::nothrow) cglbm_server;
    if (!s) return set_error(CGLBM_ENOMEM, "host allocation failed");
    s->opts = o;
    s->device = device;
    s->t0 = std::chrono::steady_clock::now();
    s->st.budget_bytes = o.hbm_budget_bytes;
    try {
        s->engine = std::thread(&cglbm_server::engine_loop, s);
        s->completer = std::thread(&cglbm_server::complete_loop, s);
        for (int i = 0; i < o.admit_threads; ++i) s->admitters.emplace_back(&cglbm_server::admit_loop, s);
    } catch (...) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv_admit.notify_all();
        for (auto &t : s->admitters) t.join();
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->admit_done = true;
            s->cv_engine.notify_all();
        }
        if (s->engine.joinable()) s->engine.join();
        if (s->completer.joinable()) s->completer.join();
        delete s;
        return set_error(CGLBM_ENOMEM, "cannot start server threads");
    }
    *out = s;
    return CGLBM_OK;
}

int cglbm_server_submit(cglbm_server *s, const cglbm_request *req, uint64_t *id) {
    if (!s || !req) return set_error(CGLBM_EINVAL, "null server or request");
    if (int rc = validate_config(&req->cfg)) return rc;
    const cglbm_config &c = req->cfg;
    if (req->storage != CGLBM_STORE_F32 && req->storage != CGLBM_STORE_F16)
        return set_error(CGLBM_EINVAL, "unknown storage mode");
    if (req->case_kind < CGLBM_CASE_BUBBLE || req->case_kind > CGLBM_CASE_STATE)
        return set_error(CGLBM_EINVAL, "unknown case kind");
    if (req->case_kind == CGLBM_CASE_JET && c.ybc != CGLBM_Y_JET)
        return set_error(CGLBM_EINVAL, "jet case needs ybc = CGLBM_Y_JET");
    if (req->case_kind == CGLBM_CASE_BUBBLE && (!(req->radius > 0) || (!req->sharp && !(req->width > 0))))
        return set_error(CGLBM_EINVAL, "bubble radius/width must be positive");
    if (req->case_kind == CGLBM_CASE_RTI &&
        (!(req->width > 0) || req->n_random_modes < 0 || req->n_random_modes > 64))
        return set_error(CGLBM_EINVAL, "bad RTI parameters");
    if (req->case_kind == CGLBM_CASE_STATE && !req->initial) return set_error(CGLBM_EINVAL, "state case needs `initial`");
    if (req->steps < 0) return set_error(CGLBM_EINVAL, "negative step count");
    if ((req->result_mask & ~0x7FFu) || (req->result_mask && !req->result))
        return set_error(CGLBM_EINVAL, "result_mask needs a result buffer (fields 0..10)");
    if (req->storage == CGLBM_STORE_F16) {
        if (c.gravity != 0.0f && c.ybc == CGLBM_Y_PERIODIC)
            return set_error(CGLBM_EINVAL, "FP16 storage: a body force needs bounded y (walls or jet planes)");
        if (c.nx < 3 || c.ny < 3 || c.nz < 3) return set_error(CGLBM_EINVAL, "FP16 storage needs 3 cells per axis");
    } else if (c.gravity != 0.0f && c.ybc == CGLBM_Y_PERIODIC) {
        return set_error(CGLBM_EINVAL, "a forced periodic case has no batched kernel");
    }
    Job *j = new (std::nothrow) Job;
    if (!j) return set_error(CGLBM_ENOMEM, "host allocation failed");
    j->req = *req;
    j->req.initial = nullptr;
    const int64_t fs = field_stride(c);
    j->bytes_life = cglbm_domain_bytes(&c, req->storage);
    if (req->storage == CGLBM_STORE_F16) {
        j->bytes_init = 44 * fs;
        if (req->result_mask & ~0x3u) j->bytes_life += 4 * fs;
    }
    j->remaining = req->steps;
    j->res.user_tag = req->user_tag;
    j->res.storage = req->storage;
    j->res.cells = static_cast<int64_t>(c.nx) * c.ny * c.nz;
    if (req->case_kind == CGLBM_CASE_STATE) {
        try {
            j->initial.assign(req->initial, req->initial + 11 * stored_cells(c));
        } catch (...) {
            delete j;
            return set_error(CGLBM_ENOMEM, "host allocation failed");
        }
    }
    std::lock_guard<std::mutex> g(s->mu);
    if (j->bytes_life + j->bytes_init > s->st.budget_bytes) {
        delete j;
        return set_error(CGLBM_ENOMEM, "case exceeds the server's HBM budget");
    }
    if (s->stop) {
        delete j;
        return set_error(CGLBM_ESTATE, "server is shutting down");
    }
    j->res.id = s->next_id++;
    j->res.t_submit = s->clock();
    if (id) *id = j->res.id;
    s->pending.push_back(j);
    ++s->st.submitted;
    ++s->st.queued;
    s->cv_admit.notify_all();
    return CGLBM_OK;
}

int cglbm_server_poll(cglbm_server *s, cglbm_result *out, int32_t max, int32_t timeout_ms, int32_t *n) {
    if (!s || !n || (max > 0 && !out)) return set_error(CGLBM_EINVAL, "bad arguments");
    *n = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    auto has = [&] { return !s->done.empty(); };
    if (timeout_ms < 0) s->cv_done.wait(lk, has);
    else if (timeout_ms > 0) s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), has);
    while (*n < max && !s->done.empty()) {
        out[(*n)++] = s->done.front();
        s->done.pop_front();
    }
    return CGLBM_OK;
}

int cglbm_server_drain(cglbm_server *s) {
    if (!s) return set_error(CGLBM_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->st.completed + s->st.failed == s->st.submitted; });
    return CGLBM_OK;
}

int cglbm_server_stats_get(cglbm_server *s, cglbm_server_stats *out) {
    if (!s || !out) return set_error(CGLBM_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    return CGLBM_OK;
}

double cglbm_server_clock(const cglbm_server *s) { return s ? s->clock() : -1.0; }

int cglbm_server_destroy(cglbm_server *s) {
    if (!s) return CGLBM_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        while (!s->pending.empty()) {  // never admitted: fail them
            Job *j = s->pending.front();
            s->pending.pop_front();
            --s->st.queued;
            ++s->st.resident;  // finish_locked counts it out again
            s->st.bytes_reserved += j->bytes_life;
            std::snprintf(j->res.message, sizeof(j->res.message), "server destroyed before admission");
            j->res.status = CGLBM_ESTATE;
            j->res.t_done = s->clock();
            s->finish_locked(j);
        }
    }
    s->cv_admit.notify_all();
    for (auto &t : s->admitters) t.join();  // each finishes the case it is initialising
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->admit_done = true;
        s->cv_engine.notify_all();
    }
    s->engine.join();
    s->completer.join();
    if (s->batch) (void)cglbm_batch_destroy(s->batch);
    delete s;
    return CGLBM_OK;
}

}  // extern "C"
