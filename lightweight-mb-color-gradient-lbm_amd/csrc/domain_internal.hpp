// Internal state of a domain, shared by the domain and batch translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "cglbm.h"
#include "kernels.hpp"
#include "state.hpp"

struct cglbm_batch;

struct cglbm_domain {
    cglbm_config cfg{};
    int storage = CGLBM_STORE_F32;
    int kernel = CGLBM_KERNEL_AUTO;
    int device = 0;
    int64_t fstride = 0;
    float *buf[2] = {nullptr, nullptr};
    cglbm_moments view[2]{};
    int cur = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *d_partials = nullptr;
    double *d_result = nullptr;
    int64_t steps = 0;
    cglbm_layout layout{};
    cglbm::MarchShape shape{0, 0, 0};  // ty == 0: tuned record or default; L == 0: planned chunks
    cglbm::MarchPlanState plan;         // balanced chunk plan (state.hpp)
    cglbm_batch *batch = nullptr;       // owning batch, if any
    hipStream_t batch_stream = nullptr; // its stream
};

// Tells the owning batch that a member's step count or current buffer changed outside it.
void batch_member_changed(cglbm_batch *b);
hipStream_t batch_stream(cglbm_batch *b);
