// Internal helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "cglbm.h"
#include "kernels.hpp"
#include "step_common.hpp"

namespace cglbm {

int set_error(int code, const char *msg);
int validate_config(const cglbm_config *c);
int check_layout(const cglbm_config *c, const cglbm_layout *L);
StepArgs make_step_args(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *c,
                        const cglbm_layout *L, int64_t step);
BcArgs make_bc_args(const cglbm_moments *B, const cglbm_moments *Aold, const cglbm_config *c,
                    const cglbm_layout *L, int64_t step);
// shape == nullptr: the tuned shape recorded for this configuration, else the default.
// store == kStoreF16: A->rhor / B->rhor are the bases of FP16-deviation states at field stride fs
// (storage.hpp); the other view pointers are not used.
// Balanced z-chunk plan of a library-owned domain (plan.hip). A marching shape with L == 0 asks
// for it: the step after every `period` steps (and the first) records the interface map and is
// followed by the planner; the steps in between take their chunks from `items`.
inline constexpr int kTrialMaxCand = 5;
inline constexpr int kTrialSteps = 3;
// A decision is kept kTrialFirstPeriod steps after a state is set, then four times as long after
// each race, up to kTrialPeriod: a young state (a jet that has just started, a drop still
// relaxing) changes fastest, and a race at its first steps can pick a shape that loses later.
inline constexpr int64_t kTrialFirstPeriod = 128;
inline constexpr int64_t kTrialPeriod = 2048;
struct MarchPlanState {
    int4 *items = nullptr;    // cap work items
    uint8_t *wmap = nullptr;  // interface map of the last recording step (plan_wmap_bytes)
    uint32_t *itime = nullptr;  // cap item durations of the last recording step
    int cap = 0;
    int ncu = 0;              // CUs of the device
    int ty = 0;               // tile rows the current plan was made for; 0 = none yet
    int period = 64;
    int since = 0;            // steps since the plan was made
    // Shape race of a domain left to the default shape: planned chunks (where they apply, once the
    // first plan and its measured correction are made) and the uniform candidates each run
    // kTrialSteps steps between two events on the domain's stream; the host then waits for the last
    // event and keeps the fastest shape for trial_period steps, after which the race runs again
    // (the interface moves), or until the state is replaced.
    // Results are bitwise the same for every shape. No events (caller plans): no race.
    hipEvent_t tev[2 * kTrialMaxCand] = {};
    int trial = 0;            // kTrial* below
    int trial_cand = 0;       // candidate being timed
    int trial_left = 0;       // steps left in its timed phase
    int64_t trial_since = 0;  // steps since the last decision
    int64_t trial_period = kTrialFirstPeriod;  // steps the current decision is kept
    MarchShape keep{0, 0, 0}; // shape kept by the last race (ty == 0: none yet)
};
enum : int {
    kTrialNone = 0,    // not raced yet (the first candidate's steps meanwhile)
    kTrialTiming = 1,  // timing candidate trial_cand
    kTrialKeep = 2,    // `keep` decided
};
// Work items of a planned step for tiles of `ty` rows: one per CU, rounded up to a multiple of
// the CU count so every column tile gets at least one.
int plan_items(const cglbm_config &c, int ty, int ncu);
// Work items to allocate (the largest plan_items over the tile-row choices) and bytes of the
// interface map for the largest tile-row choice; both together in device bytes.
int plan_capacity(const cglbm_config &c, int ncu);
size_t plan_wmap_bytes(const cglbm_config &c);
int64_t plan_device_bytes(const cglbm_config &c);

int launch_step(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                const cglbm_layout *L, int64_t step, int kernel, const MarchShape *shape, hipStream_t s,
                int store = 0, int64_t fs = 0, MarchPlanState *plan = nullptr);

// Process-wide record of autotuned z-marching shapes, keyed by (q, ybc, nx, ny, nz, storage,
// body force on/off).
bool tuned_shape_lookup(const cglbm_config &c, int store, MarchShape *out);
void tuned_shape_store(const cglbm_config &c, int store, MarchShape m);

inline int stored_ny(const cglbm_config &c) { return c.ybc == CGLBM_Y_PERIODIC ? c.ny : c.ny + 2; }
inline int64_t stored_cells(const cglbm_config &c) {
    return static_cast<int64_t>(c.nx) * stored_ny(c) * c.nz;
}
// Field stride in floats: padded to a multiple of 64 floats (256 B) so every field starts aligned.
inline int64_t field_stride(const cglbm_config &c) { return (stored_cells(c) + 63) / 64 * 64; }

}  // namespace cglbm
