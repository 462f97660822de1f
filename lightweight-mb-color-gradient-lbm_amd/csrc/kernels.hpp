// Host-side launchers of the device kernels (internal C++ interface of the library).
#pragma once

#include <hip/hip_runtime.h>

#include "step_common.hpp"

namespace cglbm {

// Direct-gather step (any layout).
hipError_t launch_gather(const StepArgs &a, int q, hipStream_t s);
// Boundary planes of B from B and Aold (jet inlet/outlet or wall copies); no-op when periodic.
hipError_t launch_ybc(const BcArgs &a, hipStream_t s);
// Halo planes of a z-slab (zslab.hip): pack = copy the planes the two neighbours need from a.A into
// lo / hi (13 * nx * nyt floats each), else write lo / hi into the slab's halo planes.
hipError_t launch_zslab_halo(const StepArgs &a, int z_lo, int z_hi, float *lo, float *hi, bool pack, hipStream_t s);

// z-marching step (any layout, fused boundary planes).
struct MarchShape {
    int ty;    // tile rows (4, 8 or 16; the tile is 32 wide)
    int L;     // planes per z-chunk
    int nbuf;  // LDS exchange buffers: 2 = one barrier per plane, 1 = two barriers, half the LDS
};
// The shape a launch runs: D3Q27 16-row tiles keep one exchange buffer (two exceed the 160 KB LDS
// of a CU), whatever was asked.
inline MarchShape effective_shape(int q, MarchShape m) {
    if (q == 27 && m.ty == 16) m.nbuf = 1;
    return m;
}
// Arguments of the marching kernel: both states as one base pointer each plus a field stride
// (every field at base + f * stride), read and written through buffer resources; or, in the
// pointer form, one base pointer per field.
struct MarchArgs {
    const float *A0;
    float *B0;
    uint32_t fsA, fsB;      // field strides (elements)
    uint32_t spanA, spanB;  // bytes addressable from A0 / B0 (buffer range)
    int nx, ny, nz;
    uint32_t py, pz;
    int ybc, outlet, conservative, forced;
    Coeffs c;
    float jet_x0, jet_z0, jet_r, jet_u, jet_noise;
    uint32_t seed;
    int64_t step;
    int store;  // kStoreF32, kStoreF16 or kStoreF32Ptr (storage.hpp); spans and strides refer to that layout
    // Field bases of both states, field f of A at Af[f] and of B at Bf[f]: what the pointer form
    // (kStoreF32Ptr) addresses; the buffer forms use A0/B0 and the strides.
    const float *Af[11];
    float *Bf[11];
    // Target planes of uniform chunks: [zlo, zhi), the whole [0, nz) for a domain. A z-slab of a
    // domain sharded along z updates its own planes between two halo planes on each side (planes
    // zlo-2 .. zhi+1 are read, none wraps); zoff turns a local plane index into the global one
    // where the update depends on it (the jet inlet's disk and perturbation key). zface > 0
    // restricts the update to the zface planes next to each end of [zlo, zhi) (the faces of a slab,
    // updated once their halos have arrived): the first half of the chunks covers the lower face,
    // the second half the upper one.
    int zlo, zhi, zoff, zface;
    // Optional work plan (single-domain launches): items[i] = {tile x, tile y, z0, z1}, one per
    // work-group (z0 == z1: idle). Null: uniform chunks of the launch shape's length.
    const int4 *items;
    int nitems;
    // Optional interface map written by the step: wmap[((tile y * ntx + tile x) * nz + z) * (TY / 2) + w]
    // = 1 when own wave w of that column tile took the interface form at source plane z, else 0.
    uint8_t *wmap;
    // Optional, with items and wmap: itime[i] = duration of work item i in wall-clock ticks
    // (100 MHz), written by the recording step for the planner's measured correction.
    uint32_t *itime;
};
// Batched launch: descriptors of member domains (one kernel class: q, ybc, forced), a list of
// work items {domain, tile x, tile y, chunk}, and the device step base.
struct BatchDesc {
    MarchArgs args;      // geometry and physics of the member (A0/B0/step set per launch)
    float *buf0;         // the member's two states
    float *buf1;
    int cur0;            // index of the current state when the descriptors were built
    int nchunk;          // z-chunks of this member
    int64_t step0;       // member steps completed when the descriptors were built
};
struct BatchLaunch {
    const BatchDesc *descs;
    const int4 *items;
    int nitems;
    const int64_t *base;
    int k;
    int grid;
    int ybc, forced, store;
    MarchShape shape;
};
hipError_t launch_batch_q19_f32(const BatchLaunch &L, hipStream_t s);
hipError_t launch_batch_q19_f16(const BatchLaunch &L, hipStream_t s);
hipError_t launch_batch_q27_f32(const BatchLaunch &L, hipStream_t s);
hipError_t launch_batch_q27_f16(const BatchLaunch &L, hipStream_t s);
hipError_t launch_batch(const BatchLaunch &L, int q, hipStream_t s);
// device base += n (one thread), ordered on the stream after a bundle of n steps
hipError_t launch_advance(int64_t *base, int n, hipStream_t s);

// Fills `m` (FP32 storage): the buffer form when the 11 fields of each state sit at one stride
// from one base and the state is below 4 GiB (the domain layout, or caller arrays carved from one
// allocation), else the pointer form (fields in separate allocations, larger states). False when
// a field holds 2^32 cells or more.
bool march_args(const StepArgs &a, MarchArgs &m);
// Same geometry and physics for an FP16-deviation state pair at field stride fs (storage.hpp).
bool march_args_f16(const StepArgs &a, const float *A0, float *B0, int64_t fs, MarchArgs &m);
bool march_supports(const StepArgs &a, int q);
MarchShape march_default_shape(const StepArgs &a, int q);
hipError_t launch_march_shape(const StepArgs &a, int q, MarchShape m, hipStream_t s);
hipError_t launch_march_args(const MarchArgs &a, int q, MarchShape m, hipStream_t s);
hipError_t launch_march_q19_f32(const MarchArgs &a, MarchShape m, hipStream_t s);
hipError_t launch_march_q19_f16(const MarchArgs &a, MarchShape m, hipStream_t s);
hipError_t launch_march_q27_f32(const MarchArgs &a, MarchShape m, hipStream_t s);
hipError_t launch_march_q27_f16(const MarchArgs &a, MarchShape m, hipStream_t s);
hipError_t launch_march_q19_f32ptr(const MarchArgs &a, MarchShape m, hipStream_t s);
hipError_t launch_march_q27_f32ptr(const MarchArgs &a, MarchShape m, hipStream_t s);
hipError_t launch_march(const StepArgs &a, int q, hipStream_t s);

// Balanced chunk plan of one domain (plan.hip): from the interface map of a recorded step
// (ntx * nty column tiles, nz planes, nw own waves per tile), writes nitems work items
// {tile x, tile y, z0, z1} covering every (column tile, plane) once; unused items are idle.
// dynamic LDS of the planner: 20 KB of item staging and sort keys, 4 B per column and per
// thread, and one byte per (column tile, plane)
constexpr int64_t kPlanMaxShared = 60 * 1024;
bool plan_fits(int ntx, int nty, int nz, int nitems);
// itime (may be null): durations of the items currently in `items`, timed by the recording step;
// the planner then scales its modelled plane costs by each item's measured over modelled cost.
hipError_t launch_plan(const uint8_t *wmap, const uint32_t *itime, int ntx, int nty, int nz, int nw, float alpha,
                       float e, int4 *items, int nitems, hipStream_t s);

// FP16-deviation states (storage.hpp, field stride fs) <-> FP32 compact fields [f0, f1), field f of
// the FP32 side at (f - f0) * xs floats; n cells.
hipError_t launch_pack_f16(const float *src, int64_t xs, float *state, int64_t fs, int64_t n, int f0, int f1,
                           hipStream_t s);
hipError_t launch_unpack_f16(const float *state, int64_t fs, float *dst, int64_t xs, int64_t n, int f0, int f1,
                             hipStream_t s);
// Compact FP32 fields (nx * nyt * nz cells each, back to back from src) into 11 caller arrays at
// cell x + y py + z pz (pointers with the layout origin applied).
struct FieldPtrs {
    float *p[11];
};
hipError_t launch_scatter_fields(const float *src, const FieldPtrs &dst, int nx, int nyt, int nz, uint32_t py,
                                 uint32_t pz, hipStream_t s);

// FP64 diagnostics: partial sums per block into `partials` (capacity kReduceBlocks * 6), then the
// fixed-order final sum into `result` (6 doubles).
inline constexpr int kReduceBlocks = 512;
hipError_t launch_reduce(const float *const *A, int nx, int ny, int nz, int y0, uint32_t py, uint32_t pz,
                         double *partials, double *result, hipStream_t s);
// Same over an FP16-deviation state (storage.hpp) at field stride fs.
hipError_t launch_reduce_f16(const float *state, int64_t fs, int nx, int ny, int nz, int y0, uint32_t py, uint32_t pz,
                             double *partials, double *result, hipStream_t s);

}  // namespace cglbm
