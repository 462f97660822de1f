/* cglbm — MI355X (gfx950) colour-gradient lattice Boltzmann update behind a plain C ABI.
 *
 * The reference has no FFI: its solver update sits behind inline host launchers that take
 * non-owning views of eleven device arrays (launch.cuh:14-94, kernels.cuh:10-30), with all
 * configuration compile-time (constants.cuh, -DLBM_D3Q19). This header replaces that boundary
 * one to one, with the configuration moved to a runtime struct:
 *
 *   MomentsDevice (memory.cuh:14-29) ..................... cglbm_moments
 *   constants.cuh:9-93 + -DLBM_D3Q19 (compile.sh:9-26) ... cglbm_config
 *   launch_RCS (launch.cuh:89-94) ........................ cglbm_launch_rcs
 *   launch_JetBoundaryConditions (launch.cuh:78-83) ...... cglbm_launch_ybc
 *   launch_InitStaticBubble/InitRTI/InitJet (launch.cuh:14-72) ... cglbm_launch_init_* (caller
 *                                                         arrays), cglbm_init_* (domains)
 *   allocate_device_memory / free_device_memory (memory.cuh:153-197) ... cglbm_domain_create/destroy
 *   swap_moments (memory.cuh:200-203) .................... inside cglbm_step
 *   copy_out_D2H (memory.cuh:209-217) .................... cglbm_copy_out
 *   write_vti_step_device (io/vtiWriter.cuh:151-164) ..... cglbm_write_vti
 *   write_vti (io/vtiWriter.cuh:92-145) ................... cglbm_write_vti_host
 *   MlupsCounter (utilities/mlups_count.cuh:14-144) ...... cglbm_time_steps
 *
 * Every entry point returns 0 on success or a negative CGLBM_E* code (no exit(), unlike
 * CUDA_CHECK in utilities/cudaUtilities.cuh:9-19); cglbm_last_error() gives the message of the
 * calling thread's last failure. Streams are hipStream_t passed as void* (NULL = the null
 * stream). Step launches are asynchronous on their stream; a domain is used from one host
 * thread at a time.
 *
 * Geometry. nx, ny, nz count interior cells (the reference's NX-2, NY-2, NZ-2). x and z are
 * periodic (modular wrap; the reference's ghost columns, which are never read or written by the
 * update, are not stored). y is either periodic, or bounded by the reference's two ghost planes
 * (stored y = 0 and y = ny + 1) carrying the jet inlet/outlet (jet/boundary_conditions.cuh) or
 * the RTI wall copies (rti/rtiIn.cuh:77-119). Arithmetic is IEEE binary32 with the reference's
 * operation order, so results match the reference's update bit for bit.
 */
#ifndef CGLBM_H
#define CGLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGLBM_ABI_VERSION 1

enum {
    CGLBM_OK = 0,
    CGLBM_EINVAL = -1,   /* invalid argument or configuration */
    CGLBM_ENOMEM = -2,   /* device or host allocation failed / over the memory budget */
    CGLBM_EHIP = -3,     /* HIP runtime error */
    CGLBM_EIO = -4,      /* file input or output failed */
    CGLBM_ESTATE = -5,   /* call not valid in the object's current state */
};

enum { CGLBM_Y_PERIODIC = 0, CGLBM_Y_JET = 1, CGLBM_Y_WALLS = 2 };
enum { CGLBM_OUTLET_CONVECTIVE = 0, CGLBM_OUTLET_ZERO_GRADIENT = 1 };
enum { CGLBM_STORE_F32 = 0, CGLBM_STORE_F16 = 1 };
enum { CGLBM_KERNEL_AUTO = 0, CGLBM_KERNEL_GATHER = 1, CGLBM_KERNEL_MARCH = 2 };

/* Eleven device arrays, same names and order as MomentsDevice (memory.cuh:14-29). */
typedef struct cglbm_moments {
    float *rhor, *rhob;
    float *ux, *uy, *uz;
    float *Pixx, *Pixy, *Piyy, *Piyz, *Pizz, *Pixz;
} cglbm_moments;

/* Runtime form of constants.cuh:9-93 and the stencil macro. */
typedef struct cglbm_config {
    int32_t nx, ny, nz;       /* interior cells */
    int32_t q;                /* 19 or 27 (stencil_ct.cuh:419-425) */
    int32_t ybc;              /* CGLBM_Y_* */
    int32_t outlet;           /* CGLBM_OUTLET_* (jet only; convective is the reference's) */
    float tau;                /* taub: one relaxation time for both phases (constants.cuh:82) */
    float sigma;              /* surface tension (constants.cuh:90) */
    float beta;               /* recolouring beta (constants.cuh:91) */
    float jet_x0, jet_z0;     /* inlet disk centre, interior coordinates ((n-1)/2 = reference) */
    float jet_radius;         /* constants.cuh:32 */
    float jet_velocity;       /* constants.cuh:35 */
    float jet_noise;          /* relative inlet velocity perturbation; 0 = reference */
    uint32_t seed;            /* keys the inlet perturbation */
    float gravity;            /* body-force acceleration along -y; 0 = reference (no forcing) */
    float atwood;             /* density-weighted forcing F = atwood*(rhob - rhor)*gravity*(-y) */
    int32_t conservative_omega2; /* 0 = reference; 1 = skip Omega2 where |F| <= 1e-4 */
} cglbm_config;

/* Where cell (x, y, z) lives inside each array: origin + x + y*pitch_y + z*pitch_z, with y the
 * stored index (0 = inlet ghost plane when ybc != periodic). The compact layout is
 * {0, nx, nx*nyt}; the reference's ghost-column layout of an NX*NY*NZ array is
 * {1 + NX*NY (+ NX for a periodic y), NX, NX*NY}. */
typedef struct cglbm_layout {
    int64_t origin, pitch_y, pitch_z;
} cglbm_layout;

const char *cglbm_last_error(void);
int cglbm_abi_version(void);
/* Fills the compact layout of a configuration. */
int cglbm_compact_layout(const cglbm_config *cfg, cglbm_layout *out);

/* ---- Launch-level entry points: caller-owned arrays (launch.cuh replacement) ------------ */

/* One fused update of every interior cell, A -> B (RCSKernel, kernels.cu:119-131). When
 * ybc != periodic the ghost planes of B are written in the same launch sequence, so this one
 * call replaces main.cu:41 and main.cu:43. `step` keys the inlet perturbation. No field of B
 * may overlap a field of A or another field of B over the cells the layout addresses
 * (CGLBM_EINVAL). kernel = CGLBM_KERNEL_*. */
int cglbm_launch_rcs(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                     const cglbm_layout *layout, int64_t step, int32_t kernel, void *stream);

/* The z-marching kernel (the fast path) takes any placement of the 11 fields: at one stride from
 * one base (field f at A->rhor + f * stride, one allocation of 11 * stride floats) it addresses a
 * state below 4 GiB as one buffer; fields in separate allocations, as the reference's
 * allocate_device (memory.cuh:59-81) makes them, it addresses through one 64-bit pointer each, with
 * the same results at 95-100% of the rate. */

/* Balanced chunk plan for launch-level calls: what a library-owned domain keeps for itself
 * (cglbm_domain_set_tile with L = 0), held by the caller for one pair of caller-owned states.
 * No counterpart in the reference (its RCSKernel launch covers the domain uniformly). */
typedef struct cglbm_plan cglbm_plan;
int cglbm_plan_create(const cglbm_config *cfg, cglbm_plan **out);
int cglbm_plan_destroy(cglbm_plan *plan);
/* Forget the current cut (the caller replaced the state): the next launch records afresh. */
int cglbm_plan_reset(cglbm_plan *plan);
/* cglbm_launch_rcs on the z-marching kernel with 16-row tiles and planned chunks: every 64
 * launches (and the first two after creation or a reset) the step records which waves took the
 * interface form and how long each work-group took, and a one-work-group planner re-cuts the
 * z-chunks so every CU gets equal work. Same results as cglbm_launch_rcs, bit for bit. Call it
 * for consecutive steps of the same pair of states, in step order, on one stream; cfg must have
 * the plan's extents and stencil. Where planning does not apply (more column tiles than one round
 * of work-groups, or a planner limit) it runs as cglbm_launch_rcs. */
int cglbm_launch_rcs_planned(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                             const cglbm_layout *layout, int64_t step, cglbm_plan *plan, void *stream);

/* cglbm_launch_rcs_planned without a plan object and without any device allocation. No
 * counterpart in the reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The plan lives in a file-scope device
 * array of 384 KiB per GPU (held until the process ends), enough for a single-round plan of 16-row
 * column tiles while ceil(nx/32) ceil(ny/16) nz <= ~48,000 (cubes up to 256^3 on 256 CUs); larger domains
 * get CGLBM_EINVAL with the bytes needed. Domains whose plan would need several rounds of
 * work-groups run as cglbm_launch_rcs (no scratch). The plan follows the last call's extents and
 * stencil: a call with another shape starts a fresh plan (results are bit for bit those of
 * cglbm_launch_rcs either way). Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (two host threads, or two streams running at once) race; a process stepping
 * several domains on one GPU should keep a cglbm_plan per domain instead. */
int cglbm_launch_rcs_planned_noalloc(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                                     const cglbm_layout *layout, int64_t step, void *stream);

/* One shard of a domain split into z-slabs (no counterpart in the reference, which runs on one
 * GPU): updates target planes [z_lo, z_hi) of B from A, reading A's planes z_lo-2 .. z_hi+1. The
 * arrays hold a slab plus two halo planes on each side (cfg->nz counts all of them), which the
 * caller fills from the neighbouring shards before the call: plane z_lo-1 (all 11 fields) and
 * z_lo-2 (rhor, rhob: the two-hop colour gradient) below, z_hi and z_hi+1 above. Requires
 * 2 <= z_lo < z_hi <= nz - 2; no plane wraps. z_offset is added to a local plane index where the
 * update depends on the global one (the jet inlet's disk and perturbation key), so the shards of a
 * domain together give bitwise the domain's own step. x and y are treated as `cfg` says (boundary
 * planes are written for the updated planes). z-marching kernel, uniform chunks of the slab. */
int cglbm_launch_rcs_zslab(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                           const cglbm_layout *layout, int32_t z_lo, int32_t z_hi, int32_t z_offset, int64_t step,
                           void *stream);

/* The faces of a z-slab in one launch (no counterpart in the reference): target planes
 * [z_lo, z_lo + width) and [z_hi - width, z_hi), as cglbm_launch_rcs_zslab would update them
 * (1 <= width <= (z_hi - z_lo) / 2). With width 2 it completes a step whose interior,
 * [z_lo + 2, z_hi - 2), was updated by cglbm_launch_rcs_zslab before the halo planes arrived: the
 * interior reads no halo plane, so the exchange can run during it. */
int cglbm_launch_rcs_zslab_faces(const cglbm_moments *A, const cglbm_moments *B, const cglbm_config *cfg,
                                 const cglbm_layout *layout, int32_t z_lo, int32_t z_hi, int32_t width,
                                 int32_t z_offset, int64_t step, void *stream);

/* Device side of the z-slab halo exchange (no counterpart in the reference). Each face moves 13
 * plane-fields, nx * nyt floats each, in one contiguous buffer: the 11 fields of the plane next to
 * the face, then rhor and rhob of the plane beyond it. cglbm_zslab_pack copies, from A, what the
 * neighbours need: planes z_lo (fields) and z_lo+1 (rhor, rhob) into send_lo for the one below, z_hi-1
 * and z_hi-2 into send_hi for the one above (slabs of at least 2 planes). cglbm_zslab_unpack writes
 * a received recv_lo into the halo planes z_lo-1 and z_lo-2 and recv_hi into z_hi and z_hi+1. One
 * launch each on `stream`; the transfer between GPUs (RCCL send/receive, peer copies) is the
 * caller's. Same range rules as cglbm_launch_rcs_zslab. */
int cglbm_zslab_pack(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, int32_t z_lo,
                     int32_t z_hi, float *send_lo, float *send_hi, void *stream);
int cglbm_zslab_unpack(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, int32_t z_lo,
                       int32_t z_hi, const float *recv_lo, const float *recv_hi, void *stream);

/* Boundary planes only: JetBoundaryConditions(B, Aold) (kernels.cu:102-114) or the wall copies.
 * B == Aold is allowed (main.cu:26). */
int cglbm_launch_ybc(const cglbm_moments *B, const cglbm_moments *Aold, const cglbm_config *cfg,
                     const cglbm_layout *layout, int64_t step, void *stream);

/* Initial conditions on caller-owned arrays: the states of cglbm_init_bubble / cglbm_init_jet /
 * cglbm_init_rti written into the 11 arrays of A at `layout` (cells the layout does not address,
 * such as the reference's ghost columns, are left untouched). Computed on the host in double,
 * uploaded through a transient staging buffer and scattered on `stream`; synchronous, as the
 * reference's launchers synchronise the device.
 *   launch_InitStaticBubble (launch.cuh:14-20; bubble/bubbleIn.cuh:15-52) ... cglbm_launch_init_bubble
 *       (sharp = 1 is the reference's sphere; sharp = 0 the diffuse tanh drop)
 *   launch_InitJet (launch.cuh:52-72; jet/jetIn.cuh:16-60) + main.cu:26 ... cglbm_launch_init_jet
 *       (ybc must be CGLBM_Y_JET; the inlet disk comes from the boundary planes, as main.cu:26)
 *   launch_InitRTI (launch.cuh:42-46; rti/rtiIn.cuh:18-119) ........... cglbm_launch_init_rti
 *       (with ybc = CGLBM_Y_WALLS the wall planes are written as InitRTIWalls does) */
int cglbm_launch_init_bubble(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, double x0,
                             double y0, double z0, double radius, double width, int32_t sharp, void *stream);
int cglbm_launch_init_jet(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout, void *stream);
int cglbm_launch_init_rti(const cglbm_moments *A, const cglbm_config *cfg, const cglbm_layout *layout,
                          double amplitude, double width, int32_t n_random_modes, double random_amplitude,
                          uint32_t seed, void *stream);

/* ---- Domain objects: library-owned ping-pong state (LbmDevice replacement) -------------- */

typedef struct cglbm_domain cglbm_domain;

/* Allocates A and B (2 x 11 arrays, zeroed) plus a private stream. storage = CGLBM_STORE_*. */
int cglbm_domain_create(const cglbm_config *cfg, int32_t storage, cglbm_domain **out);
int cglbm_domain_destroy(cglbm_domain *dom);
/* Current state (which = 0) or the other buffer (which = 1); FP32 storage only. */
int cglbm_domain_moments(cglbm_domain *dom, int32_t which, cglbm_moments *out);
int cglbm_domain_config(const cglbm_domain *dom, cglbm_config *out);
/* Device bytes held by a domain of this configuration. */
int64_t cglbm_domain_bytes(const cglbm_config *cfg, int32_t storage);
int cglbm_domain_set_kernel(cglbm_domain *dom, int32_t kernel);
/* Tile shape of the z-marching kernel: ty tile rows (4, 8 or 16; tiles are 32 cells wide),
 * L planes per z-chunk, nbuf LDS exchange buffers (2: one barrier per plane; 1: half the LDS,
 * two barriers; D3Q27 with 16 rows always runs 1, which does not fit twice). L = 0 asks for planned chunks: every 64 steps the step records which waves took
 * the interface form and a planner cuts each column tile's z range into chunks of equal modelled
 * cost, one work-group per CU (results are bitwise the same for every shape). ty = 0 returns to
 * automatic selection: the autotuned record, else planned 16-row chunks (when they average at
 * least 2 planes and the planner's LDS holds a byte per (column tile, plane): cubes from 64^3
 * to 256^3), else the size-aware uniform shape; either races the other default candidates
 * (the size-aware uniform shape, and 4-row tiles with 32-plane chunks for D3Q19) once the state is
 * set and every 2048 steps after: three steps of each timed with events on the domain's stream,
 * the fastest kept (cglbm_step then waits for those few steps once per race). */
int cglbm_domain_set_tile(cglbm_domain *dom, int32_t ty, int32_t L, int32_t nbuf);
/* Autotuner: times each candidate tile shape (planned chunks, L = 0, among them) for
 * steps_per_candidate updates from the current state (the state does not advance), keeps the
 * fastest for this domain and records it for every later domain or launch of the same
 * (q, ybc, nx, ny, nz, storage, body force on/off); launches without a plan take the size-aware
 * uniform chunk length for a recorded L = 0. Outputs may be NULL. Synchronous. */
int cglbm_domain_autotune(cglbm_domain *dom, int32_t steps_per_candidate, int32_t *ty, int32_t *L,
                          int32_t *nbuf, double *ms_per_step);
/* The domain's current chunk plan (for tiles with L = 0): *n_items work items, each four int32
 * {tile x, tile y, z0, z1} (z0 == z1: an idle work-group), copied into items when cap >= *n_items;
 * *ty = the plan's tile rows. *n_items = 0 when no plan has been made since the state was last
 * set. Waits for the domain's stream. */
int cglbm_domain_plan(cglbm_domain *dom, int32_t *items, int32_t cap, int32_t *n_items, int32_t *ty);
int64_t cglbm_domain_step_count(const cglbm_domain *dom);
void *cglbm_domain_stream(cglbm_domain *dom);

/* Initial conditions (computed on the host in double, uploaded; synchronous). */
int cglbm_init_bubble(cglbm_domain *dom, double x0, double y0, double z0, double radius, double width,
                      int32_t sharp);
int cglbm_init_jet(cglbm_domain *dom);
int cglbm_init_rti(cglbm_domain *dom, double amplitude, double width, int32_t n_random_modes,
                   double random_amplitude, uint32_t seed);

/* nsteps of (update, boundary planes, swap) on the domain's stream (main.cu:38-55). */
int cglbm_step(cglbm_domain *dom, int32_t nsteps);
/* Same, timed with HIP events recorded on the domain's stream around the launches;
 * ms_total receives the elapsed time (MlupsCounter, mlups_count.cuh:41-144). */
int cglbm_time_steps(cglbm_domain *dom, int32_t nsteps, double *ms_total);
int cglbm_sync(cglbm_domain *dom);

/* Host <-> device copies of the current state, compact layout, fields in cglbm_moments order,
 * each nx*nyt*nz floats (nyt = ny, or ny + 2 with ghost planes). field_mask bit f selects field
 * f; the selected fields are packed consecutively in `host`. Synchronous. */
int cglbm_copy_out(cglbm_domain *dom, uint32_t field_mask, float *host);
int cglbm_copy_in(cglbm_domain *dom, uint32_t field_mask, const float *host);

/* Restartable snapshots (no counterpart in the reference, whose only output is the 5-field VTI).
 * cglbm_snapshot_write stores the domain's configuration, storage format, step count and its full
 * current state (11 fields, ghost planes included; FP16 states decoded exactly) in one file;
 * cglbm_snapshot_read restores them into a domain created with the same configuration and
 * storage, after which stepping gives bitwise the results of the uninterrupted run (the inlet
 * perturbation's step index included). cglbm_snapshot_info reads a snapshot's configuration,
 * storage and step count (any output pointer may be NULL), to create the domain to restore into.
 * Files carry a checksum and this library's ABI version; a mismatch is CGLBM_EINVAL, a file that
 * cannot be opened or written CGLBM_EIO. Synchronous. */
int cglbm_snapshot_write(cglbm_domain *dom, const char *path);
int cglbm_snapshot_read(cglbm_domain *dom, const char *path);
int cglbm_snapshot_info(const char *path, cglbm_config *cfg, int32_t *storage, int64_t *steps);

/* FP64 diagnostics of the current state over interior cells, fixed summation order:
 * out[0] = sum rhor, out[1] = sum rhob, out[2..4] = sum rho*u, out[5] = count of bad cells: a
 * negative or non-finite phase density, a non-positive density, or any non-finite velocity or Pi
 * component (with FP16 storage a saturated field decodes to inf and counts). Synchronous. */
int cglbm_reduce(cglbm_domain *dom, double out[6]);

/* ASCII VTI with rhor rhob ux uy uz of the current state over the stored extent
 * (write_vti_step_device, io/vtiWriter.cuh:151-164). Synchronous. */
int cglbm_write_vti(cglbm_domain *dom, const char *path);

/* The reference's writer itself (write_vti(filename, LbmHost), io/vtiWriter.cuh:92-145): five
 * host arrays of nx*ny*nz floats, x fastest, any extent (the reference passes its full grid,
 * ghost layers included). Creates missing parent folders. Byte-identical output for the same
 * arrays. Needs no GPU. CGLBM_EIO where the reference throws std::runtime_error. */
int cglbm_write_vti_host(const char *path, int nx, int ny, int nz, const float *rhor, const float *rhob,
                         const float *ux, const float *uy, const float *uz);

/* ---- Batches: many domains stepped together ----------------------------------------------
 * The reference runs one domain per process (main.cu). A batch steps any number of domains with
 * one launch per kernel class (stencil, bounded y, body force) per step over all their tiles, on
 * the batch's own stream. Each member's result is identical to stepping it alone. Members keep
 * every domain entry point (init, copies, diagnostics, VTI); those wait for the batch stream. A
 * member cannot be stepped on its own or destroyed while it belongs to a batch. */

typedef struct cglbm_batch cglbm_batch;

int cglbm_batch_create(cglbm_batch **out);
/* Releases the batch (members are detached, not destroyed). */
int cglbm_batch_destroy(cglbm_batch *b);
int cglbm_batch_add(cglbm_batch *b, cglbm_domain *dom);
int cglbm_batch_remove(cglbm_batch *b, cglbm_domain *dom);
int32_t cglbm_batch_size(const cglbm_batch *b);
/* One tile shape for every class (ty = 0: per-class autotuned or default). */
int cglbm_batch_set_tile(cglbm_batch *b, int32_t ty, int32_t L, int32_t nbuf);
/* Issue steps as replays of a captured HIP graph of `bundle` steps (0 = plain launches). */
int cglbm_batch_set_graph(cglbm_batch *b, int32_t bundle);
/* nsteps batched steps of every member (asynchronous on the batch stream). */
int cglbm_batch_step(cglbm_batch *b, int32_t nsteps);
/* Same, timed with HIP events on the batch stream. */
int cglbm_batch_time_steps(cglbm_batch *b, int32_t nsteps, double *ms_total);
/* Baseline: the members stepped one after another by their own single-domain launches on the
 * batch stream, timed (used to report the batching speed-up). */
int cglbm_batch_time_sequential(cglbm_batch *b, int32_t nsteps, double *ms_total);
/* Autotuner per kernel class (members do not advance). Synchronous. */
int cglbm_batch_autotune(cglbm_batch *b, int32_t steps_per_candidate);
int cglbm_batch_sync(cglbm_batch *b);

/* ---- Serving: continuous batching of independent cases ------------------------------------
 * The reference has no request loop (one case per process, main.cu). A server admits requests
 * (bubble, jet, RTI or a caller-given state; D3Q19/D3Q27; FP32 or FP16-deviation storage) in
 * arrival order whenever the HBM budget and a member slot allow, creates and initialises each
 * case on admission worker threads while the GPU steps the resident ones, keeps every state
 * resident on the device, steps all of them together as one batch (graph-replayed bundles), and
 * hands each case to a completion thread as soon as its step count is reached: results are
 * copied to the caller's buffer, FP64 diagnostics computed, the device memory released. A
 * case's result is identical to running it alone. Thread-safe: any thread may submit and poll. */

typedef struct cglbm_server cglbm_server;

enum { CGLBM_CASE_BUBBLE = 0, CGLBM_CASE_JET = 1, CGLBM_CASE_RTI = 2, CGLBM_CASE_STATE = 3 };

typedef struct {
    cglbm_config cfg;
    int32_t storage;         /* CGLBM_STORE_F32 or CGLBM_STORE_F16 */
    int32_t case_kind;       /* CGLBM_CASE_* */
    double radius, width;    /* bubble: drop radius and tanh width (centre = domain middle); RTI: width */
    int32_t sharp;           /* bubble: sharp sphere instead of the tanh profile */
    double amplitude;        /* RTI: single-mode amplitude */
    int32_t n_random_modes;  /* RTI: seeded random modes */
    double random_amplitude;
    uint32_t init_seed;
    const float *initial;    /* CGLBM_CASE_STATE: 11 fields, compact layout (copied at submit) */
    int64_t steps;           /* updates to run (>= 0) */
    uint32_t result_mask;    /* fields copied to `result` when done (bit f = field f; 0 = none) */
    float *result;           /* caller buffer for popcount(result_mask) fields, or NULL */
    uint64_t user_tag;
} cglbm_request;

typedef struct {
    uint64_t id;
    uint64_t user_tag;
    int32_t status;          /* CGLBM_OK or the CGLBM_E* code of the failure */
    int32_t storage;
    int64_t steps;           /* steps completed */
    int64_t cells;           /* interior cells */
    double diag[6];          /* cglbm_reduce of the final state */
    double t_submit, t_admit, t_start, t_done;  /* seconds since cglbm_server_create */
    char message[160];       /* error text when status != CGLBM_OK */
} cglbm_result;

typedef struct {
    int64_t hbm_budget_bytes;  /* device bytes for resident cases (0: 80% of free memory) */
    int32_t max_members;       /* resident-case slots (0: 1024) */
    int32_t bundle;            /* steps per graph-replayed bundle (0: 8) */
    int32_t admit_threads;     /* admission workers (0: 2) */
    int32_t tile_ty, tile_L, tile_nbuf;  /* batch tile shape (tile_ty = 0: per-class default) */
} cglbm_server_opts;

typedef struct {
    int64_t submitted, admitted, completed, failed;
    int64_t queued, resident;  /* waiting for admission / on the device now */
    int64_t bytes_reserved, peak_bytes_reserved, budget_bytes;
    int64_t peak_resident;
    int64_t updates;           /* sum of cells x steps of completed cases */
    int64_t bundles;           /* batched step calls issued */
    double busy_seconds;       /* time with at least one resident case being stepped */
} cglbm_server_stats;

int cglbm_server_create(const cglbm_server_opts *opts, cglbm_server **out);
/* Validates and enqueues (never blocks on the GPU). CGLBM_ENOMEM if the case alone exceeds the
 * budget. The request struct may be reused after the call; `result` must stay valid until the
 * case's cglbm_result has been polled. */
int cglbm_server_submit(cglbm_server *s, const cglbm_request *req, uint64_t *id);
/* Up to `max` finished cases into `out` (*n set), waiting up to timeout_ms for the first one
 * (0: no wait, < 0: wait indefinitely). */
int cglbm_server_poll(cglbm_server *s, cglbm_result *out, int32_t max, int32_t timeout_ms, int32_t *n);
/* Blocks until every submitted case has finished (results stay available to poll). */
int cglbm_server_drain(cglbm_server *s);
int cglbm_server_stats_get(cglbm_server *s, cglbm_server_stats *out);
/* Seconds since cglbm_server_create (the clock of cglbm_result's timestamps). */
double cglbm_server_clock(const cglbm_server *s);
/* Stops the threads after the current work (pending cases are failed with CGLBM_ESTATE). */
int cglbm_server_destroy(cglbm_server *s);

#ifdef __cplusplus
}
#endif
#endif
