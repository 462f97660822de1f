/* TEST INFRASTRUCTURE ONLY (oracle). Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load this library, and only as the checker. The product never links it.
 *
 * Clean-room CPU restatement of the reference colour-gradient update (lbm.cuh:20-262 and the
 * helpers it calls), templated on the arithmetic type (float for parity, double for the FP64
 * precision oracle). State layout is the product's compact layout: 11 SoA arrays
 * (rhor rhob ux uy uz Pixx Pixy Piyy Piyz Pizz Pixz), x fastest, then y, then z, with no ghost
 * columns in x and z (periodic by modular wrap) and, when the y boundary is not periodic, one
 * ghost plane at y = 0 and one at y = ny + 1 (the reference's inlet/outlet planes). */
#ifndef CGREF_H
#define CGREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CGREF_Y_PERIODIC = 0, CGREF_Y_JET = 1, CGREF_Y_WALLS = 2 };
enum { CGREF_OUTLET_CONVECTIVE = 0, CGREF_OUTLET_ZERO_GRADIENT = 1 };

typedef struct cgref_params {
    int32_t nx, ny, nz;     /* interior cells (ny excludes the two ghost planes) */
    int32_t q;              /* 19 or 27 */
    int32_t ybc;            /* CGREF_Y_* */
    int32_t outlet;         /* CGREF_OUTLET_* (jet only) */
    float tau, sigma, beta; /* taub, sigma, beta_recolor (constants.cuh:78-93) */
    float jet_x0, jet_z0;   /* inlet disk centre, compact coordinates */
    float jet_radius, jet_velocity;
    float jet_noise;        /* relative amplitude of the inlet velocity perturbation (0 = reference) */
    uint32_t seed;
    float gravity;          /* body-force acceleration along -y (0 = reference: no forcing) */
    float atwood;           /* density-weighted forcing: F = atwood*(rhob - rhor)*g */
    int32_t conservative_omega2; /* 0 = reference; 1 = skip Omega2 when |F| <= 1e-4 */
} cgref_params;

/* One step: RCS over interior cells A -> B, then the y-boundary planes of B (jet inlet/outlet from
 * B and the old A, or wall copies), exactly the order of main.cu:38-55. `step` keys the inlet noise.
 * Returns 0, or a negative code for invalid parameters. */
int cgref_step_f32(const cgref_params *p, float *const *A, float *const *B, int64_t step);
int cgref_step_f64(const cgref_params *p, double *const *A, double *const *B, int64_t step);

/* Boundary planes only (used once after initialisation, main.cu:26, with B == Aold). */
int cgref_apply_ybc_f32(const cgref_params *p, float *const *B, float *const *Aold, int64_t step);
int cgref_apply_ybc_f64(const cgref_params *p, double *const *B, double *const *Aold, int64_t step);

/* Initial conditions (compact layout, all cells written). */
int cgref_init_bubble_f32(const cgref_params *p, float *const *A, double x0, double y0, double z0,
                          double radius, double width, int32_t sharp);
int cgref_init_jet_f32(const cgref_params *p, float *const *A);
int cgref_init_rti_f32(const cgref_params *p, float *const *A, double amplitude, double width,
                       int32_t n_random_modes, double random_amplitude, uint32_t seed);

/* Per-phase mass and momentum over interior cells, accumulated in double in a fixed order:
 * out[0] = sum rhor, out[1] = sum rhob, out[2..4] = sum (rhor+rhob) u. */
int cgref_reduce_f32(const cgref_params *p, float *const *A, double *out);

/* Float constants exactly as the reference derives them (constants.cuh:56-93), for tests. */
int cgref_constants_f32(float tau, float sigma, float *out /* 12 */);

#ifdef __cplusplus
}
#endif
#endif
