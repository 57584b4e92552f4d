/*
 * CPU oracle for the fractional-step lattice Boltzmann (FSLBM) time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what GPU results are compared
 * with: the GPU library never links it, and only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 *
 * It restates the reference update of alparmat/FlowContrast
 * (PipeDemoStandalone/FSLBM.cu) as plain C with OpenMP:
 *   - predictor (D3Q27 equilibrium gather)          FSLBM.cu:130-173
 *   - CCS reverse-diffusion corrector + blend       FSLBM.cu:231-273
 *   - periodic/wall boundary + gauge (bcond)        FSLBM.cu:280-303
 *   - in-place shift storage (FSLBM.cu:275,307-312,622-626) restated as an
 *     out-of-place update; every read sees time level n (race-free order)
 *   - monitoring sums                               FSLBM.cu:587-591,630-666
 *   - mean field / fluctuation statistics           FSLBM.cu:683-764
 *
 * Two layouts:
 *   padded : the reference layout.  lx*ly*lz cells, x fastest, mask codes
 *            0 = fluid (updated), 1 = periodic padding, 125 = wall; d2
 *            padding planes at each z end (pipe.m:49-58).
 *   folded : no padding planes.  nx*ny*nz cells, z periodic with the
 *            pressure jump +pin / -pin applied when a gather wraps below
 *            z=0 / above z=nz-1.  Mask codes 0 = fluid, 125 = wall, or no
 *            mask at all with x and y periodic too (fully periodic box).
 *
 * Precision: *_f64 uses exact rationals (8/27, 2/27, 1/54, 1/6, b/3, b/12);
 * *_f32 uses the reference's float literals (FSLBM.cu:60-61,231,599).
 * The source-order arithmetic is compiled with -ffp-contract=off.
 *
 * State arrays are SoA: four planes p, ux, uy, uz of N cells each, passed
 * as one pointer to 4*N values.  Stored velocity has the reference's sign
 * (stored u = -physical u; FSLBM.cu:137-148, avgFields.m:83).
 */
#ifndef FSLBM_ORACLE_H
#define FSLBM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double tblend; /* blend of new and old velocity, FSLBM.cu:269 */
    double pin;    /* pressure jump across the z period, FSLBM.cu:294-301 */
    double nu;     /* viscosity; dtddy = nu - 1/6, FSLBM.cu:599 */
    double blend;  /* order 2 <-> 4 blend of the CCS stencil, FSLBM.cu:231 */
    int d1;        /* coarse corrective stencil length, FSLBM.cu:243-253 */
} orc_params;

/* Reference (padded) layout, one full step: FSLBMcalc on every IO==0 cell,
 * then bcond with gauge p2.  in/out: 4*lx*ly*lz values; in != out. */
void orc_step_padded_f64(int lx, int ly, int lz, int d2, const uint8_t *IO,
                         const orc_params *prm, double p2,
                         const double *in, double *out);
void orc_step_padded_f32(int lx, int ly, int lz, int d2, const uint8_t *IO,
                         const orc_params *prm, double p2,
                         const float *in, float *out);

/* Folded layout, one full step.  mask == NULL means a fully periodic box.
 * gauge is subtracted from the new pressure of every updated cell (the
 * bcond gauge p2).  in != out. */
void orc_step_folded_f64(int nx, int ny, int nz, const uint8_t *mask,
                         const orc_params *prm, double gauge,
                         const double *in, double *out);
void orc_step_folded_f32(int nx, int ny, int nz, const uint8_t *mask,
                         const orc_params *prm, double gauge,
                         const float *in, float *out);

/* Monitoring sums over the inlet plane z=d2 and the outlet plane z=lz-d2
 * of the padded layout, FSLBM.cu:635-647.  l1 is the fluid cell count of
 * one cross-section.  Sums in double (the reference sums floats). */
void orc_monitor_padded_f64(int lx, int ly, int lz, int d2, double l1,
                            const double *state, double out4[4]);

/* Statistics accumulator with the reference's semantics, FSLBM.cu:683-764.
 * acc holds 11 arrays of n doubles in the order
 *   pav uxav uyav uzav p2av ux2av uy2av uz2av uxyav uxzav uyzav.
 * sample holds 4 arrays of n doubles: p ux uy uz of the sampled plane.
 * delta: 0 resets everything first, 1 resets the second-order sums, 2 keeps
 * both (FSLBM.cu:683-690).  cnt/cnt2 are updated. */
void orc_stats_accumulate(int n, int delta, double *acc, const double *sample,
                          int *cnt, int *cnt2);

/* Wall-clock MLUPS of nsteps folded steps (ping-pong, OpenMP) for the CPU
 * baseline.  Returns elapsed seconds; *threads gets omp_get_max_threads(). */
double orc_time_folded_f32(int nx, int ny, int nz, const uint8_t *mask,
                           const orc_params *prm, int nsteps,
                           float *a, float *b, int *threads);
double orc_time_folded_f64(int nx, int ny, int nz, const uint8_t *mask,
                           const orc_params *prm, int nsteps,
                           double *a, double *b, int *threads);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
