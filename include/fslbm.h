/*
 * fslbm.h -- C ABI of the MI355X (gfx950) FSLBM time-step library
 * (flowcontrast_amd/libfslbm.so).
 *
 * The reference (alparmat/FlowContrast, PipeDemoStandalone/FSLBM.cu) has no
 * library interface: its hot path is the per-step launch sequence
 *     38 x FSLBMcalc -> 310 x shiftback -> bcond          FSLBM.cu:622-628
 * over state reached through module-scope texture/surface references
 * (FSLBM.cu:39-40,537-543), driven by main() and a file protocol
 * (pipe.txt, param.txt in; qp.txt and the averaged planes out).
 * This header is the drop-in boundary for that path: one opaque context
 * owns the device state; plain pointers and sizes cross the boundary.
 *
 * Layout handed across the boundary ("folded"): nx*ny*nz cells, x fastest,
 * then y, then z.  The reference's d2 periodic padding planes
 * (pipe.m:49-58) are not stored: z is periodic and a gather that wraps
 * below z=0 adds +pin to the neighbour pressure, one that wraps above
 * z=nz-1 adds -pin (the copies bcond makes, FSLBM.cu:294-301).
 * Mask codes: 0 = fluid (updated), 125 = wall (state held at 0).  A NULL
 * mask selects a fully periodic box (x and y wrap too; no walls).
 * Stored velocity keeps the reference sign (stored u = -physical u).
 *
 * Conventions: every int-returning call returns 0 on success or a
 * negative FSLBM_E* code; fslbm_last_error() holds the message.  The
 * caller owns all host buffers; the library keeps no host pointer.  One
 * context per host thread at a time; many contexts per device.  Calls on a
 * context are asynchronous on its own HIP stream unless they return host
 * data (get_state, monitor, reductions, stats reads), which synchronise.
 */
#ifndef FSLBM_H
#define FSLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSLBM_ABI_VERSION 1

/* error codes */
#define FSLBM_OK 0
#define FSLBM_EINVAL (-1)   /* bad argument / geometry / parameters */
#define FSLBM_ENOMEM (-2)   /* device or host allocation failed */
#define FSLBM_EHIP (-3)     /* HIP runtime error */
#define FSLBM_ENODEV (-4)   /* no usable gfx950 device */
#define FSLBM_ESTATE (-5)   /* call not valid in the context's state */
#define FSLBM_EBUSY (-6)    /* batch admission refused: HBM budget full */

/* storage / arithmetic modes */
#define FSLBM_F64 0   /* double storage, double arithmetic, exact rationals */
#define FSLBM_F32 1   /* float storage and arithmetic, reference literals */
#define FSLBM_F16S 2  /* IEEE half storage of shifted fields, float math */
#define FSLBM_BF16S 3 /* bfloat16 storage of shifted fields, float math */

/* kernel variants (for testing; AUTO picks the fastest valid one) */
#define FSLBM_KERNEL_AUTO 0
#define FSLBM_KERNEL_PLAIN 1 /* one thread per cell, direct 27-point gather */
#define FSLBM_KERNEL_TILED 2 /* z-marching LDS tile, separable predictor */

/* Run parameters: the six values of param.txt (pipe.m:67-69, read at
 * FSLBM.cu:499-510 and reloaded at :669-677). */
typedef struct {
    double tblend; /* velocity blend factor (timestep = dx * tblend) */
    double pin;    /* pressure jump across the z period */
    double nu;     /* lattice viscosity; corrector uses nu - 1/6 */
    double blend;  /* CCS order 2 <-> 4 blend b in [0,1] */
    int d1;        /* coarse corrective stencil length (>= 1) */
    int delta;     /* statistics reset switch 0/1/2, FSLBM.cu:683-690 */
} fslbm_params;

typedef struct fslbm_ctx *fslbm_t;

/* Library / device information. */
int fslbm_abi_version(void);
const char *fslbm_error_string(int code);

/* Create a context and its device state (two state copies, mask).
 * Replaces the allocation + texture binding of FSLBM.cu:519-543,593.
 * mask: nx*ny*nz codes (0 fluid / 125 wall), or NULL for a fully
 * periodic box.  For a masked domain no fluid cell may lie on the x or y
 * faces of the box (the reference's wall frame, pipe.m:56-58).
 * The initial state is all zero. */
int fslbm_create(const fslbm_params *prm, const uint8_t *mask, int nx, int ny,
                 int nz, int precision, fslbm_t *out);
void fslbm_destroy(fslbm_t ctx);
const char *fslbm_last_error(fslbm_t ctx);

/* Replace the run parameters (param.txt hot reload, FSLBM.cu:669-677). */
int fslbm_set_params(fslbm_t ctx, const fslbm_params *prm);
int fslbm_get_params(fslbm_t ctx, fslbm_params *prm);

/* Choose the kernel variant (FSLBM_KERNEL_*).  get returns the variant
 * the next step will run (PLAIN or TILED).  The tiled kernel takes d1 of
 * 1 or 2 and a fully periodic box or a mask that is the same in every z
 * plane; other cases run the plain kernel. */
int fslbm_set_kernel(fslbm_t ctx, int variant);
int fslbm_get_kernel(fslbm_t ctx);

/* Tiled kernel schedule: z planes per work item (0 = default heuristic).
 * get returns the value in effect. */
int fslbm_set_tile_chunk(fslbm_t ctx, int chunk);
int fslbm_get_tile_chunk(fslbm_t ctx);

/* Lattice cells one step computes (no reference counterpart; for
 * bandwidth accounting).  The tiled kernel skips tiles that hold only wall
 * cells (their state stays zero), so this is the cell count of the tiles
 * holding fluid; the plain kernel covers all nx*ny*nz cells.  Inside those
 * tiles the tiled kernel also neither reads nor writes the 16-cell row
 * pieces of a walled box whose cells are all walls.  Returns a negative
 * error code on failure. */
long long fslbm_step_cells(fslbm_t ctx);

/* Host-side plan of the tiled kernel's schedule for a domain (no reference
 * counterpart; no device needed, for capacity planning and tests).
 * blend: the corrector's order-4 weight the domain is created with.
 * mask2d: one z plane of nx*ny codes (0/125), or NULL for a periodic box.
 * Outputs, any of which may be NULL: the rows of the tiles (64 columns
 * wide; 12 for FP32, or 8 in a periodic box where the step-time model
 * rates 8-row tiles faster; for FP16S and BF16S 12 in a walled box created
 * with blend 0, else 8; 8 for a mask that varies along z; 4 for FP64, or 8
 * for an FP64 domain past the pipelined kernel's 32-bit buffer offsets,
 * which runs the two-barrier kernel), the number of tiles holding fluid
 * (work items run on these only), the cells one step computes (as
 * fslbm_step_cells) and the default chunk (z planes per work item).  A
 * context keeps the tiles it was created with when its blend changes. */
int fslbm_tile_plan(int nx, int ny, int nz, int precision, double blend,
                    const uint8_t *mask2d, int *tile_rows, int *ntiles, long long *cells,
                    int *chunk);

/* Autotuner of the tiled kernel's schedule: after 2*trial_steps untimed
 * launches, times trial_steps launches for each candidate chunk twice (the
 * step-time model's best picks and a few fixed lengths; HIP events on the
 * context's stream), keeps the fastest (the model's first pick unless
 * another is at least 2% faster) and reports it (*best_chunk, *best_ms per
 * step; either may be NULL).  The solver state is not advanced: trial launches read the
 * current state and write the spare buffer.  No-op for the plain kernel. */
int fslbm_autotune(fslbm_t ctx, int trial_steps, int *best_chunk,
                   double *best_ms);

/* Upload / download the whole state as four SoA arrays of nx*ny*nz values
 * (p, ux, uy, uz).  The uploaded state is taken as the state after a
 * bcond: wall cells are forced to 0 (FSLBM.cu:290).  Replaces the restart
 * load (initphys1, FSLBM.cu:386-419) and the backup dump (:839-856). */
int fslbm_set_state(fslbm_t ctx, const double *p, const double *ux,
                    const double *uy, const double *uz);
int fslbm_get_state(fslbm_t ctx, double *p, double *ux, double *uy,
                    double *uz);
int fslbm_set_state_f32(fslbm_t ctx, const float *p, const float *ux,
                        const float *uy, const float *uz);
int fslbm_get_state_f32(fslbm_t ctx, float *p, float *ux, float *uy,
                        float *uz);

/* Exact snapshot of a context's stepping state (no reference counterpart;
 * for checkpoints that resume bitwise in every precision mode): the state
 * as stored on the device, the scales of the 16-bit storage modes, the
 * stochastic-rounding step counter and the pending gauge.  A state set
 * from doubles is re-quantised and restarts that counter, so only a
 * snapshot resumes 16-bit storage exactly.  size: the snapshot's bytes;
 * save fills buf (at least that many bytes, after the queued steps end);
 * load takes a snapshot of a context with the same precision and cells. */
long long fslbm_snapshot_size(fslbm_t ctx);
int fslbm_snapshot_save(fslbm_t ctx, void *buf, long long bytes);
int fslbm_snapshot_load(fslbm_t ctx, const void *buf, long long bytes);

/* Advance nsteps time steps: predictor + CCS corrector + blend + periodic
 * and wall boundaries (FSLBM.cu:622-628).  A pending gauge (set by
 * fslbm_monitor or fslbm_set_gauge) is subtracted from the new pressure of
 * the first step only, as bcond's p2 (FSLBM.cu:628). Asynchronous. */
int fslbm_step(fslbm_t ctx, int nsteps);

/* Set the gauge p2 for the next step (FSLBM.cu:628). */
int fslbm_set_gauge(fslbm_t ctx, double p2);

/* Monitoring, FSLBM.cu:630-647: out[0..3] = q1, q2, p1, p2, the mean axial
 * velocity and pressure over the inlet plane (padded z=d2, folded z=0) and
 * the outlet plane (padded z=lz-d2, the periodic image of folded z=0 with
 * p - pin), each divided by the fluid cell count of that cross-section.
 * If set_gauge != 0, p2 becomes the gauge of the next step (the
 * reference always does this).  Sums are FP64 on the device. */
int fslbm_monitor(fslbm_t ctx, int set_gauge, double out[4]);

/* Whole-field FP64 reductions over non-wall cells:
 * out[0] = count, out[1] = sum p, out[2..4] = sum ux, uy, uz,
 * out[5] = sum |u|^2, out[6] = sum p^2, out[7] = number of non-finite
 * values (NaN/Inf) in the state. */
int fslbm_reduce(fslbm_t ctx, double out[8]);

/* Single value at (x,y,z) (the qp.txt probe, FSLBM.cu:653-665):
 * out[0..3] = p, ux, uy, uz. */
int fslbm_probe(fslbm_t ctx, int x, int y, int z, double out[4]);

/* Extract one plane in the reference's padded layout.
 * axis 0: YZ plane at x = idx, out has ny*(nz+2*d2) values, index
 *         y + ny*zp (getYZplane, FSLBM.cu:315-324);
 * axis 2: XY plane at padded z = idx, out has nx*ny values, index x + nx*y
 *         (getXYplane, FSLBM.cu:327-336).
 * field 0..3 = p, ux, uy, uz.  Padding planes are rebuilt from the
 * periodic images with +pin (inlet) / -pin (outlet), walls are 0. */
int fslbm_get_plane(fslbm_t ctx, int axis, int idx, int d2, int field,
                    double *out);

/* Reference plane statistics, FSLBM.cu:683-764, accumulated on the device
 * in FP64 over the YZ plane x = xplane in the padded layout with d2
 * padding planes (n = ny*(nz+2*d2) points).  Each accumulate call is one
 * sample: optional reset per delta, cnt += 1, cnt2 += 1, running sums,
 * |dev| and dev products against the running mean including the sample.
 * read: acc gets 11*n values (sum p, ux, uy, uz; sum |dev| p, ux, uy, uz;
 * sum dev_x dev_y, dev_x dev_z, dev_z dev_y) and the counts; divide by cnt
 * (first four) or cnt2 (the rest) as the reference does at write time.
 * load: restore accumulators (restart, FSLBM.cu:429-468). */
int fslbm_stats_config(fslbm_t ctx, int xplane, int d2);
int fslbm_stats_accumulate(fslbm_t ctx, int delta);
int fslbm_stats_read(fslbm_t ctx, double *acc, int *cnt, int *cnt2);
int fslbm_stats_load(fslbm_t ctx, const double *acc, int cnt, int cnt2);
int fslbm_stats_points(fslbm_t ctx);

/* True moments for turbulence statistics (not in the reference, which
 * keeps mean |dev|): per cross-section cell (x,y), FP64 sums over z and
 * over samples of p, ux, uy, uz, p^2, ux^2, uy^2, uz^2, ux*uy, ux*uz,
 * uy*uz.  read: out gets 11*nx*ny values, *samples the number of
 * accumulate calls (each call adds nz values per (x,y)). */
int fslbm_moments_reset(fslbm_t ctx);
int fslbm_moments_accumulate(fslbm_t ctx);
int fslbm_moments_read(fslbm_t ctx, double *out, long *samples);

/* Wait for all work queued on the context. */
int fslbm_sync(fslbm_t ctx);

/* Time nsteps steps with HIP events recorded on the context's stream
 * around the whole sequence: *ms gets the elapsed milliseconds. */
int fslbm_time_steps(fslbm_t ctx, int nsteps, double *ms);

/* Bytes of device memory the context holds. */
int fslbm_device_bytes(fslbm_t ctx, size_t *bytes);

/* Host-side input builders (no device work).
 *
 * pipe.m geometry (pipe.m:38-58): inner box of side `diameter`, circle
 * ((i-o)^2 + (j-o)^2 <= r^2 with 1-based i,j, o = r = diameter/2), a wall
 * frame of `frame` cells, length + 2*frame planes; codes 0 fluid,
 * 1 padding (first and last `frame` planes), 125 wall.  mask has
 * (diameter+2*frame)^2*(length+2*frame) bytes, x fastest.
 * Returns the number of padding cells (code 1). */
long fslbm_pipe_m_mask(int diameter, int length, int frame, uint8_t *mask);

/* Circular pipe for the folded layout: nx*ny*nz codes, fluid where
 * (x - cx)^2 + (y - cy)^2 <= R^2 with cx = (nx-1)/2, cy = (ny-1)/2
 * (0-based), wall elsewhere.  Returns the fluid cells per cross-section. */
long fslbm_circle_mask(int nx, int ny, int nz, double radius, uint8_t *mask);

/* Seeded initial states (SplitMix64 keyed by seed, cell and component;
 * identical on every host).
 * Poiseuille: p = pin * (nz - z) / nz (the initphys ramp, FSLBM.cu:381,
 * folded), stored uz = -2*ubar*(1 - r^2/R^2) on fluid cells, plus uniform
 * noise in [-noise, noise] * 2*ubar on each velocity component; walls 0.
 * Taylor-Green (fully periodic, 2*pi box): u = U0 sin x cos y cos z,
 * v = -U0 cos x sin y cos z, w = 0 (stored with the reference sign, i.e.
 * negated), p = rho0 + 3*(U0^2/16)(cos 2x + cos 2y)(cos 2z + 2). */
int fslbm_init_poiseuille(int nx, int ny, int nz, const uint8_t *mask,
                          double radius, double ubar, double pin, double noise,
                          uint64_t seed, double *p, double *ux, double *uy,
                          double *uz);
int fslbm_init_taylor_green(int nx, int ny, int nz, double u0, double rho0,
                            double *p, double *ux, double *uy, double *uz);

/* Fold a reference (padded) pipe.txt mask of lx*ly*lz codes with d2
 * padding planes per end into nx=lx, ny=ly, nz=lz-2*d2 folded codes
 * (0/125).  Fails if the padding planes are not periodic images of the
 * first/last fluid planes. */
int fslbm_fold_mask(int lx, int ly, int lz, int d2, const uint8_t *padded,
                    uint8_t *folded);

/* ------------------------------------------------------------------ */
/* Batched domains (no counterpart in the reference, which runs one     */
/* hard-coded pipe): many independent pipes / periodic boxes advanced   */
/* together, one launch per kernel class and step over a packed table  */
/* of (domain, tile, z range) work items, one workgroup per item.       */
/* Every work item runs the same                                        */
/* device code as the single-domain tiled kernel, so a domain's result  */
/* is bitwise the same alone or in any mix.  Domains are admitted       */
/* against an HBM budget and stay device-resident until removed.        */
/* Precision FSLBM_F32, FSLBM_F16S or FSLBM_BF16S; one d1 (1 or 2) per  */
/* batch; masks must be the same in every z plane.                      */
/* ------------------------------------------------------------------ */
typedef struct fslbm_batch *fslbm_batch_t;

int fslbm_batch_create(int precision, int d1, size_t hbm_budget,
                       fslbm_batch_t *out);
void fslbm_batch_destroy(fslbm_batch_t b);
const char *fslbm_batch_last_error(fslbm_batch_t b);

/* Device bytes a domain of this size takes in a batch (admission test). */
long long fslbm_batch_domain_bytes(int precision, const uint8_t *mask, int nx,
                                   int ny, int nz);

/* Admit a domain with its run parameters, mask (NULL: periodic box) and
 * initial state (four arrays of nx*ny*nz doubles).  Returns the slot id
 * (>= 0), FSLBM_EBUSY when the budget cannot hold it, or another error. */
int fslbm_batch_add(fslbm_batch_t b, const fslbm_params *prm,
                    const uint8_t *mask, int nx, int ny, int nz,
                    const double *p, const double *ux, const double *uy,
                    const double *uz);
int fslbm_batch_remove(fslbm_batch_t b, int slot);

/* Advance every admitted domain by nsteps (asynchronous). */
int fslbm_batch_step(fslbm_batch_t b, int nsteps);
int fslbm_batch_get_state(fslbm_batch_t b, int slot, double *p, double *ux,
                          double *uy, double *uz);
int fslbm_batch_query(fslbm_batch_t b, size_t *used, size_t *budget,
                      int *ndomains, long long *cells);

/* Replay the step sequence from captured hipGraphs (16 steps per graph
 * launch; the rest of a call, and calls of fewer steps, launch directly).
 * A call of 32 steps or more over state larger than the last-level cache
 * steps the domains in cohorts that fit it, each with its own tables and
 * graphs, built once per call. */
int fslbm_batch_use_graph(fslbm_batch_t b, int enable);

/* Schedule: the longest work item in z planes (0: default, 128).  Tile
 * columns are cut into near-equal items of one length, the step-time
 * model's pick up to this cap, and run longest first.  autotune picks the
 * fastest cap by timing trial launches that do not advance any domain. */
int fslbm_batch_set_chunk(fslbm_batch_t b, int chunk);
int fslbm_batch_autotune(fslbm_batch_t b, int trial_steps, int *best_chunk,
                         double *best_ms);
int fslbm_batch_time_steps(fslbm_batch_t b, int nsteps, double *ms);
int fslbm_batch_sync(fslbm_batch_t b);

/* ------------------------------------------------------------------ */
/* Serving loop (no counterpart in the reference): solve requests of    */
/* different sizes, run parameters and step budgets arrive at any time. */
/* Each pump advances every stepping request by up to `quantum` steps   */
/* in one batched launch sequence per precision mode (the modes overlap */
/* on separate streams).  Meanwhile it copies out the results of the    */
/* requests whose budget ran out in the previous pump, and uploads the  */
/* requests that the server's encoder thread has encoded (in arrival    */
/* order, while the HBM budget holds them); the uploaded requests join  */
/* the stepping ones at the next pump (continuous batching).  A request */
/* counts against the budget from the start of its encoding and holds   */
/* device memory from upload to hand-back.  Its result is bitwise the   */
/* same as a solo run of the tiled single-domain path in its precision  */
/* mode.  The calls on one server come from one host thread at a time   */
/* (the encoder thread is internal).                                    */
/* ------------------------------------------------------------------ */
typedef struct fslbm_server *fslbm_server_t;

typedef struct fslbm_request {
    fslbm_params params; /* d1 must equal the server's                    */
    int nx, ny, nz;      /* box; the mask (if any) equal in every z plane */
    int precision;       /* FSLBM_F32, FSLBM_F16S or FSLBM_BF16S          */
    long long nsteps;    /* step budget of the solve (>= 1)               */
} fslbm_request;

typedef struct fslbm_completion {
    long long id;        /* as returned by fslbm_server_submit            */
    int status;          /* FSLBM_OK, or the error the request failed with */
    long long steps;     /* steps run                                     */
    double t_submit, t_admit, t_done; /* seconds on the server's clock    */
} fslbm_completion;

/* quantum: steps per pump at most (>= 2); hbm_budget: device bytes the
 * resident requests may hold together. */
int fslbm_server_create(int d1, size_t hbm_budget, int quantum,
                        fslbm_server_t *out);

/* Which resident requests step in a pump.  FSLBM_SERVE_SHARE (the
 * default): all of them, for the highest throughput.
 * FSLBM_SERVE_SHORTEST: those with the least work left (cells x steps
 * left) first, until their cells reach fill_cells (0: 32768 per compute
 * unit); the others wait resident (at most 32 pumps in a row), for lower
 * time to solution at some throughput.  Results do not depend on the
 * policy. */
#define FSLBM_SERVE_SHARE 0
#define FSLBM_SERVE_SHORTEST 1
int fslbm_server_set_policy(fslbm_server_t s, int policy, long long fill_cells);
void fslbm_server_destroy(fslbm_server_t s);
const char *fslbm_server_last_error(fslbm_server_t s);
/* Seconds since the server was created (the clock of the completions). */
double fslbm_server_clock(fslbm_server_t s);

/* Queue a request; the initial state (four arrays of nx*ny*nz doubles)
 * and mask (NULL: periodic box) are copied.  Returns the request id
 * (>= 0) or an error (a request larger than the whole budget is
 * FSLBM_EINVAL). */
long long fslbm_server_submit(fslbm_server_t s, const fslbm_request *req,
                              const uint8_t *mask, const double *p,
                              const double *ux, const double *uy,
                              const double *uz);

/* One scheduling round (blocks until its steps are done).  Returns the
 * number of requests that finished in it (>= 0) or an error. */
int fslbm_server_pump(fslbm_server_t s);

/* Pop the oldest unreported completion: 1 if one was written, 0 if none. */
int fslbm_server_next_done(fslbm_server_t s, fslbm_completion *out);

/* A finished request's final state (decoded to doubles); the result is
 * kept until fslbm_server_drop. */
int fslbm_server_fetch(fslbm_server_t s, long long id, double *p, double *ux,
                       double *uy, double *uz);
int fslbm_server_drop(fslbm_server_t s, long long id);

int fslbm_server_query(fslbm_server_t s, int *queued, int *running,
                       int *finished, size_t *used, size_t *budget);

/* ------------------------------------------------------------------ */
/* One step of caller-owned device state, allocating nothing.  No        */
/* counterpart in the reference; for callers that cannot allocate       */
/* device memory (code under HIP graph capture, tight serving loops).   */
/* It is fslbm_step without a context: in and out are device buffers of */
/* four SoA planes of nx*ny*nz values (p, ux, uy, uz; double for        */
/* FSLBM_F64, float for FSLBM_F32), in != out; mask2d is one z plane of */
/* nx*ny codes 0/125 on the device (the mask of every plane, with a     */
/* wall frame on the x/y faces), or NULL for a fully periodic box.  The */
/* wall cells of `in` must hold 0 in every field (the state after a     */
/* bcond, FSLBM.cu:290); every cell of `out` is written (walls with 0). */
/* gauge is subtracted from the new pressure (bcond's p2, FSLBM.cu:628).*/
/* The step is queued on `stream` (a hipStream_t; NULL: the default    */
/* stream) and the call returns without waiting.  d1 must be 1 or 2 and */
/* nz >= 16.  The per-column wall weights the kernel needs are built on */
/* the device into a fixed table of the library (FSLBM_NOALLOC_MAX_     */
/* COLUMNS columns; a walled box with more columns is FSLBM_EINVAL),    */
/* one per GPU, which every call on that GPU rewrites: calls that       */
/* overlap on one GPU race.  Errors: fslbm_last_error(NULL).            */
/* ------------------------------------------------------------------ */
#define FSLBM_NOALLOC_MAX_COLUMNS (1 << 19)
int fslbm_step_noalloc(const fslbm_params *prm, const uint8_t *mask2d, int nx, int ny,
                       int nz, int precision, const void *in, void *out, double gauge,
                       void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FSLBM_H */
