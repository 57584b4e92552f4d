// Internal declarations shared by the host API and the HIP kernels of
// libfslbm.so.  Not part of the public C ABI (include/fslbm.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fslbm.h"

namespace fslbm {

// Geometry of one folded domain.
struct Geo {
    int nx, ny, nz;
    long long N;    // nx*ny*nz
    int periodic;   // 1: fully periodic box, no mask
};

// Dequantisation of the 16-bit shifted storage modes:
//   p  = pbase + pslope*z + sp*h_p,
//   ux = h_ux,  uy = h_uy,  uz = h_uz + uref[x + nx*y].
// sp is a power of two (exact scaling); su is 1 (velocities are stored
// unscaled) and kept so that snapshots carry the whole record.  uref is
// the streamwise reference profile, one float per column of the box in
// device memory (choose_uref): the stored uz is the flow's departure from
// it, and a rounding error relative to that departure is correspondingly
// smaller.  uref is nullptr in FP64 and FP32 (and reads as 0); it is 0 on
// every column that holds a wall cell, so walls still decode to 0.
struct Quant {
    float sp, inv_sp, su, inv_su, pbase, pslope;
    const float *uref = nullptr;  // device address (never dereferenced on the host)
};

// Per-launch scalars of one step (FSLBM.cu:231,269,599,623-624), in the
// arithmetic type R of the mode.
template <typename R>
struct StepK {
    R w0, w1, w2;     // weights of the cx=0 plane: centre, face, edge
    R quarter;        // factor for cx = +-1
    R dtddy, m, c, cc;
    R c0, c1, c2;
    R pin, gauge;
    int d1;
    uint32_t skey;    // stochastic-rounding key of this step (16-bit modes)
};

// Host-side description of a step launch.
struct StepArgs {
    double tblend, pin, nu, blend, gauge;
    int d1;
    Quant q;
    uint32_t seed = 0;  // step counter of the domain (stochastic rounding)
};

// Build the per-mode constants: exact rationals for FSLBM_F64, the
// reference's float literals for the float-arithmetic modes.
StepK<double> make_consts_f64(const StepArgs &a);
StepK<float> make_consts_f32(const StepArgs &a);

size_t storage_bytes(int precision);  // bytes per stored value
// Bytes of one state buffer (four fields) of a domain.
inline size_t state_bytes(const Geo &g, int precision)
{
    return storage_bytes(precision) * 4 * (size_t)g.N;
}
// Last-level (Infinity) cache of the GPU: 256 MB on MI355X.
constexpr size_t LLC_BYTES = size_t(256) << 20;
// Whether FP32 steps store the new state non-temporal: only when the
// state one step reads and writes (`bytes`, both buffers) exceeds the
// last-level cache, which otherwise keeps it for the next step.
inline bool stream_stores(size_t bytes) { return bytes > LLC_BYTES; }

// Kernel launchers (fslbm_kernels.hip, fslbm_tiled.hip).  `in`/`out` are
// the base of four SoA planes of N stored values each.
hipError_t launch_step_plain(const Geo &g, int precision, const void *in,
                             void *out, const uint8_t *mask,
                             const StepArgs &a, hipStream_t s);
// Extra inputs of the tiled kernel: the z-invariant mask as one plane,
// the per-column wall weights omega (3 planes of nx*ny values in the
// arithmetic type: sum over wall neighbours of W_e, W_e cx, W_e cy) and
// the device list of tiles that hold a fluid cell (from active_tiles;
// null: every tile), and the rows of the tiles (plan_rows).
struct TiledAux {
    const uint8_t *mask2d;
    const void *omega;
    const int *tiles;
    int ntiles;
    int chunk;  // z planes per work item (0: default)
    int rows;
    hipStream_t stream;
    // per-cell wall codes of a mask that varies along z (launch_wall_codes;
    // nullptr: z-invariant mask or periodic box, mask2d and omega instead)
    const uint16_t *wcode = nullptr;
    // both state buffers hold zero at every wall cell (a context keeps that
    // from the upload on): the kernel may skip the loads and stores of
    // rows of wall cells
    bool walls_zero = false;
};
// Wall codes (2 bytes per cell) of a walled box's mask, for the tiled
// kernel's variant for masks that vary along z.
hipError_t launch_wall_codes(const Geo &g, const uint8_t *mask, uint16_t *codes,
                             hipStream_t s);
hipError_t launch_step_tiled(const Geo &g, int precision, const void *in,
                             void *out, const TiledAux &aux,
                             const StepArgs &a);
bool tiled_supported(const Geo &g, int precision, int d1, bool z_invariant);
// Compute units of the current device.
int device_cus();
// Rows of the tiled kernels' tiles (64 columns wide): 8 for FP64 and BF16
// storage, 12 for FP32, and for FP16 storage 12 in a walled box whose
// corrector has no order-4 arms (blend 0), else 8; 8 in every 32-bit mode
// for a walled box whose mask varies along z (zvar).  A domain keeps the
// rows it was planned with; a later blend change runs the other corrector
// variant on the same tiles.
int plan_rows(int precision, bool periodic, bool order4, bool zvar = false);
// Rows of a single domain's tiles: plan_rows, except that an FP32 periodic
// box (which steps every tile) takes 8 rows where the step-time model
// rates their best schedule faster than that of 12 (batches keep 12: one
// launch per kernel class takes one tile shape).
int solo_rows(int precision, const Geo &g, bool order4, bool zvar = false);
// Tile indices (ty * ntx + tx, tiles of 64 x rows) whose cells include a
// fluid cell of the z-invariant mask plane; empty for a periodic box
// (every tile).
std::vector<int> active_tiles(const Geo &g, int rows, const uint8_t *mask2d);
// Lattice cells (all z planes) covered by those tiles, clipped to the box.
long long tiles_cells(const Geo &g, int rows, const std::vector<int> &tiles);
// Chunk lengths (z planes per work item) ordered by the modelled step
// time over `ntiles` tiles (0: every tile); default_chunk is the first.
std::vector<int> chunk_candidates(const Geo &g, int rows, int ntiles, int count);
int default_chunk(const Geo &g, int rows, int ntiles);

// Flag on a batch work item's ze: the domain's lead item, whose workgroup
// advances the domain's step counter.
constexpr int BATCH_LEAD = 0x40000000;
// Batch work items are dealt to the XCDs in groups of this many.
constexpr int BATCH_GROUP = 16;

// Batched step over many domains (32-bit arithmetic modes).  `doms` is a
// device array of per-domain descriptors (batch_dom_bytes each, filled by
// batch_fill_dom), `work` the packed (domain, work item) table of one
// kernel class (periodic / d1 / order-4 arms); flip = steps since the
// descriptors' parities were written, mod 2.
struct BatchLaunch {
    const void *doms;
    uint32_t *seeds;  // two step counters per domain (see k_step_batch)
    const int4 *work;  // (domain, tile, zs, ze | BATCH_LEAD)
    int nwork, flip, grid;
    bool periodic, fd2;
    bool nt;  // FP32 non-temporal stores (stream_stores over the batch)
    int d1, precision;
    hipStream_t stream;
};
hipError_t launch_step_batch(const BatchLaunch &b);
size_t batch_dom_bytes(int precision);
// Returns the domain's number of work items; `split` (if not null)
// receives the item layout: tiles per chunk, and chunk c covering
// split[1] + (c < split[2]) planes.
int batch_fill_dom(void *dst, int precision, const Geo &g, const StepArgs &a, int chunk,
                   void *buf0, void *buf1, const uint8_t *mask2d, const void *omega,
                   const int *tiles, int ntiles, int parity, int split[3]);
void batch_set_parity(void *dst, int precision, int parity);
struct HostResult;      // fslbm_codec.h
struct PreparedDomain;  // fslbm_codec.h
// Validate and encode a domain on the host (no device work), then upload
// and admit it; batch_add_plane does both.
int batch_prepare(fslbm_batch *b, const fslbm_params *prm, const uint8_t *mask, int nx,
                  int ny, int nz, const double *const src[4], PreparedDomain &pd);
// The two halves of batch_prepare, which touch no batch state (any thread
// may call them): the check (FSLBM_OK, or an error with the reason in why)
// and the encoding of a checked domain.
int domain_check(int precision, int d1, const fslbm_params *prm, const uint8_t *mask, int nx,
                 int ny, int nz, bool have_state, std::string &why);
void domain_encode(int precision, const fslbm_params *prm, const uint8_t *mask, int nx, int ny,
                   int nz, const double *const src[4], PreparedDomain &pd);
// attach uploads on stream s (nullptr: the batch's) and waits for it if
// sync; otherwise the caller keeps pd's host buffers until it has.
int batch_attach(fslbm_batch *b, PreparedDomain &pd, hipStream_t s, bool sync);
// Stop stepping a domain, keeping it resident; then copy its state out on
// a stream of the caller's (asynchronous).
int batch_hold(fslbm_batch *b, int slot);
int batch_resume(fslbm_batch *b, int slot);
int batch_copy_out(fslbm_batch *b, int slot, HostResult &out, hipStream_t s);
// Admit a domain whose mask (NULL: periodic box) is one z plane of nx*ny
// codes; src holds the four initial fields.  Returns the slot or an error.
int batch_add_plane(fslbm_batch *b, const fslbm_params *prm, const uint8_t *mask, int nx,
                    int ny, int nz, const double *const src[4]);
// Copy a batch domain's current state to the host, then remove it.
int batch_take(fslbm_batch *b, int slot, HostResult &out);

// Whole-field reductions (8 doubles, see fslbm_reduce) into dst (device).
hipError_t launch_reduce(const Geo &g, int precision, const void *state,
                         const uint8_t *mask, const Quant &q, double *partials,
                         int npart, double *dst, hipStream_t s);
// Plane-0 monitoring sums: dst[0] = count, dst[1] = sum p, dst[2] = sum uz.
hipError_t launch_monitor(const Geo &g, int precision, const void *state,
                          const uint8_t *mask, const Quant &q,
                          double *partials, int npart, double *dst,
                          hipStream_t s);
// Padded-layout plane extraction (fslbm_get_plane) into dst (device).
hipError_t launch_plane(const Geo &g, int precision, const void *state,
                        const uint8_t *mask, const Quant &q, double pin,
                        int axis, int idx, int d2, int field, double *dst,
                        hipStream_t s);
// Reference statistics sample (FSLBM.cu:715-748) on the padded YZ plane.
hipError_t launch_stats(const Geo &g, int precision, const void *state,
                        const uint8_t *mask, const Quant &q, double pin,
                        int xplane, int d2, double *acc, int cnt,
                        hipStream_t s);
// Moment sums per cross-section cell.
hipError_t launch_moments(const Geo &g, int precision, const void *state,
                          const uint8_t *mask, const Quant &q, double *acc,
                          hipStream_t s);
// Message of the last failed call without a handle (fslbm_create,
// fslbm_step_noalloc), read by fslbm_last_error(NULL); per host thread.
void set_thread_error(const char *msg);

// Fill n doubles with zero (kernel, usable under graph capture).
hipError_t launch_zero(double *dst, long long n, hipStream_t s);

}  // namespace fslbm
