// fslbm_step_noalloc: one time step of caller-owned device state that
// allocates nothing.  For callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops): the caller passes the
// input and output state buffers and the mask plane, all on the device;
// the one table the tiled kernel derives from the mask, the per-column
// wall weights, is built by a kernel into `scratch`, a file-scope device
// array of fixed size that each GPU the library runs on holds until the
// process ends.  Every call on a GPU rewrites that GPU's table, so calls
// that overlap on one GPU (on different streams) race.

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "fslbm_internal.h"

namespace fslbm {
namespace {

// Wall weights of FSLBM_NOALLOC_MAX_COLUMNS columns in the widest
// arithmetic type: three planes (sum of W_e, W_e cx, W_e cy over the wall
// neighbours of each column).
constexpr size_t SCRATCH_BYTES = 3 * (size_t)FSLBM_NOALLOC_MAX_COLUMNS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// Device address of `scratch` per device, looked up once per device with
// that device current (in host code the array's name is not its address).
constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
void *g_scratch[MAX_DEVICES] = {};

hipError_t scratch_of(int dev, void **ptr)
{
    std::lock_guard<std::mutex> guard(g_lock);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *ptr = g_scratch[dev];
    return hipSuccess;
}

// The wall weights of one column, as the host builds them for a context
// (wall_weights in fslbm_codec.h): the same double sums in the same order
// (products of the 1D weights by -1, 0, 1 are exact), so the table is
// bitwise the same.  Wall columns get zeros; a neighbour outside the box
// is a wall.
template <typename R>
__global__ void k_wall_weights(const uint8_t *__restrict__ m2, int nx, int ny, R *__restrict__ om)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxy = nx * ny;
    if (c >= nxy) return;
    const int x = c % nx, y = c / nx;
    double o0 = 0, ox = 0, oy = 0;
    if (m2[c] == 0) {
        const double w[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
        for (int cy = -1; cy <= 1; cy++)
            for (int cx = -1; cx <= 1; cx++) {
                const int xx = x + cx, yy = y + cy;
                const bool wall = xx < 0 || yy < 0 || xx >= nx || yy >= ny ||
                                  m2[xx + nx * yy] != 0;
                if (!wall) continue;
                const double W = w[cx + 1] * w[cy + 1];
                o0 += W;
                ox += W * cx;
                oy += W * cy;
            }
    }
    om[c] = (R)o0;
    om[nxy + c] = (R)ox;
    om[2 * nxy + c] = (R)oy;
}

int fail(int code, const char *msg)
{
    set_thread_error(msg);
    return code;
}

}  // namespace
}  // namespace fslbm

using namespace fslbm;

extern "C" int fslbm_step_noalloc(const fslbm_params *prm, const uint8_t *mask2d, int nx,
                                  int ny, int nz, int precision, const void *in, void *out,
                                  double gauge, void *stream)
{
    set_thread_error("");
    if (!prm || !in || !out) return fail(FSLBM_EINVAL, "NULL argument");
    if (in == out) return fail(FSLBM_EINVAL, "in and out must be different buffers");
    if (precision != FSLBM_F32 && precision != FSLBM_F64)
        return fail(FSLBM_EINVAL, "fslbm_step_noalloc takes FSLBM_F32 or FSLBM_F64");
    if (nx < 1 || ny < 1 || nz < 1) return fail(FSLBM_EINVAL, "box dimensions must be positive");
    const long long nxy = (long long)nx * ny, N = nxy * nz;
    if (N >= (1ll << 31))
        return fail(FSLBM_EINVAL, "domain too large: cells are indexed with 32 bits (N < 2^31)");
    if (mask2d && nxy > FSLBM_NOALLOC_MAX_COLUMNS)
        return fail(FSLBM_EINVAL,
                    ("walled box of " + std::to_string(nxy) + " columns exceeds the " +
                     std::to_string(FSLBM_NOALLOC_MAX_COLUMNS) +
                     "-column wall-weight table of fslbm_step_noalloc")
                        .c_str());
    if (!std::isfinite(prm->tblend) || !std::isfinite(prm->pin) || !std::isfinite(prm->nu) ||
        !std::isfinite(prm->blend) || !std::isfinite(gauge))
        return fail(FSLBM_EINVAL, "non-finite run parameter");
    if (prm->d1 < 1 || 2 * prm->d1 > nz ||
        (!mask2d && (2 * prm->d1 > nx || 2 * prm->d1 > ny)))
        return fail(FSLBM_EINVAL, "d1 must be >= 1 with 2*d1 inside the box");
    Geo g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.N = N;
    g.periodic = mask2d == nullptr;
    if (!tiled_supported(g, precision, prm->d1, true))
        return fail(FSLBM_EINVAL, "fslbm_step_noalloc takes d1 = 1 or 2, nz >= 16 and fields "
                                  "below 2^32 bytes (2^31 in a periodic box)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return fail(FSLBM_ENODEV, "no usable HIP device");
    const hipStream_t s = (hipStream_t)stream;
    void *omega = nullptr;
    if (mask2d) {
        hipError_t e = scratch_of(dev, &omega);
        if (e != hipSuccess)
            return fail(FSLBM_EHIP, (std::string("scratch address lookup failed: ") +
                                     hipGetErrorString(e)).c_str());
        const int blocks = (int)((nxy + 255) / 256);
        if (precision == FSLBM_F64)
            k_wall_weights<double><<<blocks, 256, 0, s>>>(mask2d, nx, ny, (double *)omega);
        else
            k_wall_weights<float><<<blocks, 256, 0, s>>>(mask2d, nx, ny, (float *)omega);
        e = hipGetLastError();
        if (e != hipSuccess)
            return fail(FSLBM_EHIP, (std::string("wall weight launch failed: ") +
                                     hipGetErrorString(e)).c_str());
    }
    StepArgs a;
    a.tblend = prm->tblend;
    a.pin = prm->pin;
    a.nu = prm->nu;
    a.blend = prm->blend;
    a.d1 = prm->d1;
    a.gauge = gauge;
    a.q = Quant{1.f, 1.f, 1.f, 1.f, 0.f, 0.f};
    // Every tile is stepped (a context skips tiles of wall cells, whose
    // state stays zero): every cell of `out` is written, walls with 0.
    TiledAux aux{mask2d, omega, nullptr, 0, 0, solo_rows(precision, g, prm->blend != 0.0), s};
    const hipError_t e = launch_step_tiled(g, precision, in, out, aux, a);
    if (e != hipSuccess)
        return fail(FSLBM_EHIP, (std::string("step launch failed: ") + hipGetErrorString(e)).c_str());
    return FSLBM_OK;
}
