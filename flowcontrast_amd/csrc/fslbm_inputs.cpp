// Host-side input builders of the C ABI: geometry masks and seeded
// initial states.  No device work; safe to call without a GPU.

#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/fslbm.h"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): a counter-based generator, so
// every (seed, cell, component) draws the same number on every host.
inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in [-1, 1) from (seed, cell, component); 53-bit resolution.
inline double unit_noise(uint64_t seed, uint64_t cell, int comp)
{
    const uint64_t r = splitmix64(splitmix64(seed) ^ (cell * 4ull + (uint64_t)comp));
    return 2.0 * ((double)(r >> 11) * 0x1.0p-53) - 1.0;
}

}  // namespace

extern "C" {

// pipe.m:38-58.  1-based (i, j) in the inner box, o = r = diameter/2.
long fslbm_pipe_m_mask(int diameter, int length, int frame, uint8_t *mask)
{
    if (diameter < 1 || length < 1 || frame < 0 || !mask) return FSLBM_EINVAL;
    const int lx = diameter + 2 * frame, ly = lx, lz = length + 2 * frame;
    const double o = diameter / 2.0, r = o;
    long padding = 0;
    for (int z = 0; z < lz; z++) {
        const bool pad = z < frame || z >= lz - frame;  // c = [1:d l-d+1:l]
        for (int y = 0; y < ly; y++)
            for (int x = 0; x < lx; x++) {
                const long idx = x + (long)lx * (y + (long)ly * z);
                const int i = x - frame + 1, j = y - frame + 1;
                uint8_t code = 125;  // IO2 = ones(...)*125
                if (i >= 1 && i <= diameter && j >= 1 && j <= diameter) {
                    const double di = i - o, dj = j - o;
                    const bool inside = di * di + dj * dj <= r * r;
                    code = inside ? (pad ? 1 : 0) : 125;
                }
                if (code == 1) padding++;
                mask[idx] = code;
            }
    }
    return padding;
}

long fslbm_circle_mask(int nx, int ny, int nz, double radius, uint8_t *mask)
{
    if (nx < 1 || ny < 1 || nz < 1 || !mask || !(radius > 0)) return FSLBM_EINVAL;
    const double cx = (nx - 1) / 2.0, cy = (ny - 1) / 2.0;
    long fluid = 0;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const double dx = x - cx, dy = y - cy;
            const bool in = dx * dx + dy * dy <= radius * radius;
            fluid += in;
            for (int z = 0; z < nz; z++)
                mask[x + (long)nx * (y + (long)ny * z)] = in ? 0 : 125;
        }
    return fluid;
}

int fslbm_init_poiseuille(int nx, int ny, int nz, const uint8_t *mask,
                          double radius, double ubar, double pin, double noise,
                          uint64_t seed, double *p, double *ux, double *uy,
                          double *uz)
{
    if (nx < 1 || ny < 1 || nz < 1 || !mask || !p || !ux || !uy || !uz ||
        !(radius > 0))
        return FSLBM_EINVAL;
    const double cx = (nx - 1) / 2.0, cy = (ny - 1) / 2.0;
    const double umax = 2.0 * ubar;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long i = x + (long)nx * (y + (long)ny * z);
                if (mask[i] != 0) {
                    p[i] = ux[i] = uy[i] = uz[i] = 0.0;
                    continue;
                }
                // initphys ramp (FSLBM.cu:381) on the folded fluid planes
                p[i] = pin * (double)(nz - z) / (double)nz;
                const double dx = x - cx, dy = y - cy;
                double prof = 1.0 - (dx * dx + dy * dy) / (radius * radius);
                if (prof < 0) prof = 0;
                const double amp = noise * umax;
                ux[i] = amp * unit_noise(seed, (uint64_t)i, 1);
                uy[i] = amp * unit_noise(seed, (uint64_t)i, 2);
                // stored velocity = -physical velocity (flow towards +z)
                uz[i] = -umax * prof + amp * unit_noise(seed, (uint64_t)i, 3);
            }
    return FSLBM_OK;
}

int fslbm_init_taylor_green(int nx, int ny, int nz, double u0, double rho0,
                            double *p, double *ux, double *uy, double *uz)
{
    if (nx < 1 || ny < 1 || nz < 1 || !p || !ux || !uy || !uz) return FSLBM_EINVAL;
    const double two_pi = 6.283185307179586476925286766559;
    for (int z = 0; z < nz; z++) {
        const double zc = two_pi * z / nz;
        for (int y = 0; y < ny; y++) {
            const double yc = two_pi * y / ny;
            for (int x = 0; x < nx; x++) {
                const double xc = two_pi * x / nx;
                const long i = x + (long)nx * (y + (long)ny * z);
                // stored = -physical
                ux[i] = -u0 * std::sin(xc) * std::cos(yc) * std::cos(zc);
                uy[i] = u0 * std::cos(xc) * std::sin(yc) * std::cos(zc);
                uz[i] = 0.0;
                // P = p/3 (PDF eq. 5): p = rho0 + 3 (U0^2/16)(...)(...)
                p[i] = rho0 + 3.0 * (u0 * u0 / 16.0) *
                                  (std::cos(2 * xc) + std::cos(2 * yc)) *
                                  (std::cos(2 * zc) + 2.0);
            }
        }
    }
    return FSLBM_OK;
}

int fslbm_fold_mask(int lx, int ly, int lz, int d2, const uint8_t *padded,
                    uint8_t *folded)
{
    if (lx < 1 || ly < 1 || d2 < 0 || lz <= 2 * d2 || !padded || !folded)
        return FSLBM_EINVAL;
    const int L = lz - 2 * d2;
    const long lxy = (long)lx * ly;
    for (int z = 0; z < lz; z++)
        for (long c = 0; c < lxy; c++) {
            const uint8_t m = padded[c + lxy * z];
            if (m != 0 && m != 1 && m != 125) return FSLBM_EINVAL;
            if (z >= d2 && z < lz - d2) {
                if (m == 1) return FSLBM_EINVAL;  // padding code in the fluid range
                folded[c + lxy * (z - d2)] = m;
            }
        }
    // Padding planes must be the periodic images (1 where the image is
    // fluid, 125 where it is wall).
    for (int z = 0; z < lz; z++) {
        if (z >= d2 && z < lz - d2) continue;
        const int zf = z < d2 ? z - d2 + L : z - d2 - L;
        if (zf < 0 || zf >= L) return FSLBM_EINVAL;
        for (long c = 0; c < lxy; c++) {
            const uint8_t m = padded[c + lxy * z];
            const uint8_t f = folded[c + lxy * zf];
            if ((f == 0 && m != 1) || (f == 125 && m != 125)) return FSLBM_EINVAL;
        }
    }
    return FSLBM_OK;
}

}  // extern "C"
