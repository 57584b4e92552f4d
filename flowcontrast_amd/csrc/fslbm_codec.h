// Host-side conversion between caller arrays (double / float) and the
// stored state of a precision mode, shared by single contexts and batch
// domains.  Walls are stored as 0 and read back as 0.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

#include "fslbm_internal.h"
#include "fslbm_store.h"

namespace fslbm {

// Wall lookup over a full mask (nz planes) or a z-invariant plane.
struct HostMask {
    const uint8_t *m = nullptr;  // nullptr: periodic box, no walls
    long long nxy = 0;
    bool plane = false;          // m holds one plane, repeated in z
    bool wall(long long i) const
    {
        if (!m) return false;
        return (plane ? m[i % nxy] : m[i]) != 0;
    }
};

// Host threads for the codec loops: OMP_NUM_THREADS if set, else the
// hardware's count, at most 16.
inline int codec_threads()
{
    static const int n = [] {
        int t = 0;
        if (const char *e = std::getenv("OMP_NUM_THREADS")) t = std::atoi(e);
        if (t <= 0) t = (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(t, 16));
    }();
    return n;
}

// Run f(lo, hi, block) over fixed blocks of `block` cells of [0, n) on
// the codec threads.  The blocks do not depend on the thread count, so
// per-block partial results combined in block order are reproducible.
template <typename F>
void par_blocks(long long n, long long block, F f)
{
    const long long nb = (n + block - 1) / block;
    const int T = (int)std::min<long long>(codec_threads(), nb);
    auto run = [&](int t) {
        for (long long b = t; b < nb; b += T) f(b * block, std::min(n, (b + 1) * block), b);
    };
    if (T <= 1) {
        run(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(run, t);
    run(0);
    for (auto &x : th) x.join();
}

constexpr long long CODEC_BLOCK = 1 << 18;

inline float pow2_at_least(double v)
{
    int e;
    const double m = std::frexp(v, &e);  // v = m * 2^e, m in [0.5, 1)
    if (m == 0.5) e -= 1;
    return std::ldexp(1.0f, e);
}

// Shifted-storage reference and pressure scale from the state: the
// pressure reference is the mean plus the mean axial gradient -pin/nz
// (walled domains); the scale is a power of two, so scaling is exact.
template <typename T>
void choose_quant(const Geo &g, const HostMask &hm, double pin, const T *const src[4],
                  Quant &q)
{
    const long long N = g.N;
    const int nxy = g.nx * g.ny;
    const double slope = g.periodic ? 0.0 : -pin / g.nz;
    const long long nb = (N + CODEC_BLOCK - 1) / CODEC_BLOCK;
    std::vector<double> bsum(nb, 0.0);
    std::vector<long long> bcnt(nb, 0);
    par_blocks(N, CODEC_BLOCK, [&](long long lo, long long hi, long long b) {
        double sum = 0;
        long long n = 0;
        for (long long i = lo; i < hi; i++) {
            if (hm.wall(i)) continue;
            sum += (double)src[0][i] - slope * (double)(i / nxy);
            n++;
        }
        bsum[b] = sum;
        bcnt[b] = n;
    });
    double sum = 0;
    long long n = 0;
    for (long long b = 0; b < nb; b++) {
        sum += bsum[b];
        n += bcnt[b];
    }
    const double base = n ? sum / (double)n : 0.0;
    std::vector<double> bp(nb, 0.0);
    par_blocks(N, CODEC_BLOCK, [&](long long lo, long long hi, long long b) {
        double pmax = 0;
        for (long long i = lo; i < hi; i++) {
            if (hm.wall(i)) continue;
            const double dp = (double)src[0][i] - (base + slope * (double)(i / nxy));
            pmax = std::max(pmax, std::fabs(dp));
        }
        bp[b] = pmax;
    });
    double pmax = 0;
    for (long long b = 0; b < nb; b++) pmax = std::max(pmax, bp[b]);
    // Velocities are stored unscaled (su = 1, see Store<FSLBM_F16S>).  The
    // pressure deviation is scaled by a power of two sp: the largest
    // initial deviation maps into [512, 1024), 64x below the binary16
    // maximum, and sp >= 2^-15 keeps deviations up to 2 representable
    // whatever the initial state (a quiet start has none): past any
    // pressure a stable lattice flow reaches, |dp| <~ sqrt(3)|u| with |u|
    // below the lattice speed of sound.  Deviations above 2^-14 sp keep 11
    // significant bits, so the floor costs no precision that matters.
    const float sp = std::max(pow2_at_least(pmax) / 1024.0f, std::ldexp(1.0f, -15));
    q.su = 1.0f;
    q.inv_su = 1.0f;
    q.sp = sp;
    q.inv_sp = 1.0f / sp;
    q.pbase = (float)base;
    q.pslope = (float)slope;
}

// Streamwise reference profile of the 16-bit modes, one float per column
// (x, y): the z-mean of uz on a column that is fluid in every plane, 0 on
// a column holding a wall cell (so that walls, stored 0, decode to 0).
// Storing uz - uref leaves the departures from the mean profile to the
// 16-bit values: in a pipe the mean axial flow is most of |uz|, and the
// rounding of each store (an ulp relative to the stored value) shrinks
// with it.  Velocity rounding noise drives pressure noise through the
// predictor's e.u terms; an emulation of FP16 storage on a laminar
// 34x34x96 pipe (1500 steps) measured the pressure deviation's relative
// L2 error against FP64 at 11% without the reference and 2.1% with it.
template <typename T>
std::vector<float> choose_uref(const Geo &g, const HostMask &hm, const T *const src[4])
{
    const long long nxy = (long long)g.nx * g.ny;
    std::vector<float> ref((size_t)nxy, 0.0f);
    par_blocks(nxy, 4096, [&](long long lo, long long hi, long long) {
        for (long long c = lo; c < hi; c++) {
            double sum = 0;
            bool wall = false;
            for (long long z = 0; z < g.nz && !wall; z++) {
                const long long i = c + nxy * z;
                wall = hm.wall(i);
                sum += (double)src[3][i];
            }
            ref[(size_t)c] = wall ? 0.0f : (float)(sum / g.nz);
        }
    });
    return ref;
}

// Stored state from caller arrays; uref: the host copy of the streamwise
// reference profile of a 16-bit mode (nullptr: none).
template <typename T>
void encode_state(const Geo &g, const HostMask &hm, int precision, const Quant &q,
                  const T *const src[4], void *dst, const float *uref = nullptr)
{
    const long long N = g.N;
    const int nxy = g.nx * g.ny;
    auto run = [&](auto tag) {
        using ST = decltype(tag);
        using S = typename ST::S;
        using R = typename ST::R;
        S *d = (S *)dst;
        par_blocks(N, CODEC_BLOCK, [&](long long lo, long long hi, long long) {
            for (long long i = lo; i < hi; i++) {
                const bool wall = hm.wall(i);
                const int z = (int)(i / nxy);
                using L = StateLayout<ST::mode>;
                d[L::at(0, i, N)] = wall ? S(0) : ST::st_p((R)src[0][i], z, q);
                for (int f = 1; f < 3; f++)
                    d[L::at(f, i, N)] = wall ? S(0) : ST::st_u((R)src[f][i], q);
                const R w = uref ? (R)((double)src[3][i] - (double)uref[i % nxy]) : (R)src[3][i];
                d[L::at(3, i, N)] = wall ? S(0) : ST::st_u(w, q);
            }
        });
    };
    switch (precision) {
    case FSLBM_F64: run(Store<FSLBM_F64>{}); break;
    case FSLBM_F32: run(Store<FSLBM_F32>{}); break;
    case FSLBM_F16S: run(Store<FSLBM_F16S>{}); break;
    case FSLBM_BF16S: run(Store<FSLBM_BF16S>{}); break;
    }
}

template <typename T>
void decode_state(const Geo &g, const HostMask &hm, int precision, const Quant &q,
                  const void *src, T *const dst[4], const float *uref = nullptr)
{
    const long long N = g.N;
    const int nxy = g.nx * g.ny;
    auto run = [&](auto tag) {
        using ST = decltype(tag);
        using S = typename ST::S;
        const S *s = (const S *)src;
        par_blocks(N, CODEC_BLOCK, [&](long long lo, long long hi, long long) {
            for (long long i = lo; i < hi; i++) {
                const bool wall = hm.wall(i);
                const int z = (int)(i / nxy);
                using L = StateLayout<ST::mode>;
                dst[0][i] = wall ? T(0) : (T)ST::ld_p(s[L::at(0, i, N)], z, q);
                for (int f = 1; f < 3; f++)
                    dst[f][i] = wall ? T(0) : (T)ST::ld_u(s[L::at(f, i, N)], q);
                // the sum rounds once in float, as the kernels decode
                const auto w = ST::ld_u(s[L::at(3, i, N)], q);
                dst[3][i] = wall ? T(0) : uref ? (T)(w + uref[i % nxy]) : (T)w;
            }
        });
    };
    switch (precision) {
    case FSLBM_F64: run(Store<FSLBM_F64>{}); break;
    case FSLBM_F32: run(Store<FSLBM_F32>{}); break;
    case FSLBM_F16S: run(Store<FSLBM_F16S>{}); break;
    case FSLBM_BF16S: run(Store<FSLBM_BF16S>{}); break;
    }
}

// Wall weights of a z-invariant mask, per fluid column (x, y):
// omega0 = sum of W_e, omegax = sum of W_e cx, omegay = sum of W_e cy over
// the neighbours x+e that are walls (or outside the box).  The z sums of
// w(cz) are 1 and of w(cz) cz are 0, so omega_z = 0.  They restore the
// reference's wall rule (own pressure in place of a wall neighbour's,
// FSLBM.cu:137-143) in the separable predictor of the tiled kernels.
template <typename T>
std::vector<T> wall_weights(int nx, int ny, const uint8_t *m2)
{
    const double w[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
    const size_t nxy = (size_t)nx * ny;
    std::vector<T> om(3 * nxy, T(0));
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const size_t c = x + (size_t)nx * y;
            if (m2[c] != 0) continue;
            double o0 = 0, ox = 0, oy = 0;
            for (int cy = -1; cy <= 1; cy++)
                for (int cx = -1; cx <= 1; cx++) {
                    const int xx = x + cx, yy = y + cy;
                    const bool wall = xx < 0 || yy < 0 || xx >= nx || yy >= ny ||
                                      m2[xx + (size_t)nx * yy] != 0;
                    if (!wall) continue;
                    const double W = w[cx + 1] * w[cy + 1];
                    o0 += W;
                    ox += W * cx;
                    oy += W * cy;
                }
            om[c] = (T)o0;
            om[nxy + c] = (T)ox;
            om[2 * nxy + c] = (T)oy;
        }
    return om;
}

// A domain validated and encoded on the host, ready to be uploaded into a
// batch (its precision mode's storage).
struct PreparedDomain {
    Geo g{};
    fslbm_params prm{};
    int precision = FSLBM_F32;
    Quant q{1.f, 1.f, 1.f, 1.f, 0.f, 0.f};
    std::vector<uint8_t> mask2d;      // one z plane; empty for a periodic box
    std::vector<unsigned char> host;  // encoded state, 4 planes
    std::vector<float> omega;         // wall weights (walled domains)
    std::vector<int> tiles;           // tiles holding fluid (walled domains)
    std::vector<float> uref;          // streamwise reference profile (16-bit modes)
    size_t bytes = 0;                 // device bytes in the batch
};

// A domain's stored state copied to the host, with what decoding it
// needs: the geometry, the precision mode, its quantisation and its
// z-invariant mask plane (empty for a periodic box).
struct HostResult {
    Geo g{};
    int precision = FSLBM_F32;
    Quant q{1.f, 1.f, 1.f, 1.f, 0.f, 0.f};
    std::vector<uint8_t> mask2d;
    std::vector<float> uref;  // streamwise reference profile (empty: none)
    std::vector<unsigned char> raw;
    long steps = 0;

    void decode(double *const dst[4]) const
    {
        HostMask hm;
        hm.m = g.periodic ? nullptr : mask2d.data();
        hm.nxy = (long long)g.nx * g.ny;
        hm.plane = true;
        decode_state(g, hm, precision, q, raw.data(), dst, uref.empty() ? nullptr : uref.data());
    }
};

}  // namespace fslbm
