// Plain FSLBM step kernel (one thread per cell, direct 27-point gather)
// and the auxiliary kernels: reductions, monitoring, plane extraction,
// reference statistics and moment accumulation.
//
// Step semantics follow PipeDemoStandalone/FSLBM.cu:
//   predictor  FSLBM.cu:130-173   f_e = W_e (p~ + 3 e.u + 4.5 (e.u)^2
//                                 - 1.5 |u|^2) of the neighbour at x+e,
//                                 p~ = own pressure if that neighbour is
//                                 a wall; P* = sum f_e, u* = sum e f_e
//   corrector  FSLBM.cu:231-257   L = c0 u + c1 sum u(x +- d1 e_a)
//                                        + c2 sum u(x +- 2 d1 e_a)
//   blend      FSLBM.cu:269-273   u' = tb (u* + dtddy L m) + (1-tb) u,
//                                 p' = P* - gauge
//   boundary   FSLBM.cu:280-303   walls hold 0; z periodic with +pin/-pin
//                                 on the pressure of a wrapped neighbour
// The reference updates in place through a shifted plane index and
// copies back (FSLBM.cu:275,307-312,622-626); here the step reads one
// state buffer and writes the other, so every read sees time level n.

#include "fslbm_internal.h"
#include "fslbm_store.h"

namespace fslbm {

StepK<double> make_consts_f64(const StepArgs &a)
{
    StepK<double> k;
    k.w0 = 8.0 / 27.0;
    k.w1 = 2.0 / 27.0;
    k.w2 = 1.0 / 54.0;
    k.quarter = 0.25;
    k.dtddy = a.nu - 1.0 / 6.0;
    k.m = 1.0 / (double)(a.d1 * a.d1);
    k.c = a.tblend;
    k.cc = 1.0 - a.tblend;
    k.c0 = -6.0 - 1.5 * a.blend;
    k.c1 = 1.0 + a.blend / 3.0;
    k.c2 = -a.blend / 12.0;
    k.pin = a.pin;
    k.gauge = a.gauge;
    k.d1 = a.d1;
    k.skey = sr_key(a.seed);
    return k;
}

StepK<float> make_consts_f32(const StepArgs &a)
{
    // FSLBM.cu:60-61 (weights), :599 (dtddy), :624 (m), :231 (c0..c2),
    // :269 (c, cc); run parameters are floats read with %e.
    StepK<float> k;
    const float tb = (float)a.tblend, nu = (float)a.nu, b = (float)a.blend;
    k.w0 = 0.296296296f;
    k.w1 = 0.07407407407f;
    k.w2 = 0.01851851851f;
    k.quarter = 0.25f;
    k.dtddy = nu - 0.16666666f;
    k.m = (float)(1.0 / (double)(float)(a.d1 * a.d1));
    k.c = tb;
    k.cc = (float)(1.0 - (double)tb);
    k.c0 = -6.0f - 1.5f * b;
    k.c1 = 1.0f + 0.3333333333f * b;
    k.c2 = -0.0833333333f * b;
    k.pin = (float)a.pin;
    k.gauge = (float)a.gauge;
    k.d1 = a.d1;
    k.skey = sr_key(a.seed);
    return k;
}

size_t storage_bytes(int precision)
{
    switch (precision) {
    case FSLBM_F64: return 8;
    case FSLBM_F32: return 4;
    default: return 2;
    }
}

// ---------------------------------------------------------------------
// Plain step
// ---------------------------------------------------------------------
template <int MODE, bool PERIODIC>
__global__ __launch_bounds__(256) void k_step_plain(
    const typename Store<MODE>::S *__restrict__ in,
    typename Store<MODE>::S *__restrict__ out,
    const uint8_t *__restrict__ mask, int nx, int ny, int nz,
    StepK<typename Store<MODE>::R> k, Quant q)
{
    using ST = Store<MODE>;
    using R = typename ST::R;
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= nx || y >= ny) return;
    const int nxy = nx * ny;
    const int N = nxy * nz;
    const int i = x + nx * y + nxy * z;
    if (!PERIODIC && mask[i] != 0) return;  // walls stay 0 in both buffers

    using L = StateLayout<MODE>;
    // field f of cell j (64-bit offsets: 4*N may exceed 2^31)
    auto F = [&](int f, int j) { return in[L::at(f, j, N)]; };
    // uz of cell j in column c: the 16-bit modes add the column's
    // streamwise reference
    auto UZ = [&](int j, int c) -> R {
        const R w = ST::ld_u(F(3, j), q);
        if constexpr (ST::SR) return q.uref ? w + q.uref[c] : w;
        else return w;
    };
    const int col = x + nx * y;
    const R ps = ST::ld_p(F(0, i), z, q);
    const R us = ST::ld_u(F(1, i), q), vs = ST::ld_u(F(2, i), q), ws = UZ(i, col);

    R rt = 0, ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int cz = -1; cz <= 1; cz++) {
        int zz = z + cz;
        R jump = 0;
        if (zz < 0) { zz += nz; jump = k.pin; }
        else if (zz >= nz) { zz -= nz; jump = -k.pin; }
#pragma unroll
        for (int cy = -1; cy <= 1; cy++) {
            int yy = y + cy;
            if (PERIODIC) yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
            const R wyz = (cy == 0 && cz == 0) ? k.w0
                          : (cy == 0 || cz == 0) ? k.w1 : k.w2;
#pragma unroll
            for (int cx = -1; cx <= 1; cx++) {
                int xx = x + cx;
                if (PERIODIC) xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
                const int j = xx + nx * yy + nxy * zz;
                const bool wall = !PERIODIC && mask[j] != 0;
                const R pn = wall ? ps : ST::ld_p(F(0, j), zz, q) + jump;
                const R un = ST::ld_u(F(1, j), q), vn = ST::ld_u(F(2, j), q),
                        wn = UZ(j, xx + nx * yy);
                const R e_u = (R)cx * un + (R)cy * vn + (R)cz * wn;
                const R t = (R)3 * e_u;
                const R uu = (R)3 * (un * un + vn * vn + wn * wn);
                R f = wyz * (pn + t + (R)0.5 * (t * t - uu));
                if (cx != 0) f *= k.quarter;
                rt += f;
                if (cx) ax += (R)cx * f;
                if (cy) ay += (R)cy * f;
                if (cz) az += (R)cz * f;
            }
        }
    }

    // CCS corrector: star stencil of reach 2*d1 on the time-n velocity.
    // x and y clamp to the box (texture clamp addressing of the
    // reference) or wrap in a periodic box; z wraps.
    R Lx = k.c0 * us, Ly = k.c0 * vs, Lz = k.c0 * ws;
#pragma unroll
    for (int r = 1; r <= 2; r++) {
        const int d = r * k.d1;
        const R cr = r == 1 ? k.c1 : k.c2;
        int xm = x - d, xp = x + d, ym = y - d, yp = y + d;
        int zm = z - d, zp = z + d;
        if (PERIODIC) {
            xm = xm < 0 ? xm + nx : xm; xp = xp >= nx ? xp - nx : xp;
            ym = ym < 0 ? ym + ny : ym; yp = yp >= ny ? yp - ny : yp;
        } else {
            xm = xm < 0 ? 0 : xm; xp = xp >= nx ? nx - 1 : xp;
            ym = ym < 0 ? 0 : ym; yp = yp >= ny ? ny - 1 : yp;
        }
        zm = zm < 0 ? zm + nz : zm;
        zp = zp >= nz ? zp - nz : zp;
        const int oc[6] = {xp + nx * y, xm + nx * y, x + nx * yp, x + nx * ym, col, col};
        const int o[6] = {oc[0] + nxy * z, oc[1] + nxy * z, oc[2] + nxy * z, oc[3] + nxy * z,
                          col + nxy * zp, col + nxy * zm};
        R sx = 0, sy = 0, sz = 0;
#pragma unroll
        for (int t = 0; t < 6; t++) {
            sx += ST::ld_u(F(1, o[t]), q);
            sy += ST::ld_u(F(2, o[t]), q);
            sz += UZ(o[t], oc[t]);
        }
        Lx += cr * sx;
        Ly += cr * sy;
        Lz += cr * sz;
    }

    const R g = k.dtddy * k.m;
    const SrWords rw = ST::SR ? sr_words(sr_salt((uint32_t)col), sr_plane(k.skey, (uint32_t)z))
                              : SrWords{{0u, 0u, 0u, 0u}};
    out[L::at(0, i, N)] = ST::st_p_sr(rt - k.gauge, z, q, rw.r[0]);
    out[L::at(1, i, N)] = ST::st_u_sr((ax + g * Lx) * k.c + us * k.cc, q, rw.r[1]);
    out[L::at(2, i, N)] = ST::st_u_sr((ay + g * Ly) * k.c + vs * k.cc, q, rw.r[2]);
    R wn = (az + g * Lz) * k.c + ws * k.cc;
    if constexpr (ST::SR)
        if (q.uref) wn -= q.uref[col];
    out[L::at(3, i, N)] = ST::st_u_sr(wn, q, rw.r[3]);
}

template <int MODE>
static hipError_t launch_plain_mode(const Geo &g, const void *in, void *out,
                                    const uint8_t *mask,
                                    const StepK<typename Store<MODE>::R> &k,
                                    const Quant &q, hipStream_t s)
{
    using S = typename Store<MODE>::S;
    dim3 block(64, 4, 1);
    dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    if (g.periodic)
        k_step_plain<MODE, true><<<grid, block, 0, s>>>(
            (const S *)in, (S *)out, mask, g.nx, g.ny, g.nz, k, q);
    else
        k_step_plain<MODE, false><<<grid, block, 0, s>>>(
            (const S *)in, (S *)out, mask, g.nx, g.ny, g.nz, k, q);
    return hipGetLastError();
}

hipError_t launch_step_plain(const Geo &g, int precision, const void *in,
                             void *out, const uint8_t *mask,
                             const StepArgs &a, hipStream_t s)
{
    switch (precision) {
    case FSLBM_F64:
        return launch_plain_mode<FSLBM_F64>(g, in, out, mask,
                                            make_consts_f64(a), a.q, s);
    case FSLBM_F32:
        return launch_plain_mode<FSLBM_F32>(g, in, out, mask,
                                            make_consts_f32(a), a.q, s);
    case FSLBM_F16S:
        return launch_plain_mode<FSLBM_F16S>(g, in, out, mask,
                                             make_consts_f32(a), a.q, s);
    case FSLBM_BF16S:
        return launch_plain_mode<FSLBM_BF16S>(g, in, out, mask,
                                              make_consts_f32(a), a.q, s);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------
// Deterministic two-stage FP64 reductions
// ---------------------------------------------------------------------
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *sh /* 256*NV */)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < NV; c++) sh[c * 256 + t] = v[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int c = 0; c < NV; c++) sh[c * 256 + t] += sh[c * 256 + t + w];
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < NV; c++) v[c] = sh[c * 256];
}

template <int NV>
__global__ __launch_bounds__(256) void k_finish(const double *partials,
                                                int npart, double *dst)
{
    __shared__ double sh[256 * NV];
    double v[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) v[c] = 0;
    for (int b = threadIdx.x; b < npart; b += 256)
#pragma unroll
        for (int c = 0; c < NV; c++) v[c] += partials[b * NV + c];
    block_sum<NV>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < NV; c++) dst[c] = v[c];
}

template <int MODE>
__global__ __launch_bounds__(256) void k_reduce(
    const typename Store<MODE>::S *__restrict__ st,
    const uint8_t *__restrict__ mask, int nxy, int N, Quant q,
    double *partials)
{
    using ST = Store<MODE>;
    __shared__ double sh[256 * 8];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // 64-bit loop index: i + the grid stride may pass 2^31 for N < 2^31
    for (long long i = blockIdx.x * 256 + threadIdx.x; i < N; i += (long long)gridDim.x * 256) {
        if (mask && mask[i] != 0) continue;
        using L = StateLayout<MODE>;
        const double p = ST::ld_p(st[L::at(0, i, N)], i / nxy, q);
        const double u = ST::ld_u(st[L::at(1, i, N)], q);
        const double w = ST::ld_u(st[L::at(2, i, N)], q);
        const double z = ld_uz<MODE>(st[L::at(3, i, N)], q, i % nxy);
        v[0] += 1.0;
        v[1] += p;
        v[2] += u;
        v[3] += w;
        v[4] += z;
        v[5] += u * u + w * w + z * z;
        v[6] += p * p;
        v[7] += (isfinite(p) && isfinite(u) && isfinite(w) && isfinite(z)) ? 0.0
                : (double)(!isfinite(p) + !isfinite(u) + !isfinite(w) + !isfinite(z));
    }
    block_sum<8>(v, sh);
    if (threadIdx.x == 0)
        for (int c = 0; c < 8; c++) partials[blockIdx.x * 8 + c] = v[c];
}

hipError_t launch_reduce(const Geo &g, int precision, const void *state,
                         const uint8_t *mask, const Quant &q, double *partials,
                         int npart, double *dst, hipStream_t s)
{
    const int nxy = g.nx * g.ny, N = (int)g.N;
    switch (precision) {
    case FSLBM_F64:
        k_reduce<FSLBM_F64><<<npart, 256, 0, s>>>((const double *)state, mask, nxy, N, q, partials);
        break;
    case FSLBM_F32:
        k_reduce<FSLBM_F32><<<npart, 256, 0, s>>>((const float *)state, mask, nxy, N, q, partials);
        break;
    case FSLBM_F16S:
        k_reduce<FSLBM_F16S><<<npart, 256, 0, s>>>((const _Float16 *)state, mask, nxy, N, q, partials);
        break;
    case FSLBM_BF16S:
        k_reduce<FSLBM_BF16S><<<npart, 256, 0, s>>>((const uint16_t *)state, mask, nxy, N, q, partials);
        break;
    default:
        return hipErrorInvalidValue;
    }
    k_finish<8><<<1, 256, 0, s>>>(partials, npart, dst);
    return hipGetLastError();
}

// Monitoring over folded plane z = 0 (FSLBM.cu:635-645).
template <int MODE>
__global__ __launch_bounds__(256) void k_monitor(
    const typename Store<MODE>::S *__restrict__ st,
    const uint8_t *__restrict__ mask, int nxy, int N, Quant q,
    double *partials)
{
    using ST = Store<MODE>;
    __shared__ double sh[256 * 3];
    double v[3] = {0, 0, 0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nxy; i += gridDim.x * 256) {
        if (mask && mask[i] != 0) continue;
        v[0] += 1.0;
        v[1] += ST::ld_p(st[StateLayout<MODE>::at(0, i, N)], 0, q);
        v[2] += ld_uz<MODE>(st[StateLayout<MODE>::at(3, i, N)], q, i);
    }
    block_sum<3>(v, sh);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) partials[blockIdx.x * 3 + c] = v[c];
}

hipError_t launch_monitor(const Geo &g, int precision, const void *state,
                          const uint8_t *mask, const Quant &q,
                          double *partials, int npart, double *dst,
                          hipStream_t s)
{
    const int nxy = g.nx * g.ny, N = (int)g.N;
    switch (precision) {
    case FSLBM_F64:
        k_monitor<FSLBM_F64><<<npart, 256, 0, s>>>((const double *)state, mask, nxy, N, q, partials);
        break;
    case FSLBM_F32:
        k_monitor<FSLBM_F32><<<npart, 256, 0, s>>>((const float *)state, mask, nxy, N, q, partials);
        break;
    case FSLBM_F16S:
        k_monitor<FSLBM_F16S><<<npart, 256, 0, s>>>((const _Float16 *)state, mask, nxy, N, q, partials);
        break;
    case FSLBM_BF16S:
        k_monitor<FSLBM_BF16S><<<npart, 256, 0, s>>>((const uint16_t *)state, mask, nxy, N, q, partials);
        break;
    default:
        return hipErrorInvalidValue;
    }
    k_finish<3><<<1, 256, 0, s>>>(partials, npart, dst);
    return hipGetLastError();
}

// ---------------------------------------------------------------------
// Padded-layout views of the folded state
// ---------------------------------------------------------------------
// Value of field f at folded (x, y) and padded plane zp with d2 padding
// planes: the inlet padding holds the periodic image with +pin, the
// outlet padding the image with -pin (bcond, FSLBM.cu:294-301).
template <int MODE>
__device__ inline double padded_value(const typename Store<MODE>::S *st,
                                      const uint8_t *mask, int nx, int nxy,
                                      int nz, int N, const Quant &q,
                                      double pin, int x, int y, int zp,
                                      int d2, int f)
{
    using ST = Store<MODE>;
    int z = zp - d2;
    double jump = 0;
    if (z < 0) { z += nz; jump = pin; }
    else if (z >= nz) { z -= nz; jump = -pin; }
    const int i = x + nx * y + nxy * z;
    if (mask && mask[i] != 0) return 0.0;
    if (f == 0) return (double)ST::ld_p(st[StateLayout<MODE>::at(0, i, N)], z, q) + jump;
    if (f == 3) return (double)ld_uz<MODE>(st[StateLayout<MODE>::at(3, i, N)], q, x + nx * y);
    return (double)ST::ld_u(st[StateLayout<MODE>::at(f, i, N)], q);
}

template <int MODE>
__global__ void k_plane(const typename Store<MODE>::S *__restrict__ st,
                        const uint8_t *__restrict__ mask, int nx, int ny,
                        int nz, Quant q, double pin, int axis, int idx, int d2,
                        int f, double *dst)
{
    const int nxy = nx * ny, N = nxy * nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (axis == 0) {
        const int lz = nz + 2 * d2;
        if (t >= ny * lz) return;
        const int y = t % ny, zp = t / ny;
        dst[t] = padded_value<MODE>(st, mask, nx, nxy, nz, N, q, pin, idx, y,
                                    zp, d2, f);
    } else {
        if (t >= nxy) return;
        const int x = t % nx, y = t / nx;
        dst[t] = padded_value<MODE>(st, mask, nx, nxy, nz, N, q, pin, x, y,
                                    idx, d2, f);
    }
}

hipError_t launch_plane(const Geo &g, int precision, const void *state,
                        const uint8_t *mask, const Quant &q, double pin,
                        int axis, int idx, int d2, int field, double *dst,
                        hipStream_t s)
{
    const int n = axis == 0 ? g.ny * (g.nz + 2 * d2) : g.nx * g.ny;
    const int blocks = (n + 255) / 256;
    switch (precision) {
    case FSLBM_F64:
        k_plane<FSLBM_F64><<<blocks, 256, 0, s>>>((const double *)state, mask, g.nx, g.ny, g.nz, q, pin, axis, idx, d2, field, dst);
        break;
    case FSLBM_F32:
        k_plane<FSLBM_F32><<<blocks, 256, 0, s>>>((const float *)state, mask, g.nx, g.ny, g.nz, q, pin, axis, idx, d2, field, dst);
        break;
    case FSLBM_F16S:
        k_plane<FSLBM_F16S><<<blocks, 256, 0, s>>>((const _Float16 *)state, mask, g.nx, g.ny, g.nz, q, pin, axis, idx, d2, field, dst);
        break;
    case FSLBM_BF16S:
        k_plane<FSLBM_BF16S><<<blocks, 256, 0, s>>>((const uint16_t *)state, mask, g.nx, g.ny, g.nz, q, pin, axis, idx, d2, field, dst);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Reference statistics, one sample (FSLBM.cu:715-748).  The sample is
// the padded YZ plane at x = xplane, widened to double as the reference
// widens its float plane.  acc: 11 arrays of n = ny*(nz+2*d2) doubles.
template <int MODE>
__global__ void k_stats(const typename Store<MODE>::S *__restrict__ st,
                        const uint8_t *__restrict__ mask, int nx, int ny,
                        int nz, Quant q, double pin, int xplane, int d2,
                        double *acc, int cnt)
{
    const int nxy = nx * ny, N = nxy * nz;
    const int lz = nz + 2 * d2, n = ny * lz;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int y = k % ny, zp = k / ny;
    double s[4];
    for (int f = 0; f < 4; f++)
        s[f] = padded_value<MODE>(st, mask, nx, nxy, nz, N, q, pin, xplane, y,
                                  zp, d2, f);
    const double c = (double)cnt;
    double dd[4];
    for (int f = 0; f < 4; f++) {
        double a = acc[f * n + k] + s[f];
        acc[f * n + k] = a;
        dd[f] = a / c - s[f];
    }
    for (int f = 0; f < 4; f++) acc[(4 + f) * n + k] += fabs(dd[f]);
    acc[8 * n + k] += dd[1] * dd[2];
    acc[9 * n + k] += dd[1] * dd[3];
    acc[10 * n + k] += dd[3] * dd[2];
}

hipError_t launch_stats(const Geo &g, int precision, const void *state,
                        const uint8_t *mask, const Quant &q, double pin,
                        int xplane, int d2, double *acc, int cnt,
                        hipStream_t s)
{
    const int n = g.ny * (g.nz + 2 * d2);
    const int blocks = (n + 255) / 256;
    switch (precision) {
    case FSLBM_F64:
        k_stats<FSLBM_F64><<<blocks, 256, 0, s>>>((const double *)state, mask, g.nx, g.ny, g.nz, q, pin, xplane, d2, acc, cnt);
        break;
    case FSLBM_F32:
        k_stats<FSLBM_F32><<<blocks, 256, 0, s>>>((const float *)state, mask, g.nx, g.ny, g.nz, q, pin, xplane, d2, acc, cnt);
        break;
    case FSLBM_F16S:
        k_stats<FSLBM_F16S><<<blocks, 256, 0, s>>>((const _Float16 *)state, mask, g.nx, g.ny, g.nz, q, pin, xplane, d2, acc, cnt);
        break;
    case FSLBM_BF16S:
        k_stats<FSLBM_BF16S><<<blocks, 256, 0, s>>>((const uint16_t *)state, mask, g.nx, g.ny, g.nz, q, pin, xplane, d2, acc, cnt);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Moment sums per cross-section cell over all z: acc has 11 arrays of
// nx*ny doubles (p, ux, uy, uz, p^2, ux^2, uy^2, uz^2, uxuy, uxuz, uyuz).
template <int MODE>
__global__ __launch_bounds__(256) void k_moments(
    const typename Store<MODE>::S *__restrict__ st,
    const uint8_t *__restrict__ mask, int nxy, int nz, Quant q, double *acc)
{
    using ST = Store<MODE>;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nxy) return;
    const int N = nxy * nz;
    double s[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int z = 0; z < nz; z++) {
        const int i = c + nxy * z;
        if (mask && mask[i] != 0) continue;
        using L = StateLayout<MODE>;
        const double p = ST::ld_p(st[L::at(0, i, N)], z, q);
        const double u = ST::ld_u(st[L::at(1, i, N)], q);
        const double v = ST::ld_u(st[L::at(2, i, N)], q);
        const double w = ld_uz<MODE>(st[L::at(3, i, N)], q, c);
        s[0] += p; s[1] += u; s[2] += v; s[3] += w;
        s[4] += p * p; s[5] += u * u; s[6] += v * v; s[7] += w * w;
        s[8] += u * v; s[9] += u * w; s[10] += v * w;
    }
    for (int f = 0; f < 11; f++) acc[f * nxy + c] += s[f];
}

hipError_t launch_moments(const Geo &g, int precision, const void *state,
                          const uint8_t *mask, const Quant &q, double *acc,
                          hipStream_t s)
{
    const int nxy = g.nx * g.ny;
    const int blocks = (nxy + 255) / 256;
    switch (precision) {
    case FSLBM_F64:
        k_moments<FSLBM_F64><<<blocks, 256, 0, s>>>((const double *)state, mask, nxy, g.nz, q, acc);
        break;
    case FSLBM_F32:
        k_moments<FSLBM_F32><<<blocks, 256, 0, s>>>((const float *)state, mask, nxy, g.nz, q, acc);
        break;
    case FSLBM_F16S:
        k_moments<FSLBM_F16S><<<blocks, 256, 0, s>>>((const _Float16 *)state, mask, nxy, g.nz, q, acc);
        break;
    case FSLBM_BF16S:
        k_moments<FSLBM_BF16S><<<blocks, 256, 0, s>>>((const uint16_t *)state, mask, nxy, g.nz, q, acc);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void k_zero(double *dst, long long n)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        dst[i] = 0.0;
}

hipError_t launch_zero(double *dst, long long n, hipStream_t s)
{
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    k_zero<<<(int)blocks, 256, 0, s>>>(dst, n);
    return hipGetLastError();
}

}  // namespace fslbm
