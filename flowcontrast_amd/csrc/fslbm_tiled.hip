// Tiled, z-marching FSLBM step kernel for gfx950.
//
// Same update as the plain kernel (FSLBM.cu:130-303 semantics), organised
// for HBM bandwidth:
//
// * One workgroup owns a 64 x TY (x, y) tile and marches over a chunk of
//   z planes: TY = 8 in the two-barrier FP64 kernel (k_step_tiled); in
//   the pipelined kernel (k_step_pipe, further below) 12 (FP32, and the
//   16-bit storage modes in a walled box at blend 0), 8 (other 16-bit
//   variants, masks that vary along z) or 4 (FP64).  Every state value is
//   written once per step and read from HBM about once (halo and chunk-end
//   re-reads mostly hit L2), except wall cells, whose zeros the pipelined
//   kernel does not move where a 16-cell row piece is all walls.
// * The 27-point D3Q27 predictor is evaluated as three separable 3-tap
//   passes.  D3Q27 weights are products w(cx) w(cy) w(cz) with
//   w(0) = 2/3, w(+-1) = 1/6, and the equilibrium
//     g_e = A + 3 e.u + 4.5 (e.u)^2,   A = p - 1.5 |u|^2
//   is a quadratic polynomial in e, so sum_e W_e g_e(x+e) (1, e) becomes:
//     z pass  (registers, own column): 12 coefficients of a polynomial
//             in (cx, cy) from the planes z-1, z, z+1;
//     x pass  (LDS in k_step_tiled, DPP lane shifts in k_step_pipe):
//             each column publishes that polynomial evaluated at
//             cx = +1 and cx = -1 (10 values), neighbours gather;
//     y pass  (LDS): each row publishes the cy = +1 / -1 values of the
//             three remaining polynomials (6 values).
//   Weights are carried unnormalised (x6 per axis) and scaled by 1/216
//   at the end.  Wall cells hold zero state, so they add nothing to the
//   passes; the reference's wall rule (own pressure in place of a wall
//   neighbour's, FSLBM.cu:137-143) is added back as p_self * omega with
//   omega = sum over wall neighbours of W_e (1, e), precomputed per
//   column for z-invariant masks.
// * The CCS corrector reads its x/y arms from a double-buffered LDS tile
//   of the plane (halo 2*d1, clamped or wrapped like the reference's
//   texture reads) and its z arms from a register queue that runs
//   several planes ahead of the output plane (loads in flight).
// * Workgroups are mapped XCD by XCD: the 8 XCDs each take a contiguous
//   run of work items (z-chunk major), so tiles that share halos share
//   an L2.
// * Work items run over the tiles that hold a fluid cell only: a tile of
//   wall cells would read zeros and store zeros over the zeros both state
//   buffers hold at walls from the upload on.  For the same reason the
//   pipelined kernel neither reads nor writes the 16-lane quarter-rows of
//   a tile whose columns are all walls.  The z planes split
//   into chunks whose lengths differ by at most one, and the chunk count
//   is chosen so that the items fill whole rounds of workgroups.
//
// Roles of the 704 threads (11 waves) of k_step_tiled:
//   waves 0-7  interior columns (64 x 8): everything
//   waves 8-9  halo rows y0-1 and y0+8: z pass, x pass, y messages
//   wave 10    halo columns x0-1 and x0+64 (20 lanes): z pass, x messages

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <utility>

#include "fslbm_internal.h"
#include "fslbm_store.h"

namespace fslbm {

namespace {

constexpr int TX = 64, TY = 8;
constexpr int RX = TX + 2, RY = TY + 2;
constexpr int NI = TX * TY;       // interior threads
constexpr int NH = 2 * TX;        // halo-row threads
constexpr int NS = 2 * RY;        // halo-column threads
constexpr int NT = 704;           // 11 waves
// tile rows of the pipelined kernel: 12 for FP32 (15 waves fit the
// 128-register budget without spills), 8 for the 16-bit storage modes
// (their decode needs more registers).  Waves go to the four SIMDs round
// robin; these counts keep the busiest SIMD's VALU load lowest (13 rows
// or 9 rows put a fourth or third interior wave on one SIMD).
constexpr int pipe_rows(int mode) { return mode == FSLBM_F32 ? 12 : 8; }
// Both 16-bit storage modes in a walled box without order-4 arms fit 12
// rows (15 waves) in 128 registers too: 256x256x1024 pipe, FP16 0.341 ms
// and BF16 0.324 ms per step (BF16 0.361 ms on 8 rows).  Their periodic
// and order-4 variants spill at 12 rows and keep 8.
constexpr int F16S_WALLED_ROWS = 12;
// FP32 periodic boxes may take 8-row tiles (solo_rows): 11 waves, a
// larger halo share per cell, but more tiles to fill whole rounds of
// workgroups.  An 8-row tile-plane takes F32_BOX_ROWS_COST of a 12-row
// one's time: 256^3 box, 135.3 us for 256 items of 128 planes against
// 144.6 us for 440 items of 52 planes.
constexpr int F32_BOX_ROWS = 8;
constexpr double F32_BOX_ROWS_COST = 0.79;
// Masks that vary along z (wall codes per cell) take 8-row tiles in every
// storage mode: the code queues and the wall-weight arithmetic need more
// registers than 12-row tiles (15 waves, 128 registers) leave.
constexpr int M3_ROWS = 8;
// FP64 on the pipelined kernel: 4-row tiles, 7 waves (two per SIMD) in
// 238 registers without spills; at 6 or 8 rows (three waves per SIMD, at
// most 168 registers) it spills 380 bytes per lane.
constexpr int F64_PIPE_ROWS = 4;

template <typename R>
struct Msg {
    R A, ux, uy, xx, yy, xy, aP, aM, bP, bM, cP, cM;
};

template <typename R>
__device__ __forceinline__ Msg<R> make_msg(R p, R ux, R uy, R uz)
{
    Msg<R> m;
    const R uxx = ux * ux, uyy = uy * uy, uzz = uz * uz;
    m.A = p - R(1.5) * (uxx + uyy + uzz);
    m.ux = ux;
    m.uy = uy;
    m.xx = uxx;
    m.yy = uyy;
    m.xy = ux * uy;
    const R uxz = ux * uz, uyz = uy * uz;
    const R t = m.A + R(4.5) * uzz;
    m.aP = t + R(3) * uz;
    m.aM = t - R(3) * uz;
    m.bP = ux + R(3) * uxz;
    m.bM = ux - R(3) * uxz;
    m.cP = uy + R(3) * uyz;
    m.cM = uy - R(3) * uyz;
    return m;
}

template <typename R>
__device__ __forceinline__ Msg<R> zero_msg()
{
    Msg<R> m;
    m.A = m.ux = m.uy = m.xx = m.yy = m.xy = R(0);
    m.aP = m.aM = m.bP = m.bM = m.cP = m.cM = R(0);
    return m;
}

// z pass over planes z-1 (mm), z (m0), z+1 (mp) and the x messages.
template <typename R>
struct XMsg {
    R v[10];      // P0+, P0-, P1+, P1-, P2, Z0+, Z0-, Z1+, Z1-, Z2
    R cA, cC, dA, dC;
};

template <typename R>
__device__ __forceinline__ XMsg<R> zpass(const Msg<R> &mm, const Msg<R> &m0,
                                         const Msg<R> &mp)
{
    const R Ca = R(4) * m0.A + (mp.aP + mm.aM), Da = mp.aP - mm.aM;
    const R Cb = R(4) * m0.ux + (mp.bP + mm.bM), Db = mp.bP - mm.bM;
    const R Cc = R(4) * m0.uy + (mp.cP + mm.cM), Dc = mp.cP - mm.cM;
    const R Cd = R(4) * m0.xx + (mp.xx + mm.xx), Dd = mp.xx - mm.xx;
    const R Ce = R(4) * m0.yy + (mp.yy + mm.yy), De = mp.yy - mm.yy;
    const R Ch = R(4) * m0.xy + (mp.xy + mm.xy), Dh = mp.xy - mm.xy;
    XMsg<R> x;
    const R t = Ca + R(4.5) * Cd;
    x.v[0] = t + R(3) * Cb;
    x.v[1] = t - R(3) * Cb;
    const R c3 = R(3) * Cc;
    x.v[2] = c3 + R(9) * Ch;
    x.v[3] = c3 - R(9) * Ch;
    x.v[4] = R(4.5) * Ce;
    const R s = Da + R(4.5) * Dd;
    x.v[5] = s + R(3) * Db;
    x.v[6] = s - R(3) * Db;
    const R d3 = R(3) * Dc;
    x.v[7] = d3 + R(9) * Dh;
    x.v[8] = d3 - R(9) * Dh;
    x.v[9] = R(4.5) * De;
    x.cA = Ca;
    x.cC = c3;
    x.dA = Da;
    x.dC = d3;
    return x;
}

template <typename R>
struct Vec3 {
    R x, y, z;
};

// LDS cell of the corrector tile: float4 for FP32 (one 16-byte read),
// three doubles for FP64.
template <typename R>
struct FdCell;
template <>
struct alignas(16) FdCell<float> {
    float x, y, z, w;
};
template <>
struct FdCell<double> {
    double x, y, z;
};

template <typename R>
struct TiledParams {
    int nx, ny, nz, nxy;
    int ntx, nty, chunk, nchunks, nwork, per_xcd;
    // work items: chunk-major over ntile tiles (indices ty*ntx + tx from
    // `tiles`, or all tiles in order when it is null); chunk c covers
    // cq + (c < cr) planes from c*cq + min(c, cr)
    int ntile, cq, cr;
    // 1: both state buffers hold zero at every wall cell, so half-rows of
    // wall columns need no loads or stores (pipe_item)
    int wskip;
    const int *tiles;
    R pin, gauge;
    R c0, c1, c2, g, tb, cc;
    // blend folded into the corrector and the 1/216 scale of the
    // unnormalised passes: u' = ku*u + ks1*S1 + ks2*S2 + kx*X + tb*omega*p
    R ku, ks1, ks2, kx;
    Quant q;
    uint32_t skey;  // stochastic-rounding key of this step (16-bit modes)
};

// Tile index and z range [zs, ze) of work item w.
template <typename R>
__device__ __forceinline__ void item_of(const TiledParams<R> &P, int w, int &tile, int &zs,
                                        int &ze)
{
    const int c = w / P.ntile;
    const int slot = w - c * P.ntile;
    tile = P.tiles ? P.tiles[slot] : slot;
    zs = c * P.cq + min(c, P.cr);
    ze = zs + P.cq + (c < P.cr ? 1 : 0);
}

__device__ __forceinline__ int wrapi(int v, int n)
{
    return v < 0 ? v + n : (v >= n ? v - n : v);
}
__device__ __forceinline__ int clampi(int v, int n)
{
    return v < 0 ? 0 : (v >= n ? n - 1 : v);
}
// Periodic index for set-up coordinates, which may lie several periods
// away when the box is narrower than the tile.
__device__ __forceinline__ int modp(int v, int n)
{
    const int r = v % n;
    return r < 0 ? r + n : r;
}

constexpr int lcm3(int q) { return q % 3 == 0 ? q : 3 * q; }

// Buffer-instruction access: a resource (base, size) in scalar registers,
// a scalar plane offset and a 32-bit lane offset, so a load or store
// costs no per-lane address arithmetic.  Word 3 of the resource is the
// gfx9 untyped-buffer format word.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mkbuf(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, bytes, 0x00020000);
}
template <typename S>
__device__ __forceinline__ S bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// With NT, FP32 state is stored non-temporal (cache-policy bit NT):
// nothing in the launch reads it back, so its lines do not displace the
// input planes that neighbouring tiles still read from L2 (halo rows and
// columns).  That pays only when the two state buffers exceed the 256 MB
// Infinity Cache (stream_stores): then +4.5% on a 2 GB pipe and +6% on a
// 393 MB one.  Below it, ordinary stores leave the new state in that cache
// for the next step to read, and NT measured 35% slower (138 MB pipe) and
// 22% slower (226 MB Taylor-Green box).  16-bit stores stay ordinary: with
// NT they wrote 8% more bytes and the FP16-storage Taylor-Green step ran
// 10% slower.  Non-temporal loads lose the halo sharing: 30% slower.
constexpr int STORE_NT = 2;
template <bool NT>
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff,
                                          NT ? STORE_NT : 0);
}
template <bool NT>
__device__ __forceinline__ void bstore(double v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, NT ? STORE_NT : 0);
}
template <bool NT>
__device__ __forceinline__ void bstore(_Float16 v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, voff, soff, 0);
}
template <bool NT>
__device__ __forceinline__ void bstore(uint16_t v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, voff, soff, 0);
}

// Planes of look-ahead beyond the corrector reach, chosen so the queue
// length 4*d1 + 1 + LA has a small common multiple with 3 (static slots
// of the three-plane message queue after unrolling).
template <int D1> struct Ahead;
template <> struct Ahead<1> { static constexpr int LA = 1; };  // queue 6
template <> struct Ahead<2> { static constexpr int LA = 3; };  // queue 12

template <int MODE, bool PERIODIC, int D1>
__global__ __launch_bounds__(NT) void k_step_tiled(
    const typename Store<MODE>::S *__restrict__ in,
    typename Store<MODE>::S *__restrict__ out,
    const uint8_t *__restrict__ mask2d,
    const typename Store<MODE>::R *__restrict__ omega,
    TiledParams<typename Store<MODE>::R> P)
{
    using ST = Store<MODE>;
    using R = typename ST::R;
    using S = typename ST::S;
    constexpr int HF = 2 * D1;              // corrector reach
    constexpr int LA = Ahead<D1>::LA;       // extra look-ahead planes
    constexpr int QL = 2 * HF + 1 + LA;     // planes zo-HF .. zo+HF+LA
    constexpr int U = lcm3(QL);             // unroll: static queue slots
    static_assert(U % 2 == 0, "two-deep prefetch slots need an even unroll");
    constexpr int FX = TX + 2 * HF, FY = TY + 2 * HF;

    __shared__ R xmsg[10][RY][RX];
    __shared__ R ymsg[6][RY][TX];
    __shared__ FdCell<R> fd[2][FY][FX];

    // ---- work item (XCD-major order) -------------------------------------
    const int bid = blockIdx.x;
    const int w = (bid & 7) * P.per_xcd + (bid >> 3);
    if (w >= P.nwork) return;
    int tile, zs, ze;
    item_of(P, w, tile, zs, ze);
    const int ty = tile / P.ntx, tx = tile - ty * P.ntx;
    const int x0 = tx * TX, y0 = ty * TY;
    const int nx = P.nx, ny = P.ny, nz = P.nz, nxy = P.nxy;
    const int N = nxy * nz;

    // ---- role and column ---------------------------------------------------
    const int tid = threadIdx.x;
    int role;  // 0 interior, 1 halo row, 2 halo column, 3 none
    int x = 0, y = 0, rx = 0, ry = 0, ix = 0, iy = 0;
    if (tid < NI) {
        role = 0;
        ix = tid & (TX - 1);
        iy = tid / TX;
        x = x0 + ix;
        y = y0 + iy;
        rx = ix + 1;
        ry = iy + 1;
    } else if (tid < NI + NH) {
        role = 1;
        const int j = tid - NI;
        ix = j & (TX - 1);
        const int hr = j / TX;
        x = x0 + ix;
        y = hr ? y0 + TY : y0 - 1;
        rx = ix + 1;
        ry = hr ? TY + 1 : 0;
    } else if (tid < NI + NH + NS) {
        role = 2;
        const int k = tid - NI - NH;
        const int side = k / RY, r = k - side * RY;
        x = side ? x0 + TX : x0 - 1;
        y = y0 - 1 + r;
        rx = side ? TX + 1 : 0;
        ry = r;
    } else {
        role = 3;
    }
    const bool inside = x >= 0 && x < nx && y >= 0 && y < ny;
    const int xm = PERIODIC ? modp(x, nx) : clampi(x, nx);
    const int ym = PERIODIC ? modp(y, ny) : clampi(y, ny);
    const int col = xm + nx * ym;
    bool fluid = false;
    if (role != 3) fluid = PERIODIC ? true : (inside && mask2d[col] == 0);
    R om0 = 0, omx = 0, omy = 0;
    if (!PERIODIC && role == 0 && fluid) {
        om0 = omega[col];
        omx = omega[nxy + col];
        omy = omega[2 * nxy + col];
    }
    const bool store = role == 0 && inside && fluid;

    // ---- corrector halo arms: (fy, fx) tile position and memory column ----
    int arm_fy = -1, arm_fx = 0, arm_col = 0;
    if (tid < 2 * HF * TX) {  // y arms: rows above and below the tile
        const int ar = tid / TX, ac = tid - ar * TX;
        arm_fy = ar < HF ? ar : TY + ar;
        arm_fx = HF + ac;
    } else if (tid >= NI && tid < NI + 2 * HF * TY) {  // x arms
        const int b = tid - NI;
        const int br = b / (2 * HF), bc = b - br * (2 * HF);
        arm_fy = HF + br;
        arm_fx = bc < HF ? bc : TX + bc;
    }
    if (arm_fy >= 0) {
        const int gx = x0 - HF + arm_fx, gy = y0 - HF + arm_fy;
        const int ax = PERIODIC ? modp(gx, nx) : clampi(gx, nx);
        const int ay = PERIODIC ? modp(gy, ny) : clampi(gy, ny);
        arm_col = ax + nx * ay;
    }

    static_assert(!StateLayout<MODE>::PAIRS, "the two-barrier kernel steps four-plane states");
    const S *__restrict__ Pi = in;
    const S *__restrict__ Ui = in + N;
    const S *__restrict__ Vi = in + 2ll * N;
    const S *__restrict__ Wi = in + 3ll * N;

    // logical plane -> (physical plane, pressure jump)
    auto phys = [&](int zl, R &jump) -> int {
        jump = R(0);
        if (zl < 0) { jump = P.pin; return zl + nz; }
        if (zl >= nz) { jump = -P.pin; return zl - nz; }
        return zl;
    };

    // register queue of the own column (interior threads): planes
    // zo-HF .. zo+HF+LA in slots (i + j - HF - LA) mod QL
    R qp[QL], qx[QL], qy[QL], qz[QL];
    Msg<R> mq[3];
    // two-deep prefetch of halo-row / halo-column threads, by plane parity
    R hp[2] = {0, 0}, hx[2] = {0, 0}, hy[2] = {0, 0}, hz[2] = {0, 0};
    // two-deep prefetch of the corrector arm, by plane parity
    R ax_[2] = {0, 0}, ay_[2] = {0, 0}, az_[2] = {0, 0};

#pragma unroll
    for (int k = 0; k < QL; k++) qp[k] = qx[k] = qy[k] = qz[k] = R(0);
#pragma unroll
    for (int k = 0; k < 3; k++) mq[k] = zero_msg<R>();

    const int nit = (ze - zs) + 2 * HF + LA;
    const int zfirst = zs - 2 * HF - LA;

    for (int ib = 0; ib < nit; ib += U) {
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int i = ib + k;
            if (i >= nit) break;
            const int zo = zfirst + i;

            // -- halo threads: message of plane zo+1, then prefetch plane
            //    zo+3 into the slot it used (same parity) --
            if (role == 1 || role == 2) {
                const int sl = (k + 1) & 1;  // parity slot of planes zo+1, zo+3
                if (zo + 1 >= zs - 1 && zo + 1 <= ze)
                    mq[(k + 1) % 3] = fluid ? make_msg<R>(hp[sl], hx[sl], hy[sl], hz[sl])
                                            : zero_msg<R>();
                const int zl = zo + 3;
                if (zl >= zs - 1 && zl <= ze) {
                    R jump;
                    const int zp = phys(zl, jump);
                    const int gi = col + nxy * zp;
                    hp[sl] = ST::ld_p(Pi[gi], zp, P.q) + jump;
                    hx[sl] = ST::ld_u(Ui[gi], P.q);
                    hy[sl] = ST::ld_u(Vi[gi], P.q);
                    hz[sl] = ST::ld_u(Wi[gi], P.q);
                }
            }
            // -- interior threads: load plane zo+HF+LA into the queue --
            if (role == 0) {
                R jump;
                const int zl = zo + HF + LA;
                const int zp = phys(zl, jump);
                const int gi = col + nxy * zp;
                const int s = k % QL;
                qp[s] = fluid ? ST::ld_p(Pi[gi], zp, P.q) + jump : R(0);
                qx[s] = ST::ld_u(Ui[gi], P.q);
                qy[s] = ST::ld_u(Vi[gi], P.q);
                qz[s] = ST::ld_u(Wi[gi], P.q);
                // message of plane zo+1 (loaded HF+LA-1 iterations ago)
                if (zo + 1 >= zs - 1) {
                    const int s1 = (k + 1 - HF - LA + 2 * QL) % QL;
                    mq[(k + 1) % 3] = fluid ? make_msg<R>(qp[s1], qx[s1], qy[s1], qz[s1])
                                            : zero_msg<R>();
                }
            }
            // -- corrector arms: write plane zo, prefetch plane zo+2 --
            if (arm_fy >= 0 && zo >= zs - 2) {
                const int t = k & 1;  // parity slot of planes zo and zo+2
                if (zo >= zs) {
                    FdCell<R> c;
                    c.x = ax_[t];
                    c.y = ay_[t];
                    c.z = az_[t];
                    fd[zo & 1][arm_fy][arm_fx] = c;
                }
                if (zo + 2 < ze) {
                    const int gi = arm_col + nxy * wrapi(zo + 2, nz);
                    ax_[t] = ST::ld_u(Ui[gi], P.q);
                    ay_[t] = ST::ld_u(Vi[gi], P.q);
                    az_[t] = ST::ld_u(Wi[gi], P.q);
                }
            }
            if (zo < zs) continue;

            // ================= output plane zo =================
            const int buf = zo & 1;
            const int s0 = (k - HF - LA + 2 * QL) % QL;  // queue slot of plane zo
            if (role == 0) {
                FdCell<R> c;
                c.x = qx[s0];
                c.y = qy[s0];
                c.z = qz[s0];
                fd[buf][HF + iy][HF + ix] = c;
            }
            XMsg<R> xm_;
            if (role <= 2) {
                xm_ = zpass<R>(mq[(k + 2) % 3], mq[k % 3], mq[(k + 1) % 3]);
#pragma unroll
                for (int v = 0; v < 10; v++) xmsg[v][ry][rx] = xm_.v[v];
            }
            __syncthreads();

            R b0 = 0, g0 = 0, d0 = 0;
            if (role <= 1) {
                R e[10], wv[10];
#pragma unroll
                for (int v = 0; v < 10; v++) {
                    e[v] = xmsg[v][ry][rx + 1];
                    wv[v] = xmsg[v][ry][rx - 1];
                }
                b0 = R(4) * xm_.cA + (e[0] + wv[1]);
                const R b1 = R(4) * xm_.cC + (e[2] + wv[3]);
                const R b2 = R(4) * xm_.v[4] + (e[4] + wv[4]);
                g0 = e[0] - wv[1];
                const R g1 = e[2] - wv[3];
                const R g2 = e[4] - wv[4];
                d0 = R(4) * xm_.dA + (e[5] + wv[6]);
                const R d1_ = R(4) * xm_.dC + (e[7] + wv[8]);
                const R d2 = R(4) * xm_.v[9] + (e[9] + wv[9]);
                const R tb_ = b0 + b2, tg = g0 + g2, td = d0 + d2;
                ymsg[0][ry][ix] = tb_ + b1;
                ymsg[1][ry][ix] = tb_ - b1;
                ymsg[2][ry][ix] = tg + g1;
                ymsg[3][ry][ix] = tg - g1;
                ymsg[4][ry][ix] = td + d1_;
                ymsg[5][ry][ix] = td - d1_;
            }
            __syncthreads();

            if (role == 0) {
                const R QPp = ymsg[0][ry + 1][ix], QPm = ymsg[1][ry - 1][ix];
                const R QXp = ymsg[2][ry + 1][ix], QXm = ymsg[3][ry - 1][ix];
                const R QZp = ymsg[4][ry + 1][ix], QZm = ymsg[5][ry - 1][ix];
                constexpr R inv216 = R(1) / R(216);
                const R ps = qp[s0];
                R Ps = (R(4) * b0 + (QPp + QPm)) * inv216;
                R Xs = (R(4) * g0 + (QXp + QXm)) * inv216;
                R Ys = (QPp - QPm) * inv216;
                R Zs = (R(4) * d0 + (QZp + QZm)) * inv216;
                if (!PERIODIC) {
                    Ps += ps * om0;
                    Xs += ps * omx;
                    Ys += ps * omy;
                }
                // CCS corrector on the time-n velocity
                const R ux = qx[s0], uy = qy[s0], uz = qz[s0];
                const int cy_ = HF + iy, cx_ = HF + ix;
                const FdCell<R> xp1 = fd[buf][cy_][cx_ + D1], xm1 = fd[buf][cy_][cx_ - D1];
                const FdCell<R> yp1 = fd[buf][cy_ + D1][cx_], ym1 = fd[buf][cy_ - D1][cx_];
                const FdCell<R> xp2 = fd[buf][cy_][cx_ + 2 * D1], xm2 = fd[buf][cy_][cx_ - 2 * D1];
                const FdCell<R> yp2 = fd[buf][cy_ + 2 * D1][cx_], ym2 = fd[buf][cy_ - 2 * D1][cx_];
                constexpr int B = HF + LA;
                const int zp1 = (k + D1 - B + 2 * QL) % QL, zm1 = (k - D1 - B + 2 * QL) % QL;
                const int zp2 = (k + 2 * D1 - B + 2 * QL) % QL, zm2 = (k - 2 * D1 - B + 2 * QL) % QL;
                const R Lx = P.c0 * ux +
                             P.c1 * (xp1.x + xm1.x + yp1.x + ym1.x + qx[zp1] + qx[zm1]) +
                             P.c2 * (xp2.x + xm2.x + yp2.x + ym2.x + qx[zp2] + qx[zm2]);
                const R Ly = P.c0 * uy +
                             P.c1 * (xp1.y + xm1.y + yp1.y + ym1.y + qy[zp1] + qy[zm1]) +
                             P.c2 * (xp2.y + xm2.y + yp2.y + ym2.y + qy[zp2] + qy[zm2]);
                const R Lz = P.c0 * uz +
                             P.c1 * (xp1.z + xm1.z + yp1.z + ym1.z + qz[zp1] + qz[zm1]) +
                             P.c2 * (xp2.z + xm2.z + yp2.z + ym2.z + qz[zp2] + qz[zm2]);
                if (store) {
                    const int go = col + nxy * zo;
                    out[go] = ST::st_p(Ps - P.gauge, zo, P.q);
                    out[(long long)N + go] = ST::st_u((Xs + P.g * Lx) * P.tb + ux * P.cc, P.q);
                    out[2ll * N + go] = ST::st_u((Ys + P.g * Ly) * P.tb + uy * P.cc, P.q);
                    out[3ll * N + go] = ST::st_u((Zs + P.g * Lz) * P.tb + uz * P.cc, P.q);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------
// Software-pipelined variant for the 32-bit-arithmetic modes (FP32 and the
// 16-bit shifted-storage modes).  Tile 64 x TYP, (TYP + 3) waves:
//   waves 0 .. TYP-1   interior rows
//   waves TYP, TYP+1   halo rows y0-1, y0+TYP (+ corrector y arms)
//   wave  TYP+2        halo columns x0-1, x0+64 (+ corrector x arms)
// Each iteration t advances three planes and ends in one barrier:
//   stage A, plane a = zs+t : z pass -> x messages kept in registers
//   stage B, plane a-1      : x pass, neighbours x+-1 through DPP wave
//                             shifts (the two tile-edge lanes take the
//                             halo-column messages from LDS), y messages
//                             -> ymsg[(a-1)&1]
//   stage C, plane a-2      : y pass + corrector + blend + store
// LDS buffers are double-buffered by plane parity.  Global loads are
// unconditional buffer loads at clamped or wrapped addresses, decoded
// where used, so the queues run several planes ahead.
// ---------------------------------------------------------------------

// Queue length 12 for both supported d1; 12 is a multiple of the 3-plane
// message queue, the 4-plane halo prefetch and the 3-plane arm prefetch,
// so every slot index is static in a 12-fold unrolled loop.

// x messages used from the right neighbour (lane+1) and the left one
// (lane-1) in the x pass, and their order in the LDS edge blocks.
constexpr int XR[6] = {0, 2, 4, 5, 7, 9};
constexpr int XL[6] = {1, 3, 4, 6, 8, 9};

constexpr int DPP_WAVE_SHL1 = 0x130;  // lane i reads lane i+1
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane i reads lane i-1

// value of `src` from the neighbouring lane; lanes without a neighbour in
// the wave keep `edge`
template <int CTRL>
__device__ __forceinline__ float lane_shift(float src, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
        __builtin_bit_cast(int, edge), __builtin_bit_cast(int, src), CTRL, 0xf, 0xf, false));
}
// the same for a double: one DPP move per 32-bit half
template <int CTRL>
__device__ __forceinline__ double lane_shift(double src, double edge)
{
    typedef int i2 __attribute__((ext_vector_type(2)));
    const i2 sv = __builtin_bit_cast(i2, src), ev = __builtin_bit_cast(i2, edge);
    i2 r;
    r.x = __builtin_amdgcn_update_dpp(ev.x, sv.x, CTRL, 0xf, 0xf, false);
    r.y = __builtin_amdgcn_update_dpp(ev.y, sv.y, CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, r);
}

// Per-cell wall codes of a mask that varies along z (k_wall_codes): the
// wall rule's weights of one cell, omega = sum over its wall neighbours of
// W_e (1, e) x 216, split by z plane.  For the 3 x 3 in-plane neighbourhood
// of a cell (weights x36: corner 1, edge 4, centre 16; a neighbour outside
// the box counts as a wall) the code holds
//   bits 0-5    S0 = sum of the weights of the walls (centre included)
//   bit  6      the cell itself is a wall
//   bits 8-11   Sx + 6, Sx = sum of weight * cx over the walls
//   bits 12-15  Sy + 6
// and the kernel combines the planes c-1, c, c+1 with the z weights
// (1, 4, 1): omega0 = S0(c-1) + 4 S0(c) + S0(c+1) (the centre of plane c is
// the fluid cell itself), omegax, omegay likewise, and
// omegaz = S0(c+1) - S0(c-1).  The fields sum without carries in the bytes
// of a 32-bit word (wall_word): 152, 72 and 72 at most.
constexpr uint32_t WC_WALL = 0x40u;
__device__ __forceinline__ uint32_t wall_word(uint32_t code)
{
    return (code & 0x0f3fu) | ((code & 0xf000u) << 4) | ((code & WC_WALL) << 24);
}
constexpr uint32_t WW_WALL = WC_WALL << 24;

// One work item of the pipelined step: tile `tile` (ty * ntx + tx) over
// the z planes [zs, ze).  Called by the single-domain kernel and by the
// batched kernel; all threads of the workgroup pass the same item.
// M3: the mask varies along z; every wave reads the wall code of each cell
// it loads (wcode) instead of the column's mask byte (mask2d), and the
// interior rows build the wall weights from the codes of three planes
// instead of reading the column's (omega).
template <int MODE, bool PERIODIC, int D1, bool FD2, int TYP, bool NTS, bool M3>
__device__ __forceinline__ void pipe_item(
    const typename Store<MODE>::S *__restrict__ in,
    typename Store<MODE>::S *__restrict__ out,
    const uint8_t *__restrict__ mask2d,
    const typename Store<MODE>::R *__restrict__ omega,
    const uint16_t *__restrict__ wcode,
    const TiledParams<typename Store<MODE>::R> &P, const int tile, const int zs,
    const int ze, const uint32_t skey)
{
    using ST = Store<MODE>;
    using R = typename ST::R;
    using S = typename ST::S;
    static_assert(!(M3 && PERIODIC), "a periodic box has no mask");
    // Corrector reach: d1, and 2*d1 when the order-4 arms carry weight
    // (blend != 0).  The z queue holds the planes c-HF .. a+HF+LL of the
    // stage-C plane c = a-2 in 12 slots, so the loads can run up to
    // LA = 12 - 2*HF - 3 planes ahead (1 for reach 4, 5 for reach 2, 7 for
    // reach 1).  They run LL = min(LA, 2) ahead: fewer loads in flight per
    // wave measured faster (6% for reach 2 at 2 ahead against 5).
    constexpr int HF = FD2 ? 2 * D1 : D1;
    constexpr int QL = 12;
    constexpr int LA = QL - 2 * HF - 3;
    static_assert(LA >= 1, "queue too short for the corrector reach");
    constexpr int LL = LA < 2 ? LA : 2;
    constexpr int U = 12;
    constexpr int PD = 4;  // halo column prefetch distance (planes)
    constexpr int AD = 3;  // corrector arm prefetch distance (planes)
    constexpr int RYP = TYP + 2;
    constexpr int FX = TX + 2 * HF, FY = TYP + 2 * HF;
    constexpr int NXA = 2 * HF * TYP;          // x-arm cells
    constexpr int XPL = (NXA + 63) / 64;       // x-arm cells per lane

    __shared__ R ymsg[2][6][RYP][TX];
    __shared__ FdCell<R> fd[2][FY][FX];
    __shared__ R edge[2][RYP][2][8];  // [side 0: column x0-1 | 1: x0+64]

    const int ty = tile / P.ntx, tx = tile - ty * P.ntx;
    const int x0 = tx * TX, y0 = ty * TYP;
    const int n = ze - zs;
    const int nx = P.nx, ny = P.ny, nz = P.nz, nxy = P.nxy;
    const int N = nxy * nz;

    // State layout (StateLayout): FP32 four planes of fields, one buffer
    // resource each; the 16-bit modes two planes of pairs (p, ux) and
    // (uy, uz), where rP = rU and rV = rW and one 32-bit access moves a
    // pair.  CB: bytes per cell of one resource.
    constexpr bool PR = StateLayout<MODE>::PAIRS;
    constexpr unsigned CB = PR ? 4u : (unsigned)sizeof(S);
    typedef S S2 __attribute__((ext_vector_type(2)));
    const unsigned fbytes = (unsigned)N * CB;
    const __amdgpu_buffer_rsrc_t rP = mkbuf(in, fbytes), rV = mkbuf(in + 2ll * N, fbytes);
    const __amdgpu_buffer_rsrc_t rU = PR ? rP : mkbuf(in + N, fbytes),
                                 rW = PR ? rV : mkbuf(in + 3ll * N, fbytes);
    auto poff = [&](int zp) { return (unsigned)zp * (unsigned)nxy * CB; };
    // Quarter-rows of wall columns.  In a walled box, a row wave's quarter
    // (16 lanes: 64 B of a 32-bit field, 128 B of an FP64 one) whose
    // columns are walls in every plane reads zeros and stores
    // zeros over the zeros both buffers hold at walls (P.wskip: contexts
    // and batches keep that from the upload on).  Its loads and stores then
    // take an offset past the field, where the buffer range check returns
    // zeros for the loads (the values memory holds) and drops the stores,
    // without touching memory.  The offset plus a plane offset stays below
    // 2^32 and past the field when the field is at most 2^31 bytes.  On the
    // 256x256x1024 pipe (R 120) 17.9% of the cells of the stepped tiles lie
    // in such quarters (12.8% in all-wall halves): the FP32 step runs 4.5%
    // faster, the 16-bit ones 2-3% (half-rows: 4% in FP32).
    // With a mask that varies along z (M3), mask2d, where given, marks the
    // columns that are walls in every plane; the 16-bit M3 variants, bound
    // by vector issue, measured 0.6% slower with the skip on the rough
    // 256x256x1024 pipe (FP32 5.4% faster) and keep every access.
    constexpr unsigned OOB = 0x80000000u;
    constexpr bool SKIP = !PERIODIC && !(M3 && PR);
    const bool wskip = SKIP && P.wskip != 0 && fbytes <= OOB && mask2d != nullptr;
    // whether the quarter-row of this lane's wave holding column c (the
    // one the lane loads) is all walls; called by full waves only (wskip
    // is the same on every lane)
    auto wall_quarter = [&](unsigned c) __attribute__((always_inline)) -> bool {
        if constexpr (!SKIP) {
            (void)c;
            return false;
        } else {
            if (!wskip) return false;
            const uint64_t b = __builtin_amdgcn_ballot_w64(mask2d[c] != 0);
            const uint32_t q = (uint32_t)(b >> (16 * ((threadIdx.x & 63) >> 4))) & 0xffffu;
            return q == 0xffffu;
        }
    };
    // the same for a single cell's loads (the halo column wave's scattered
    // reads): since the neighbouring tiles skip their wall quarter-rows,
    // such a read would fetch a line for its one zero
    auto wall_cell = [&](unsigned c) __attribute__((always_inline)) -> bool {
        if constexpr (!SKIP) {
            (void)c;
            return false;
        } else {
            return wskip && mask2d[c] != 0;
        }
    };
    // fields 0, 1 (k = 0) or 2, 3 (k = 1) of the cell at column offset cb
    // and plane offset so
    auto ldpair = [&](int k, unsigned cb, unsigned so) __attribute__((always_inline)) -> S2 {
        S2 v;
        if constexpr (PR) {
            v = __builtin_bit_cast(S2, __builtin_amdgcn_raw_buffer_load_b32(k ? rV : rP, cb, so, 0));
        } else {
            v.x = bload<S>(k ? rV : rP, cb, so);
            v.y = bload<S>(k ? rW : rU, cb, so);
        }
        return v;
    };
    // Exact half-to-float of the low (HI false) or high half of a 32-bit
    // word by integer operations: the half moved to the top of a word and
    // shifted right by 3 with sign extension puts its exponent and
    // mantissa in place for a float and its sign in bits 31-28; clearing
    // bits 30-28 (and, for the high half, the low half's bits) leaves a
    // float of the half's value times 2^-112 (the exponent bias
    // difference; exact for normal and subnormal halves, as float
    // denormals are preserved).  Two operations for the high half, three
    // for the low one.  The conversion instructions (v_cvt_f32_f16,
    // v_fma_mix_f32) need hazard padding and cost more than these
    // operations: the FP16-storage step measured 8% faster without them.
    auto h2bits = [](uint32_t w, auto hi) __attribute__((always_inline)) -> float {
        constexpr bool HI = decltype(hi)::value;
        const uint32_t bits = HI ? (uint32_t)((int32_t)w >> 3) & 0x8fffe000u
                                 : (uint32_t)((int32_t)(w << 16) >> 3) & 0x8fffffffu;
        return __builtin_bit_cast(float, bits);
    };
    auto h2f = [&](uint32_t w, auto hi) __attribute__((always_inline)) -> float {
        return h2bits(w, hi) * 0x1p112f;
    };
    // the same plus `add` in one fused multiply-add (the product is exact,
    // so the result is the sum rounded once, as h2f(w) + add)
    auto h2f_add = [&](uint32_t w, auto hi, float add) __attribute__((always_inline)) -> float {
        return __builtin_fmaf(h2bits(w, hi), 0x1p112f, add);
    };
    // a stored FP16 pressure, base + sp * value: the 2^112 of the decode
    // joins the power-of-two scale sp (an exact product), so the value is
    // one fused multiply-add of the decoded bits, rounded once as before
    const float sp112 = P.q.sp * 0x1p112f;
    auto h2p = [&](uint32_t w, float base) __attribute__((always_inline)) -> float {
        return __builtin_fmaf(sp112, h2bits(w, std::false_type{}), base);
    };
    using LO_ = std::integral_constant<bool, false>;
    using HI_ = std::integral_constant<bool, true>;
    // A corrector arm cell (ux, uy, uz) as the loads bring it: three
    // fields (FP32) or the pairs (p, ux) and (uy, uz) (16-bit modes).
    using AV = std::conditional_t<PR, S2, S>;
    constexpr int NA = PR ? 2 : 3;
    auto arm_load = [&](AV (&v)[NA], unsigned cb, unsigned so) __attribute__((always_inline)) {
        if constexpr (PR) {
            v[0] = ldpair(0, cb, so);
            v[1] = ldpair(1, cb, so);
        } else {
            v[0] = bload<S>(rU, cb, so);
            v[1] = bload<S>(rV, cb, so);
            v[2] = bload<S>(rW, cb, so);
        }
    };
    // RAWZ: the corrector tile and the decoded queue hold uz's departure
    // from the reference, and the interior adds the references of the arms
    // (constant along z) in one term.  For BF16 storage the decode has no
    // multiply to fuse the reference into, and the adds on the halo waves'
    // arm path cost 5% of the step; FP16 storage fuses them (h2f_add).
    constexpr bool RAWZ = MODE == FSLBM_BF16S;
    // streamwise reference of column c (16-bit modes; 0 in FP32)
    auto uref_of = [&](unsigned c) __attribute__((always_inline)) -> R {
        if constexpr (PR) return P.q.uref ? P.q.uref[c] : R(0);
        else return R(0);
    };
    // ur: the cell's column's streamwise reference (16-bit modes)
    auto arm_cell = [&](const AV (&v)[NA], R ur) __attribute__((always_inline)) -> FdCell<R> {
        FdCell<R> cc;
        if constexpr (MODE == FSLBM_F16S) {
            cc.x = h2f(__builtin_bit_cast(uint32_t, v[0]), HI_{});
            cc.y = h2f(__builtin_bit_cast(uint32_t, v[1]), LO_{});
            cc.z = h2f_add(__builtin_bit_cast(uint32_t, v[1]), HI_{}, ur);
        } else if constexpr (PR) {
            cc.x = ST::ld_u(v[0].y, P.q);
            cc.y = ST::ld_u(v[1].x, P.q);
            cc.z = ST::ld_u(v[1].y, P.q);  // BF16: the departure (RAWZ)
            (void)ur;
        } else {
            cc.x = ST::ld_u(v[0], P.q);
            cc.y = ST::ld_u(v[1], P.q);
            cc.z = ST::ld_u(v[2], P.q);
        }
        return cc;
    };

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;

    // The physical plane of plane zl (z wraps) and the pressure jump of the
    // wrap.  zl is the same on every lane: the jump is selected as bits,
    // which keeps the selects on the scalar unit.
    using RB = std::conditional_t<sizeof(R) == 8, uint64_t, uint32_t>;  // R's bits
    auto phys = [&](int zl, R &jump) -> int {
        const RB pb = __builtin_bit_cast(RB, P.pin);
        const RB sign = RB(1) << (8 * sizeof(R) - 1);
        const RB jb = zl < 0 ? pb : zl >= nz ? pb ^ sign : RB(0);
        jump = __builtin_bit_cast(R, jb);
        return zl < 0 ? zl + nz : zl >= nz ? zl - nz : zl;
    };
    auto colx = [&](int v) { return PERIODIC ? modp(v, nx) : clampi(v, nx); };
    auto coly = [&](int v) { return PERIODIC ? modp(v, ny) : clampi(v, ny); };
    // wall code (M3) of column `col` in plane zl (wrapped): 2 bytes per
    // cell, below 2^32 bytes as N < 2^31
    const __amdgpu_buffer_rsrc_t rC = mkbuf(wcode, M3 ? (unsigned)N * 2u : 0u);
    auto cload = [&](unsigned col, int zl) __attribute__((always_inline)) -> uint32_t {
        R jump;
        const unsigned so = (unsigned)phys(zl, jump) * (unsigned)nxy * 2u;
        return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rC, col * 2u, so, 0);
    };
    // Wall cells hold zero velocity in every storage mode, so only their
    // pressure (the decoded reference line and the jump) needs masking for
    // the whole message to vanish.
    // (ur: the column's streamwise reference, 16-bit modes)
    auto msg_of = [&](S rp, S ru, S rv, S rw, int zl, bool fl, R ur) -> Msg<R> {
        R jump;
        const int zp = phys(zl, jump);
        if constexpr (MODE == FSLBM_F16S) {
            auto w = [](S v) { return (uint32_t)__builtin_bit_cast(uint16_t, v); };
            const R pp = fl ? h2p(w(rp), P.q.pbase + P.q.pslope * (float)zp) + jump : R(0);
            return make_msg<R>(pp, h2f(w(ru), LO_{}), h2f(w(rv), LO_{}), h2f_add(w(rw), LO_{}, ur));
        } else {
            const R pp = fl ? ST::ld_p(rp, zp, P.q) + jump : R(0);
            R uz = ST::ld_u(rw, P.q);
            if constexpr (PR) uz += ur;
            const R ux = ST::ld_u(ru, P.q), uy = ST::ld_u(rv, P.q);
            return make_msg<R>(pp, ux, uy, uz);
        }
    };
    // x pass of a row wave: neighbours by DPP, edge lanes from LDS
    auto xpass = [&](int buf, int ry, int ix, const R (&xv)[10], const R (&cen)[6],
                     R &b0, R &g0, R &d0) {
        R le[6], re[6];
#pragma unroll
        for (int j = 0; j < 6; j++) {
            le[j] = edge[buf][ry][0][j];
            re[j] = edge[buf][ry][1][j];
        }
        R e[6], wv[6];
#pragma unroll
        for (int j = 0; j < 6; j++) {
            e[j] = lane_shift<DPP_WAVE_SHL1>(xv[XR[j]], re[j]);
            wv[j] = lane_shift<DPP_WAVE_SHR1>(xv[XL[j]], le[j]);
        }
        // e[] = values 0,2,4,5,7,9 of x+1; wv[] = values 1,3,4,6,8,9 of x-1
        b0 = R(4) * cen[0] + (e[0] + wv[0]);
        const R b1 = R(4) * cen[1] + (e[1] + wv[1]);
        const R b2 = R(4) * cen[2] + (e[2] + wv[2]);
        g0 = e[0] - wv[0];
        const R g1 = e[1] - wv[1];
        const R g2 = e[2] - wv[2];
        d0 = R(4) * cen[3] + (e[3] + wv[3]);
        const R d1_ = R(4) * cen[4] + (e[4] + wv[4]);
        const R d2 = R(4) * cen[5] + (e[5] + wv[5]);
        const R tb_ = b0 + b2, tg = g0 + g2, td = d0 + d2;
        ymsg[buf][0][ry][ix] = tb_ + b1;
        ymsg[buf][1][ry][ix] = tb_ - b1;
        ymsg[buf][2][ry][ix] = tg + g1;
        ymsg[buf][3][ry][ix] = tg - g1;
        ymsg[buf][4][ry][ix] = td + d1_;
        ymsg[buf][5][ry][ix] = td - d1_;
    };
    auto keep = [&](const XMsg<R> &xm, R (&xv)[10], R (&cen)[6]) {
#pragma unroll
        for (int v = 0; v < 10; v++) xv[v] = xm.v[v];
        cen[0] = xm.cA;
        cen[1] = xm.cC;
        cen[2] = xm.v[4];
        cen[3] = xm.dA;
        cen[4] = xm.dC;
        cen[5] = xm.v[9];
    };

    // Iterations t = 0 .. n+1.  t = 0 and 1 are peeled (stages B, C not yet
    // active); from t = 2 on every stage runs: the stage A / B work of
    // t = n, n+1 concerns planes past the chunk and only touches registers
    // and LDS buffers that are not read again.  Keeping the main loop free
    // of branches around memory instructions keeps the memory counters
    // exact, so queue loads stay in flight for several iterations.
    const int nit = n + 2;

    if (wave < TYP) {
        // ================= interior rows =================
        const int ix = lane, iy = wave;
        const int x = x0 + ix, y = y0 + iy;
        const int ry = iy + 1;
        const bool inside = x < nx && y < ny;
        const unsigned col = colx(x) + nx * coly(y);
        // M3: `fluid` is the column part of the flag (inside the box); each
        // plane's own flag adds its wall code's
        const bool fluid = PERIODIC ? true : M3 ? inside : (inside && mask2d[col] == 0);
        R om0 = 0, omx = 0, omy = 0;
        if (!PERIODIC && !M3) {
            om0 = omega[col];
            omx = omega[nxy + col];
            omy = omega[2 * nxy + col];
        }
        // Wall codes (M3): raw codes of the planes a+1 and a+2 in
        // cq[(q - zs) & 1], loaded one iteration ahead of the message of
        // their plane; wall words of the planes a-3 .. a in ww[(q - zs) & 3]
        uint32_t cq[2] = {0u, 0u}, ww[4] = {0u, 0u, 0u, 0u};
        auto flc = [&](uint32_t code) __attribute__((always_inline)) {
            return M3 ? fluid && !(code & WC_WALL) : fluid;
        };
        const R urc = uref_of(col);  // the column's streamwise reference
        // The stored uz' = uz - urc is folded into the update's first sum:
        // nz_ = uz ku + S1 ks1 (+ S2 ks2) + ... - urc.  With RAWZ, uz and the
        // arm sums S1, S2 are sums of departures, and the references they
        // leave out (urc ku, and ks1, ks2 times the arms' references: four
        // in-plane columns at distance d1 (2 d1) and twice urc for the z
        // arms) join the constant.
        R czr = -urc;
        if constexpr (RAWZ) {
            auto arms = [&](int d) {
                return (uref_of(colx(x + d) + nx * coly(y)) + uref_of(colx(x - d) + nx * coly(y))) +
                       (uref_of(colx(x) + nx * coly(y + d)) + uref_of(colx(x) + nx * coly(y - d))) +
                       R(2) * urc;
            };
            czr = urc * P.ku - urc + arms(D1) * P.ks1;
            if constexpr (FD2) czr += arms(2 * D1) * P.ks2;
        }
        // Lanes that own no fluid cell store 0 to their (wall) column: in
        // a walled box an outside lane's clamped column is a wall.  In a
        // periodic box an outside lane computes a wrapped copy of an inside
        // cell, but along another path (its neighbours come from other
        // lanes and waves), so its result may differ in the last bit: its
        // store offset is moved past the end of the buffer, where the
        // buffer range check drops the write.  (offset + plane offset
        // stays below 2^32: a periodic box's fields are below 2^31 bytes,
        // tiled_supported.)
        const bool store = PERIODIC ? true : (inside && fluid);
        // stochastic-rounding salt of the column (16-bit modes)
        const uint32_t salt = ST::SR ? sr_salt(col) : 0u;
        const unsigned cb_ = wall_quarter(col) ? OOB : col * CB;
        const unsigned sb_ = (PERIODIC && !inside) ? OOB : cb_;
        const __amdgpu_buffer_rsrc_t oP = mkbuf(out, fbytes), oV = mkbuf(out + 2ll * N, fbytes);
        const __amdgpu_buffer_rsrc_t oU = PR ? oP : mkbuf(out + N, fbytes),
                                     oW = PR ? oV : mkbuf(out + 3ll * N, fbytes);
        // Queue of raw values, two fields per entry, (p, ux) and (uy, uz):
        // plane q in slot (q - zs + 2 + HF) mod QR.  FP32 reads every value
        // from it where used (QR = 12).  The 16-bit modes convert each
        // stored value once: plane a+DA is decoded at iteration a into
        // the decoded queue dq (slot (q - zs + 2 + HF) mod DQ), which then
        // serves the message, the corrector tile, the own values and the z
        // arms (a half-to-float conversion costs several issue slots, and
        // read where used each value was converted up to five times).  Its
        // raw queue holds only the planes in flight: at most
        // HF + LL - DA + 1 <= 4 of them, so a load never overwrites a plane
        // not yet decoded.  Both lengths divide the 12-fold unroll, so every
        // slot index is a constant.
        constexpr int QR = PR ? 4 : QL;
        constexpr int DA = PR ? (HF - 2 > 1 ? HF - 2 : 1) : 0;
        constexpr int DQ = PR ? (HF + DA + 3 <= 6 ? 6 : 12) : 1;
        static_assert(!PR || HF + LL - DA <= 3, "raw queue too short");
        S2 qa[QR], qb[QR];
        R dq[DQ][4];
        auto qload = [&](int q, int s) __attribute__((always_inline)) {
            R jump;
            const unsigned so = poff(phys(q, jump));
            qa[s] = ldpair(0, cb_, so);
            qb[s] = ldpair(1, cb_, so);
        };
        // decode plane q from raw slot s into decoded slot d (pressure
        // without the jump of a wrapped plane: msg_dec adds it)
        auto qdecode = [&](int q, int s, int d) __attribute__((always_inline)) {
            R jump;
            const int zp = phys(q, jump);
            if constexpr (MODE == FSLBM_F16S) {
                const uint32_t wa = __builtin_bit_cast(uint32_t, qa[s]),
                               wb = __builtin_bit_cast(uint32_t, qb[s]);
                dq[d][0] = h2p(wa, P.q.pbase + P.q.pslope * (float)zp);
                dq[d][1] = h2f(wa, HI_{});
                dq[d][2] = h2f(wb, LO_{});
                dq[d][3] = h2f_add(wb, HI_{}, urc);
            } else {
                dq[d][0] = ST::ld_p(qa[s].x, zp, P.q);
                dq[d][1] = ST::ld_u(qa[s].y, P.q);
                dq[d][2] = ST::ld_u(qb[s].x, P.q);
                dq[d][3] = RAWZ ? ST::ld_u(qb[s].y, P.q) : ST::ld_u(qb[s].y, P.q) + urc;
            }
        };
        auto msg_dec = [&](int d, int zl, bool fl) __attribute__((always_inline)) -> Msg<R> {
            R jump;
            (void)phys(zl, jump);
            const R pp = fl ? dq[d][0] + jump : R(0);
            return make_msg<R>(pp, dq[d][1], dq[d][2], RAWZ ? dq[d][3] + urc : dq[d][3]);
        };
        bool flm = fluid, fl0 = fluid;  // planes zs-1 and zs
        if constexpr (M3) {
            const uint32_t cm = cload(col, zs - 1), c0 = cload(col, zs);
            cq[1] = cload(col, zs + 1);
            ww[3] = wall_word(cm);
            ww[0] = wall_word(c0);
            flm = flc(cm);
            fl0 = flc(c0);
        }
        Msg<R> mq[3];
        if constexpr (PR) {
            // planes zs-HF .. zs+DA-1 decoded before the loop, the rest of
            // the look-ahead in flight
#pragma unroll
            for (int j = 2; j < HF + 2 + DA; j++) {
                qload(zs - 2 - HF + j, 0);
                qdecode(zs - 2 - HF + j, 0, j % DQ);
            }
#pragma unroll
            for (int j = HF + 2 + DA; j < 2 * HF + 2 + LL; j++) qload(zs - 2 - HF + j, j % QR);
            mq[0] = msg_dec((HF + 1) % DQ, zs - 1, flm);
            mq[1] = msg_dec((HF + 2) % DQ, zs, fl0);
        } else {
#pragma unroll
            for (int j = 2; j < 2 * HF + 2 + LL; j++) qload(zs - 2 - HF + j, j);
            mq[0] = msg_of(qa[HF + 1].x, qa[HF + 1].y, qb[HF + 1].x, qb[HF + 1].y, zs - 1, flm, urc);
            mq[1] = msg_of(qa[HF + 2].x, qa[HF + 2].y, qb[HF + 2].x, qb[HF + 2].y, zs, fl0, urc);
        }
        // the last plane any stage of this chunk reads: corrector arm of
        // the last output plane ze-1
        const int qlast = ze - 1 + HF;
        mq[2] = zero_msg<R>();
        R xv[10], cen[6];
#pragma unroll
        for (int v = 0; v < 10; v++) xv[v] = 0;
#pragma unroll
        for (int v = 0; v < 6; v++) cen[v] = 0;
        R b0c = 0, g0c = 0, d0c = 0;

        // one iteration; k = t mod 12 (a constant after unrolling)
        auto iter = [&](int t, int k, bool doB, bool doC) __attribute__((always_inline)) {
            const int a = zs + t;
            // stage C's LDS operands first: their latency overlaps the
            // message work of stage A
            R QPp = 0, QPm = 0, QXp = 0, QXm = 0, QZp = 0, QZm = 0;
            FdCell<R> xp1{}, xm1{}, yp1{}, ym1{};
            if (doC) {
                const int cb = (a - 2) & 1;
                QPp = ymsg[cb][0][ry + 1][ix]; QPm = ymsg[cb][1][ry - 1][ix];
                QXp = ymsg[cb][2][ry + 1][ix]; QXm = ymsg[cb][3][ry - 1][ix];
                QZp = ymsg[cb][4][ry + 1][ix]; QZm = ymsg[cb][5][ry - 1][ix];
                const int cy_ = HF + iy, cx_ = HF + ix;
                xp1 = fd[cb][cy_][cx_ + D1]; xm1 = fd[cb][cy_][cx_ - D1];
                yp1 = fd[cb][cy_ + D1][cx_]; ym1 = fd[cb][cy_ - D1][cx_];
            }
            __builtin_amdgcn_sched_barrier(0);
            // The wall code of plane a+2 is loaded before this iteration's
            // state loads, so the wait for it next iteration leaves those
            // in flight.  Codes are read up to plane ze, the last plane a
            // result of the chunk depends on.
            if constexpr (M3) cq[k & 1] = cload(col, a + 2 < ze ? a + 2 : ze);
            // Planes past qlast feed only results past the chunk: their
            // loads re-read plane qlast (a cache hit) instead of fetching
            // planes the next chunk reads again.
            qload(a + HF + LL < qlast ? a + HF + LL : qlast, (k + 2 * HF + 2 + LL) % QR);
            const bool fl1 = flc(cq[(k + 1) & 1]);  // plane a+1
            if constexpr (PR) {
                qdecode(a + DA, (k + DA + 2 + HF) % QR, (k + DA + 2 + HF) % DQ);
                mq[(k + 2) % 3] = msg_dec((k + 3 + HF) % DQ, a + 1, fl1);
            } else {
                const int s1 = (k + 3 + HF) % QL;  // plane a+1
                mq[(k + 2) % 3] = msg_of(qa[s1].x, qa[s1].y, qb[s1].x, qb[s1].y, a + 1, fl1, urc);
            }
            __builtin_amdgcn_sched_barrier(0);
            if (doC) {  // stage C: plane c = a-2
                const int c = a - 2;
                const int cb = c & 1;
                const int sc = (k + HF) % QR;
                const int dc = (k + HF) % DQ;
                constexpr R inv216 = R(1) / R(216);
                // velocity of field f (1..3) of the plane in queue
                // position `off` from c (a constant after unrolling)
                auto U = [&](int f, int off) __attribute__((always_inline)) -> R {
                    if constexpr (PR) {
                        return dq[(k + HF + off + DQ) % DQ][f];
                    } else {
                        const int sl = (k + HF + off + QL) % QL;
                        return ST::ld_u(f == 1 ? qa[sl].y : f == 2 ? qb[sl].x : qb[sl].y, P.q);
                    }
                };
                const R ux = U(1, 0), uy = U(2, 0), uz = U(3, 0);
                // unnormalised predictor sums (x216)
                const R Ps = R(4) * b0c + (QPp + QPm);
                const R Xs = R(4) * g0c + (QXp + QXm);
                const R Ys = QPp - QPm;
                const R Zs = R(4) * d0c + (QZp + QZm);
                // wall rule terms: own pressure times the wall weights
                R w0 = 0, wx = 0, wy = 0, wz = 0;
                bool st_c = store;
                if (!PERIODIC) {
                    R pc;
                    if constexpr (PR) pc = dq[dc][0];
                    else pc = ST::ld_p(qa[sc].x, c, P.q);
                    if constexpr (M3) {
                        // wall words of the planes c-1, c, c+1
                        const uint32_t Wm = ww[(k + 1) & 3], Wc = ww[(k + 2) & 3],
                                       Wp = ww[(k + 3) & 3];
                        st_c = fluid && !(Wc & WW_WALL);
                        const uint32_t T = Wm + Wp + 4u * Wc;
                        const R pw = (st_c ? pc : R(0)) * inv216;
                        w0 = pw * (R)(T & 0xffu);
                        wx = pw * ((R)((T >> 8) & 0xffu) - R(36));
                        wy = pw * ((R)((T >> 16) & 0xffu) - R(36));
                        wz = pw * ((R)(Wp & 0xffu) - (R)(Wm & 0xffu));
                    } else {
                        const R ps = fluid ? pc : R(0);
                        w0 = ps * om0;
                        wx = ps * omx;
                        wy = ps * omy;
                    }
                }
                const int cy_ = HF + iy, cx_ = HF + ix;
                // corrector arm sums at distance d1 (and 2*d1)
                const R s1x = xp1.x + xm1.x + yp1.x + ym1.x + U(1, D1) + U(1, -D1);
                const R s1y = xp1.y + xm1.y + yp1.y + ym1.y + U(2, D1) + U(2, -D1);
                const R s1z = xp1.z + xm1.z + yp1.z + ym1.z + U(3, D1) + U(3, -D1);
                // u' = tb (u*/216 + wall + g (c0 u + c1 S1 + c2 S2)) + (1 - tb) u
                R nx_ = ux * P.ku + s1x * P.ks1;
                R ny_ = uy * P.ku + s1y * P.ks1;
                // (the 16-bit modes store uz - uref: folded into this sum)
                R nz_ = PR ? uz * P.ku + (s1z * P.ks1 + czr) : uz * P.ku + s1z * P.ks1;
                if constexpr (FD2) {  // order-4 arms (blend != 0)
                    const FdCell<R> xp2 = fd[cb][cy_][cx_ + 2 * D1], xm2 = fd[cb][cy_][cx_ - 2 * D1];
                    const FdCell<R> yp2 = fd[cb][cy_ + 2 * D1][cx_], ym2 = fd[cb][cy_ - 2 * D1][cx_];
                    nx_ += P.ks2 * (xp2.x + xm2.x + yp2.x + ym2.x + U(1, 2 * D1) + U(1, -2 * D1));
                    ny_ += P.ks2 * (xp2.y + xm2.y + yp2.y + ym2.y + U(2, 2 * D1) + U(2, -2 * D1));
                    nz_ += P.ks2 * (xp2.z + xm2.z + yp2.z + ym2.z + U(3, 2 * D1) + U(3, -2 * D1));
                }
                nx_ += P.tb * wx;
                ny_ += P.tb * wy;
                if constexpr (M3) nz_ += P.tb * wz;
                nx_ += Xs * P.kx;
                ny_ += Ys * P.kx;
                nz_ += Zs * P.kx;
                const R np_ = (Ps * inv216 + w0) - P.gauge;
                const unsigned so = poff(c);
                const S z0 = S(0);
                const SrWords rw = ST::SR ? sr_words(salt, sr_plane(skey, (uint32_t)c))
                                          : SrWords{{0u, 0u, 0u, 0u}};
                // Encoded on every lane, then selected: written as
                // `store ? encode(..) : 0`, the stochastic-rounding
                // conversions (which the compiler does not speculate)
                // became four exec-masked branches per plane that broke
                // the loop's single scheduling region.
                if constexpr (PR) {  // two packed pairs, one store each
                    const uint32_t A = ST::sr2(ST::enc_p(np_, c, P.q), rw.r[0], nx_, rw.r[1]);
                    const uint32_t B = ST::sr2(ny_, rw.r[2], nz_, rw.r[3]);
                    __builtin_amdgcn_raw_buffer_store_b32(st_c ? A : 0u, oP, sb_, so, 0);
                    __builtin_amdgcn_raw_buffer_store_b32(st_c ? B : 0u, oV, sb_, so, 0);
                } else {
                    const S ep = ST::st_p_sr(np_, c, P.q, rw.r[0]);
                    const S ex = ST::st_u_sr(nx_, P.q, rw.r[1]);
                    const S ey = ST::st_u_sr(ny_, P.q, rw.r[2]);
                    const S ez = ST::st_u_sr(nz_, P.q, rw.r[3]);
                    bstore<NTS>(st_c ? ep : z0, oP, sb_, so);
                    bstore<NTS>(st_c ? ex : z0, oU, sb_, so);
                    bstore<NTS>(st_c ? ey : z0, oV, sb_, so);
                    bstore<NTS>(st_c ? ez : z0, oW, sb_, so);
                }
            }
            // wall word of plane a+1 (its slot held plane a-3, last read
            // by stage C above)
            if constexpr (M3) ww[(k + 1) & 3] = wall_word(cq[(k + 1) & 1]);
            if (doB) {  // stage B: plane a-1; corrector tile of a-1
                xpass((a - 1) & 1, ry, ix, xv, cen, b0c, g0c, d0c);
                FdCell<R> cc;
                if constexpr (PR) {
                    const int db = (k + 1 + HF) % DQ;
                    cc.x = dq[db][1];
                    cc.y = dq[db][2];
                    cc.z = dq[db][3];
                } else {
                    const int sb = (k + 1 + HF) % QL;
                    cc.x = ST::ld_u(qa[sb].y, P.q);
                    cc.y = ST::ld_u(qb[sb].x, P.q);
                    cc.z = ST::ld_u(qb[sb].y, P.q);
                }
                fd[(a - 1) & 1][HF + iy][HF + ix] = cc;
            }
            // stage A: plane a
            keep(zpass<R>(mq[k % 3], mq[(k + 1) % 3], mq[(k + 2) % 3]), xv, cen);
            __syncthreads();
        };
        iter(0, 0, false, false);
        iter(1, 1, true, false);
        for (int ib = 2; ib < nit; ib += U) {
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int t = ib + k;
                if (t >= nit) break;
                iter(t, (k + 2) % U, true, true);
            }
        }
    } else if (wave < TYP + 2) {
        // ================= halo rows y0-1 and y0+TYP =================
        const int hr = wave - TYP;
        const int ix = lane;
        const int x = x0 + ix;
        const int y = hr ? y0 + TYP : y0 - 1;
        const int ry = hr ? TYP + 1 : 0;
        const bool inside = x < nx && y >= 0 && y < ny;
        const unsigned col = colx(x) + nx * coly(y);
        const bool fluid = PERIODIC ? true : M3 ? inside : (inside && mask2d[col] == 0);
        const unsigned cb_ = wall_quarter(col) ? OOB : col * CB;
        // own column prefetch: plane m in slot (m - zs + 1) mod PD
        S2 ha[PD], hb[PD];
        uint32_t hc[PD];  // wall codes (M3)
        auto hload = [&](int m, int s) __attribute__((always_inline)) {
            R jump;
            const unsigned so = poff(phys(m, jump));
            if constexpr (M3) hc[s] = cload(col, m);
            ha[s] = ldpair(0, cb_, so);
            hb[s] = ldpair(1, cb_, so);
        };
        auto hfl = [&](int s) __attribute__((always_inline)) {
            return M3 ? fluid && !(hc[s] & WC_WALL) : fluid;
        };
        const R hur = uref_of(col);  // the column's streamwise reference
        // corrector y arms: HF rows of this side, column ix; plane q in
        // slot (q - zs + 1) mod AD
        unsigned acol[HF];
        R aref[HF];  // their columns' streamwise references
#pragma unroll
        for (int r = 0; r < HF; r++) {
            const int fy = hr ? TYP + HF + r : r;
            const unsigned ac = colx(x0 + ix) + nx * coly(y0 - HF + fy);
            acol[r] = wall_quarter(ac) ? OOB : ac * CB;
            aref[r] = uref_of(ac);
        }
        AV ar[AD][HF][NA];
        auto aload = [&](int q, int s) __attribute__((always_inline)) {
            const unsigned so = poff(wrapi(q, nz));
#pragma unroll
            for (int r = 0; r < HF; r++) arm_load(ar[s][r], acol[r], so);
        };
#pragma unroll
        for (int j = 0; j < PD; j++) hload(zs - 1 + j, j);
#pragma unroll
        for (int j = 0; j < AD; j++) aload(zs - 1 + j, j);
        Msg<R> mq[3];
        mq[0] = msg_of(ha[0].x, ha[0].y, hb[0].x, hb[0].y, zs - 1, hfl(0), hur);
        mq[1] = msg_of(ha[1].x, ha[1].y, hb[1].x, hb[1].y, zs, hfl(1), hur);
        mq[2] = zero_msg<R>();
        hload(zs - 1 + PD, 0);
        hload(zs + PD, 1);
        R xv[10], cen[6];
#pragma unroll
        for (int v = 0; v < 10; v++) xv[v] = 0;
#pragma unroll
        for (int v = 0; v < 6; v++) cen[v] = 0;
        R b0c, g0c, d0c;

        auto iter = [&](int t, int k, bool doB) __attribute__((always_inline)) {
            const int a = zs + t;
            {
                const int s1 = (k + 2) % PD;  // plane a+1
                mq[(k + 2) % 3] = msg_of(ha[s1].x, ha[s1].y, hb[s1].x, hb[s1].y, a + 1, hfl(s1), hur);
                // messages past plane ze and arms past ze-1 feed only
                // results past the chunk: re-read those planes instead
                hload(a + 1 + PD < ze ? a + 1 + PD : ze, s1);
            }
            if (doB) {
                xpass((a - 1) & 1, ry, ix, xv, cen, b0c, g0c, d0c);
                const int sa = k % AD;  // plane a-1
#pragma unroll
                for (int r = 0; r < HF; r++) {
                    const int fy = hr ? TYP + HF + r : r;
                    fd[(a - 1) & 1][fy][HF + ix] = arm_cell(ar[sa][r], aref[r]);
                }
            }
            aload(a + 2 < ze - 1 ? a + 2 : ze - 1, k % AD);
            keep(zpass<R>(mq[k % 3], mq[(k + 1) % 3], mq[(k + 2) % 3]), xv, cen);
            __syncthreads();
        };
        iter(0, 0, false);
        iter(1, 1, true);
        for (int ib = 2; ib < nit; ib += U) {
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int t = ib + k;
                if (t >= nit) break;
                iter(t, (k + 2) % U, true);
            }
        }
    } else {
        // ===== halo columns x0-1, x0+64 (lanes < 2*RYP) and the x arms =====
        const int sl = lane < 2 * RYP ? lane : 2 * RYP - 1;
        const int side = sl / RYP, r = sl - side * RYP;
        const int x = side ? x0 + TX : x0 - 1;
        const int y = y0 - 1 + r;
        const int ry = r;
        const bool inside = x >= 0 && x < nx && y >= 0 && y < ny;
        const unsigned col = colx(x) + nx * coly(y);
        const bool fluid = PERIODIC ? true : M3 ? inside : (inside && mask2d[col] == 0);
        const bool writer = lane < 2 * RYP;
        const unsigned cb_ = wall_cell(col) ? OOB : col * CB;
        S2 ha[PD], hb[PD];
        uint32_t hc[PD];  // wall codes (M3)
        auto hload = [&](int m, int s) __attribute__((always_inline)) {
            R jump;
            const unsigned so = poff(phys(m, jump));
            if constexpr (M3) hc[s] = cload(col, m);
            ha[s] = ldpair(0, cb_, so);
            hb[s] = ldpair(1, cb_, so);
        };
        auto hfl = [&](int s) __attribute__((always_inline)) {
            return M3 ? fluid && !(hc[s] & WC_WALL) : fluid;
        };
        const R hur = uref_of(col);  // the column's streamwise reference
        // x arms: NXA cells, XPL per lane
        int afy[XPL], afx[XPL];
        bool aw[XPL];
        unsigned acol[XPL];
        R aref[XPL];  // their columns' streamwise references
#pragma unroll
        for (int j = 0; j < XPL; j++) {
            const int cell = lane + 64 * j;
            aw[j] = cell < NXA;
            const int cl = aw[j] ? cell : NXA - 1;
            const int br = cl / (2 * HF), bc = cl - br * (2 * HF);
            afy[j] = HF + br;
            afx[j] = bc < HF ? bc : TX + bc;
            const unsigned ac = colx(x0 - HF + afx[j]) + nx * coly(y0 - HF + afy[j]);
            acol[j] = wall_cell(ac) ? OOB : ac * CB;
            aref[j] = uref_of(ac);
        }
        AV ar[AD][XPL][NA];
        auto aload = [&](int q, int s) __attribute__((always_inline)) {
            const unsigned so = poff(wrapi(q, nz));
#pragma unroll
            for (int j = 0; j < XPL; j++) arm_load(ar[s][j], acol[j], so);
        };
#pragma unroll
        for (int j = 0; j < PD; j++) hload(zs - 1 + j, j);
#pragma unroll
        for (int j = 0; j < AD; j++) aload(zs - 1 + j, j);
        Msg<R> mq[3];
        mq[0] = msg_of(ha[0].x, ha[0].y, hb[0].x, hb[0].y, zs - 1, hfl(0), hur);
        mq[1] = msg_of(ha[1].x, ha[1].y, hb[1].x, hb[1].y, zs, hfl(1), hur);
        mq[2] = zero_msg<R>();
        hload(zs - 1 + PD, 0);
        hload(zs + PD, 1);

        auto iter = [&](int t, int k, bool doB) __attribute__((always_inline)) {
            const int a = zs + t;
            {
                const int s1 = (k + 2) % PD;
                mq[(k + 2) % 3] = msg_of(ha[s1].x, ha[s1].y, hb[s1].x, hb[s1].y, a + 1, hfl(s1), hur);
                // messages past plane ze and arms past ze-1 feed only
                // results past the chunk: re-read those planes instead
                hload(a + 1 + PD < ze ? a + 1 + PD : ze, s1);
            }
            if (doB) {
                const int sa = k % AD;
#pragma unroll
                for (int j = 0; j < XPL; j++) {
                    if (!aw[j]) continue;
                    fd[(a - 1) & 1][afy[j]][afx[j]] = arm_cell(ar[sa][j], aref[j]);
                }
            }
            aload(a + 2 < ze - 1 ? a + 2 : ze - 1, k % AD);
            if (writer) {
                // edge block of this column: the values its neighbour lane
                // reads (left column -> XL, right column -> XR); two
                // branches with static indexes (a select between array
                // elements would become a dynamic index)
                const XMsg<R> xm = zpass<R>(mq[k % 3], mq[(k + 1) % 3], mq[(k + 2) % 3]);
                if (side) {
#pragma unroll
                    for (int j = 0; j < 6; j++) edge[a & 1][ry][1][j] = xm.v[XR[j]];
                } else {
#pragma unroll
                    for (int j = 0; j < 6; j++) edge[a & 1][ry][0][j] = xm.v[XL[j]];
                }
            }
            __syncthreads();
        };
        iter(0, 0, false);
        iter(1, 1, true);
        for (int ib = 2; ib < nit; ib += U) {
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int t = ib + k;
                if (t >= nit) break;
                iter(t, (k + 2) % U, true);
            }
        }
    }
}

template <int MODE, bool PERIODIC, int D1, bool FD2, int TYP, bool NTS, bool M3>
__global__ __launch_bounds__((TYP + 3) * 64) void k_step_pipe(
    const typename Store<MODE>::S *__restrict__ in,
    typename Store<MODE>::S *__restrict__ out,
    const uint8_t *__restrict__ mask2d,
    const typename Store<MODE>::R *__restrict__ omega,
    const uint16_t *__restrict__ wcode,
    TiledParams<typename Store<MODE>::R> P)
{
    // XCD-major work order: the 8 XCDs each take a contiguous run
    const int bid = blockIdx.x;
    const int w = (bid & 7) * P.per_xcd + (bid >> 3);
    if (w >= P.nwork) return;
    int tile, zs, ze;
    item_of(P, w, tile, zs, ze);
    pipe_item<MODE, PERIODIC, D1, FD2, TYP, NTS, M3>(in, out, mask2d, omega, wcode, P, tile, zs,
                                                    ze, P.skey);
}

// ---------------------------------------------------------------------
// Batched step: many independent domains advanced by one launch over a
// packed table of work items (domain, tile, z range), one workgroup per
// item.  The host orders the items longest first and places them so that
// groups of neighbouring tiles land on one XCD (workgroup b takes table
// entry b and runs on XCD b mod 8) while every XCD gets an even share,
// run longest first.
// Every item runs pipe_item, the same code as the single-domain kernel,
// so a domain's result does not depend on what else is in the batch.
// ---------------------------------------------------------------------
template <typename R>
struct BatchDom {
    TiledParams<R> P;
    void *buf[2];           // the domain's two state buffers
    const uint8_t *mask2d;  // z-invariant mask plane (nullptr: periodic)
    const R *omega;         // wall weights (nullptr: periodic)
    int parity;             // buffer holding the state when flip == 0
    int pad_;
};

// Step counters (stochastic rounding) live beside the read-only
// descriptors, two per domain: the launch with flip f reads seeds[2d + f];
// the workgroup of the domain's lead item (one per domain, flagged in the
// table) writes seeds[2d + (f ^ 1)] = that + 1 for the next launch (no
// launch reads and writes the same slot), so hipGraph replays of the two
// launches keep counting.
template <int MODE, bool PERIODIC, int D1, bool FD2, int TYP, bool NTS>
__global__ __launch_bounds__((TYP + 3) * 64) void k_step_batch(
    const BatchDom<typename Store<MODE>::R> *__restrict__ doms,
    uint32_t *__restrict__ seeds, const int4 *__restrict__ work, int nwork, int flip)
{
    using S = typename Store<MODE>::S;
    const int w = blockIdx.x;
    if (w >= nwork) return;
    const int4 it = work[w];  // (domain, tile, zs, ze | lead flag)
    if (it.x < 0) return;     // an empty position of the table
    const BatchDom<typename Store<MODE>::R> &D = doms[it.x];
    // The parameters are copied out of the descriptor before the first
    // store: read through the reference, every use in the z loop would be
    // reloaded after each buffer store (which may alias the descriptor),
    // and each reload's wait also waits for the LDS reads in flight.
    const TiledParams<typename Store<MODE>::R> P = D.P;
    const int par = D.parity ^ flip;
    const uint32_t seed = seeds[2 * it.x + flip];
    if ((it.w & BATCH_LEAD) && threadIdx.x == 0) seeds[2 * it.x + (flip ^ 1)] = seed + 1u;
    pipe_item<MODE, PERIODIC, D1, FD2, TYP, NTS, false>((const S *)D.buf[par],
                                                       (S *)D.buf[par ^ 1], D.mask2d, D.omega,
                                                       nullptr, P, it.y, it.z,
                                                       it.w & ~BATCH_LEAD, sr_key(seed));
}

template <int MODE, int TYP, bool NTS>
hipError_t launch_batch_rows(const BatchLaunch &b)
{
    using R = typename Store<MODE>::R;
    const auto *doms = (const BatchDom<R> *)b.doms;
    const dim3 grid(b.grid), block((TYP + 3) * 64);
#define FSLBM_BATCH(PER, DD, F2)                                                        \
    k_step_batch<MODE, PER, DD, F2, TYP, NTS><<<grid, block, 0, b.stream>>>(doms, b.seeds, b.work, \
                                                                        b.nwork, b.flip)
    if (b.periodic) {
        if (b.d1 == 1) { if (b.fd2) FSLBM_BATCH(true, 1, true); else FSLBM_BATCH(true, 1, false); }
        else { if (b.fd2) FSLBM_BATCH(true, 2, true); else FSLBM_BATCH(true, 2, false); }
    } else {
        if (b.d1 == 1) { if (b.fd2) FSLBM_BATCH(false, 1, true); else FSLBM_BATCH(false, 1, false); }
        else { if (b.fd2) FSLBM_BATCH(false, 2, true); else FSLBM_BATCH(false, 2, false); }
    }
#undef FSLBM_BATCH
    return hipGetLastError();
}

// A kernel class's tile rows follow from its precision, periodicity and
// corrector variant, as the rows its domains' tiles were planned with.
template <int MODE>
hipError_t launch_batch_mode(const BatchLaunch &b)
{
    const int rows = plan_rows(MODE, b.periodic, b.fd2);
    if constexpr (MODE == FSLBM_F16S || MODE == FSLBM_BF16S)
        if (rows == F16S_WALLED_ROWS) return launch_batch_rows<MODE, F16S_WALLED_ROWS, false>(b);
    if (rows != pipe_rows(MODE)) return hipErrorInvalidValue;
    if constexpr (MODE == FSLBM_F32)
        if (b.nt) return launch_batch_rows<MODE, pipe_rows(MODE), true>(b);
    return launch_batch_rows<MODE, pipe_rows(MODE), false>(b);
}

// Host-side fill of the per-launch parameters of one domain.
template <int MODE>
void fill_params(TiledParams<typename Store<MODE>::R> &P, const Geo &g, const StepArgs &a,
                 int chunk_req, const int *tiles, int ntiles, int rows)
{
    using R = typename Store<MODE>::R;
    const StepK<R> k = [&] {
        if constexpr (MODE == FSLBM_F64) return make_consts_f64(a);
        else return make_consts_f32(a);
    }();
    P.nx = g.nx;
    P.ny = g.ny;
    P.nz = g.nz;
    P.nxy = g.nx * g.ny;
    P.ntx = (g.nx + TX - 1) / TX;
    P.nty = (g.ny + rows - 1) / rows;
    P.tiles = tiles && ntiles > 0 ? tiles : nullptr;
    P.ntile = P.tiles ? ntiles : P.ntx * P.nty;
    P.chunk = chunk_req > 0 ? chunk_req : default_chunk(g, rows, P.ntile);
    if (P.chunk > g.nz) P.chunk = g.nz;
    P.nchunks = (g.nz + P.chunk - 1) / P.chunk;
    P.cq = g.nz / P.nchunks;
    P.cr = g.nz - P.cq * P.nchunks;
    P.nwork = P.ntile * P.nchunks;
    P.per_xcd = (P.nwork + 7) / 8;
    P.wskip = 0;
    P.pin = k.pin;
    P.gauge = k.gauge;
    P.c0 = k.c0;
    P.c1 = k.c1;
    P.c2 = k.c2;
    P.g = k.dtddy * k.m;
    P.tb = k.c;
    P.cc = k.cc;
    P.ku = P.tb * P.g * P.c0 + P.cc;
    P.ks1 = P.tb * P.g * P.c1;
    P.ks2 = P.tb * P.g * P.c2;
    P.kx = P.tb / R(216);
    P.q = a.q;
    P.skey = k.skey;
}

template <int MODE, bool PERIODIC, int D1, int TYP, bool NTS, bool M3>
hipError_t launch_pipe_nts(const void *in, void *out, const TiledAux &aux,
                           const TiledParams<typename Store<MODE>::R> &P, int grid,
                           hipStream_t s)
{
    using S = typename Store<MODE>::S;
    using R = typename Store<MODE>::R;
    if (P.c2 != R(0))
        k_step_pipe<MODE, PERIODIC, D1, true, TYP, NTS, M3><<<grid, (TYP + 3) * 64, 0, s>>>(
            (const S *)in, (S *)out, aux.mask2d, (const R *)aux.omega, aux.wcode, P);
    else  // blend = 0: the order-4 arms have zero weight
        k_step_pipe<MODE, PERIODIC, D1, false, TYP, NTS, M3><<<grid, (TYP + 3) * 64, 0, s>>>(
            (const S *)in, (S *)out, aux.mask2d, (const R *)aux.omega, aux.wcode, P);
    return hipGetLastError();
}

// FP32 and FP64 steps store non-temporal when the domain's state exceeds
// the last-level cache (stream_stores); the 16-bit modes never do.
template <int MODE, bool PERIODIC, int D1, int TYP, bool M3>
hipError_t launch_pipe(const Geo &g, const void *in, void *out, const TiledAux &aux,
                       const TiledParams<typename Store<MODE>::R> &P, int grid, hipStream_t s)
{
    if constexpr (MODE == FSLBM_F32 || MODE == FSLBM_F64)
        if (stream_stores(2 * state_bytes(g, MODE)))
            return launch_pipe_nts<MODE, PERIODIC, D1, TYP, true, M3>(in, out, aux, P, grid, s);
    return launch_pipe_nts<MODE, PERIODIC, D1, TYP, false, M3>(in, out, aux, P, grid, s);
}

template <int MODE, bool PERIODIC, int D1>
hipError_t launch_t(const Geo &g, const void *in, void *out,
                    const TiledAux &aux,
                    const TiledParams<typename Store<MODE>::R> &P,
                    hipStream_t s)
{
    using S = typename Store<MODE>::S;
    using R = typename Store<MODE>::R;
    const int grid = 8 * P.per_xcd;
    // FP64 runs the pipelined kernel on 4-row tiles where its buffer
    // offsets reach (f64_pipe_fits), else the two-barrier kernel, whose
    // field offsets are 64-bit.
    if constexpr (MODE == FSLBM_F64) {
        if (aux.wcode) return hipErrorInvalidValue;  // z-invariant masks only
        if (aux.rows == F64_PIPE_ROWS)
            return launch_pipe<MODE, PERIODIC, D1, F64_PIPE_ROWS, false>(g, in, out, aux, P, grid, s);
        if (aux.rows != TY) return hipErrorInvalidValue;
        k_step_tiled<MODE, PERIODIC, D1><<<grid, NT, 0, s>>>(
            (const S *)in, (S *)out, aux.mask2d, (const R *)aux.omega, P);
    }
    else if constexpr (!PERIODIC) {
        // a mask that varies along z: wall codes per cell
        if (aux.wcode) {
            if (aux.rows != M3_ROWS) return hipErrorInvalidValue;
            return launch_pipe<MODE, PERIODIC, D1, M3_ROWS, true>(g, in, out, aux, P, grid, s);
        }
        if constexpr (MODE == FSLBM_F16S || MODE == FSLBM_BF16S)
            if (aux.rows == F16S_WALLED_ROWS)
                return launch_pipe<MODE, PERIODIC, D1, F16S_WALLED_ROWS, false>(g, in, out, aux, P,
                                                                            grid, s);
    } else if constexpr (MODE == FSLBM_F32) {
        if (aux.rows == F32_BOX_ROWS)
            return launch_pipe<MODE, PERIODIC, D1, F32_BOX_ROWS, false>(g, in, out, aux, P, grid, s);
    }
    if constexpr (MODE != FSLBM_F64) {
        if (aux.rows != pipe_rows(MODE)) return hipErrorInvalidValue;
        return launch_pipe<MODE, PERIODIC, D1, pipe_rows(MODE), false>(g, in, out, aux, P, grid, s);
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(const Geo &g, const void *in, void *out,
                       const TiledAux &aux, const StepArgs &a)
{
    using R = typename Store<MODE>::R;
    TiledParams<R> P;
    fill_params<MODE>(P, g, a, aux.chunk, aux.tiles, aux.ntiles, aux.rows);
    P.wskip = aux.walls_zero ? 1 : 0;
    const bool per = g.periodic != 0;
    if (a.d1 == 1)
        return per ? launch_t<MODE, true, 1>(g, in, out, aux, P, aux.stream)
                   : launch_t<MODE, false, 1>(g, in, out, aux, P, aux.stream);
    return per ? launch_t<MODE, true, 2>(g, in, out, aux, P, aux.stream)
               : launch_t<MODE, false, 2>(g, in, out, aux, P, aux.stream);
}

}  // namespace

int device_cus()
{
    // per device; contexts on several host threads may ask at once
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int v = cached[dev].load(std::memory_order_relaxed);
    if (v <= 0) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            v <= 0)
            v = 256;  // MI355X
        cached[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

namespace {

// Cost of starting a work item (pipeline fill, first loads, dispatch) in
// units of one plane of a tile.  With 3 the model ranks the measured step
// times of the 256x256x1024 pipe (78 or 112 tiles) over chunk lengths 64
// to 342 in their measured order.
constexpr int ITEM_OVERHEAD = 3;

// Modelled duration of a step with `nch` chunks per tile: one workgroup
// per CU at a time, each XCD running its contiguous run of items in
// rounds.
double chunk_cost(int nz, long long ntile, int nch, int cus)
{
    const long long items = ntile * nch;
    const long long per_xcd = (items + 7) / 8;
    const long long slots = std::max(1, cus / 8);
    const long long rounds = (per_xcd + slots - 1) / slots;
    const int len = (nz + nch - 1) / nch;
    return (double)rounds * (len + ITEM_OVERHEAD);
}

}  // namespace

// The pipelined kernel's buffer offsets reach an FP64 domain: each field
// below 2^32 bytes (2^31 in a periodic box, whose out-of-box lanes store at
// an offset past 2^31 that the range check drops).
static bool f64_pipe_fits(const Geo &g)
{
    const unsigned long long bytes = 8ull * (unsigned long long)g.N;
    return bytes < (g.periodic ? (1ull << 31) : (1ull << 32));
}

int plan_rows(int precision, bool periodic, bool order4, bool zvar)
{
    if (precision == FSLBM_F64) return F64_PIPE_ROWS;
    if (zvar && !periodic) return M3_ROWS;
    if ((precision == FSLBM_F16S || precision == FSLBM_BF16S) && !periodic && !order4)
        return F16S_WALLED_ROWS;
    return pipe_rows(precision);
}

int solo_rows(int precision, const Geo &g, bool order4, bool zvar)
{
    // FP64 past the pipelined kernel's offsets: the two-barrier kernel
    if (precision == FSLBM_F64 && !f64_pipe_fits(g)) return TY;
    const int rows = plan_rows(precision, g.periodic != 0, order4, zvar);
    if (precision != FSLBM_F32 || !g.periodic) return rows;
    // every tile of a periodic box is stepped: compare the modelled step
    // times of the best schedules on either tile shape
    auto best = [&](int r) {
        const long long T = (long long)((g.nx + TX - 1) / TX) * ((g.ny + r - 1) / r);
        const int len = chunk_candidates(g, r, (int)T, 1)[0];
        return chunk_cost(g.nz, T, (g.nz + len - 1) / len, device_cus());
    };
    return best(F32_BOX_ROWS) * F32_BOX_ROWS_COST < best(rows) ? F32_BOX_ROWS : rows;
}

std::vector<int> chunk_candidates(const Geo &g, int rows, int ntiles, int count)
{
    const long long T = ntiles > 0 ? ntiles
                                   : (long long)((g.nx + TX - 1) / TX) * ((g.ny + rows - 1) / rows);
    const int cus = device_cus();
    // chunks of at least 4 planes: a domain too small to fill the GPU
    // runs fastest in short chunks, whose pipelines run side by side
    const int minlen = std::min(4, g.nz);
    std::vector<std::pair<double, int>> c;  // (cost, chunk length)
    for (int nch = 1; nch <= g.nz / minlen; nch++) {
        const int len = (g.nz + nch - 1) / nch;
        const int real = (g.nz + len - 1) / len;  // chunks of that length
        if (real != nch) continue;
        c.push_back({chunk_cost(g.nz, T, nch, cus), len});
    }
    if (c.empty()) c.push_back({0.0, g.nz});
    // cheapest first; among equal costs the longer chunk (fewer items)
    std::sort(c.begin(), c.end(), [](const auto &a, const auto &b) {
        return a.first < b.first || (a.first == b.first && a.second > b.second);
    });
    std::vector<int> out;
    for (const auto &e : c) {
        if ((int)out.size() >= count) break;
        out.push_back(e.second);
    }
    return out;
}

int default_chunk(const Geo &g, int rows, int ntiles)
{
    return chunk_candidates(g, rows, ntiles, 1)[0];
}

std::vector<int> active_tiles(const Geo &g, int rows, const uint8_t *mask2d)
{
    std::vector<int> out;
    if (g.periodic || !mask2d) return out;
    const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + rows - 1) / rows;
    for (int ty = 0; ty < nty; ty++)
        for (int tx = 0; tx < ntx; tx++) {
            bool fluid = false;
            for (int y = ty * rows; y < std::min(g.ny, (ty + 1) * rows) && !fluid; y++)
                for (int x = tx * TX; x < std::min(g.nx, (tx + 1) * TX); x++)
                    if (mask2d[(size_t)y * g.nx + x] == 0) { fluid = true; break; }
            if (fluid) out.push_back(ty * ntx + tx);
        }
    // a mask without fluid still gets one item, which stores nothing
    if (out.empty()) out.push_back(0);
    return out;
}

long long tiles_cells(const Geo &g, int rows, const std::vector<int> &tiles)
{
    const int ntx = (g.nx + TX - 1) / TX;
    long long n = 0;
    for (int t : tiles) {
        const int ty = t / ntx, tx = t - ty * ntx;
        n += (long long)std::min(TX, g.nx - tx * TX) * std::min(rows, g.ny - ty * rows);
    }
    return n * g.nz;
}

bool tiled_supported(const Geo &g, int precision, int d1, bool z_invariant)
{
    if (d1 != 1 && d1 != 2) return false;
    // a mask that varies along z: the pipelined kernel's M3 variant
    if (!g.periodic && !z_invariant && precision == FSLBM_F64) return false;
    // The pipelined kernel addresses each field (FP32) or pair plane
    // (16-bit modes) through a buffer resource with 32-bit byte offsets:
    // below 2^32 bytes, and below 2^31 in a periodic box, whose out-of-box
    // lanes store at an offset past 2^31 that the range check drops.  The
    // FP64 kernel indexes with 64-bit field offsets.
    if (precision != FSLBM_F64) {
        const unsigned long long bytes = 4ull * (unsigned long long)g.N;
        if (bytes >= (g.periodic ? (1ull << 31) : (1ull << 32))) return false;
    }
    // The queue runs up to 2*d1 + 3 planes past the chunk end; a single
    // periodic wrap must reach every such plane.
    if (g.nz < 16) return false;
    return true;
}

namespace {

// Wall codes of every cell of a walled box (the layout above wall_word).
__global__ void k_wall_codes(const uint8_t *__restrict__ m, int nx, int ny, long long N,
                             uint16_t *__restrict__ code)
{
    const long long nxy = (long long)nx * ny;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const long long pl = i / nxy * nxy;  // first cell of the plane
        const int c = (int)(i - pl);
        const int x = c % nx, y = c / nx;
        int s0 = 0, sx = 0, sy = 0;
        for (int cy = -1; cy <= 1; cy++)
            for (int cx = -1; cx <= 1; cx++) {
                const int xx = x + cx, yy = y + cy;
                const bool wall = xx < 0 || yy < 0 || xx >= nx || yy >= ny ||
                                  m[pl + xx + (long long)nx * yy] != 0;
                if (!wall) continue;
                const int w = (cx ? 1 : 4) * (cy ? 1 : 4);  // corner 1, edge 4, centre 16
                s0 += w;
                sx += w * cx;
                sy += w * cy;
            }
        code[i] = (uint16_t)(s0 | (m[i] != 0 ? WC_WALL : 0u) | ((sx + 6) << 8) | ((sy + 6) << 12));
    }
}

}  // namespace

hipError_t launch_wall_codes(const Geo &g, const uint8_t *mask, uint16_t *codes, hipStream_t s)
{
    const long long blocks = std::min<long long>((g.N + 255) / 256, 65536);
    k_wall_codes<<<(int)blocks, 256, 0, s>>>(mask, g.nx, g.ny, g.N, codes);
    return hipGetLastError();
}

hipError_t launch_step_tiled(const Geo &g, int precision, const void *in,
                             void *out, const TiledAux &aux, const StepArgs &a)
{
    switch (precision) {
    case FSLBM_F64: return launch_mode<FSLBM_F64>(g, in, out, aux, a);
    case FSLBM_F32: return launch_mode<FSLBM_F32>(g, in, out, aux, a);
    case FSLBM_F16S: return launch_mode<FSLBM_F16S>(g, in, out, aux, a);
    case FSLBM_BF16S: return launch_mode<FSLBM_BF16S>(g, in, out, aux, a);
    }
    return hipErrorInvalidValue;
}

size_t batch_dom_bytes(int precision)
{
    return precision == FSLBM_F64 ? sizeof(BatchDom<double>) : sizeof(BatchDom<float>);
}

int batch_fill_dom(void *dst, int precision, const Geo &g, const StepArgs &a, int chunk,
                   void *buf0, void *buf1, const uint8_t *mask2d, const void *omega,
                   const int *tiles, int ntiles, int parity, int split[3])
{
    // the domain's tiles were planned with these rows (batch_prepare)
    const int rows = plan_rows(precision, g.periodic != 0, a.blend != 0.0);
    auto fill = [&](auto *D, auto tag) -> int {
        constexpr int M = decltype(tag)::value;
        fill_params<M>(D->P, g, a, chunk, tiles, ntiles, rows);
        D->P.wskip = 1;  // a batch uploads each domain into both buffers
        if (split) {
            split[0] = D->P.ntile;
            split[1] = D->P.cq;
            split[2] = D->P.cr;
        }
        D->buf[0] = buf0;
        D->buf[1] = buf1;
        D->mask2d = mask2d;
        D->omega = (decltype(D->omega))omega;
        D->parity = parity;
        D->pad_ = 0;
        return D->P.nwork;
    };
    switch (precision) {
    case FSLBM_F32: return fill((BatchDom<float> *)dst, std::integral_constant<int, FSLBM_F32>{});
    case FSLBM_F16S: return fill((BatchDom<float> *)dst, std::integral_constant<int, FSLBM_F16S>{});
    case FSLBM_BF16S: return fill((BatchDom<float> *)dst, std::integral_constant<int, FSLBM_BF16S>{});
    }
    return -1;
}

void batch_set_parity(void *dst, int precision, int parity)
{
    if (precision == FSLBM_F64) ((BatchDom<double> *)dst)->parity = parity;
    else ((BatchDom<float> *)dst)->parity = parity;
}

hipError_t launch_step_batch(const BatchLaunch &b)
{
    switch (b.precision) {
    case FSLBM_F32: return launch_batch_mode<FSLBM_F32>(b);
    case FSLBM_F16S: return launch_batch_mode<FSLBM_F16S>(b);
    case FSLBM_BF16S: return launch_batch_mode<FSLBM_BF16S>(b);
    }
    return hipErrorInvalidValue;
}

}  // namespace fslbm
