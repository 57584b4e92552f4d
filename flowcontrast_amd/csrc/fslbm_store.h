// Storage traits of the four precision modes: how a stored value becomes
// an arithmetic value and back.  Device and host inline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "fslbm_internal.h"

namespace fslbm {

__host__ __device__ inline uint16_t f32_to_bf16_bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u)  // Inf / NaN: truncate, keep NaN
        return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);        // round to nearest even
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf16_bits_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// 32-bit integer hash ("lowbias32", C. Wellons): the random bits of the
// stochastic rounding of the 16-bit storage modes.
__host__ __device__ inline uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// Per-launch key of step `seed` (a step counter).
__host__ __device__ inline uint32_t sr_key(uint32_t seed)
{
    return mix32(seed * 0x9e3779b9u + 0x632be5abu);
}

// Random operands of the stochastic rounding of one cell in one step,
// one per field (p, ux, uy, uz).  The SR conversions take their random
// bits from the top of the operand (13 bits for binary16, 16 for
// bfloat16).
//
// They come from two words per cell, each a per-column salt (a hash of the
// column index x + nx*y, computed once per column) combined with a
// per-plane key (a hash of the step key and the plane index z, the same
// for every column of the plane): x = (salt ^ key) * M, then its high half
// folded into its low half, w = x ^ (x >> 16).  Fields 0 and 2 draw the
// top half of their word w, fields 1 and 3 its bottom half (w << 16).  A
// word is uniform over 2^32 values, so its two halves are independent;
// the fold makes the bottom half depend on every bit of the product (the
// product's own low bits depend on the low bits of its operand only).  For a fixed cell the keys are fresh hashes in
// every step, so each draw is uniform and independent from step to step
// (stochastic rounding stays unbiased); columns have unrelated salts and
// planes unrelated keys, so neighbouring cells along x, y and z draw
// unrelated values (tests/test_rounding_words.py measures both).  In the
// kernels the plane keys are uniform across a wave and computed on the
// scalar unit, leaving one vector multiply per word (against two per word
// for a full hash of the cell index).
struct SrWords {
    uint32_t r[4];
};
struct SrPlane {
    uint32_t a, b;
};
__host__ __device__ inline uint32_t sr_salt(uint32_t col)
{
    return mix32(col + 0x2545f491u);
}
__host__ __device__ inline SrPlane sr_plane(uint32_t key, uint32_t z)
{
    const uint32_t a = mix32(key ^ (z * 0x9e3779b9u));
    return {a, mix32(a + 0x632be5abu)};
}
__host__ __device__ inline SrWords sr_words(uint32_t salt, SrPlane k)
{
    uint32_t x = (salt ^ k.a) * 0x846ca68bu;
    uint32_t y = (salt ^ k.b) * 0x7feb352du;
    x ^= x >> 16;
    y ^= y >> 16;
    return {{x, x << 16, y, y << 16}};
}

template <int MODE>
struct Store;

template <>
struct Store<FSLBM_F64> {
    static constexpr int mode = FSLBM_F64;
    using S = double;
    using R = double;
    __host__ __device__ static R ld_p(S v, int, const Quant &) { return v; }
    __host__ __device__ static R ld_u(S v, const Quant &) { return v; }
    __host__ __device__ static S st_p(R v, int, const Quant &) { return v; }
    __host__ __device__ static S st_u(R v, const Quant &) { return v; }
    static constexpr bool SR = false;
    __device__ static S st_p_sr(R v, int, const Quant &, uint32_t) { return v; }
    __device__ static S st_u_sr(R v, const Quant &, uint32_t) { return v; }
};

template <>
struct Store<FSLBM_F32> {
    static constexpr int mode = FSLBM_F32;
    using S = float;
    using R = float;
    __host__ __device__ static R ld_p(S v, int, const Quant &) { return v; }
    __host__ __device__ static R ld_u(S v, const Quant &) { return v; }
    __host__ __device__ static S st_p(R v, int, const Quant &) { return v; }
    __host__ __device__ static S st_u(R v, const Quant &) { return v; }
    static constexpr bool SR = false;
    __device__ static S st_p_sr(R v, int, const Quant &, uint32_t) { return v; }
    __device__ static S st_u_sr(R v, const Quant &, uint32_t) { return v; }
};

// IEEE binary16 storage of (p - (pbase + pslope*z)) / sp and of u itself.
// The velocity is stored unscaled: a binary16 value keeps 11 significant
// bits whatever its scale, |u| < 65504 never overflows, and below 2^-14
// (6.1e-5) the subnormals keep an absolute precision of 2^-24 (6e-8), far
// below the relative error of any velocity that matters to the flow.  So
// decoding and encoding a velocity is a conversion alone.  The pressure is
// stored as its deviation from the mean line, a value far smaller than p,
// and scaled so that small deviations stay in the normal range.
template <>
struct Store<FSLBM_F16S> {
    static constexpr int mode = FSLBM_F16S;
    using S = _Float16;
    using R = float;
    __host__ __device__ static R ld_p(S v, int z, const Quant &q)
    {
        return (q.pbase + q.pslope * (float)z) + q.sp * (float)v;
    }
    __host__ __device__ static R ld_u(S v, const Quant &) { return (float)v; }
    __host__ __device__ static S st_p(R v, int z, const Quant &q)
    {
        return (S)((v - (q.pbase + q.pslope * (float)z)) * q.inv_sp);
    }
    __host__ __device__ static S st_u(R v, const Quant &) { return (S)v; }
    // Device stores round stochastically (random word r): the expected
    // stored value is the unrounded one, so per-step increments far below
    // half an FP16 ulp are not lost (the host's initial encode rounds to
    // nearest).
    static constexpr bool SR = true;
    __device__ static S sr(float x, uint32_t r)
    {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        h2 t = {0, 0};
        t = __builtin_amdgcn_cvt_sr_f16_f32(t, x, r, false);
        return t.x;
    }
    __device__ static S st_p_sr(R v, int z, const Quant &q, uint32_t r)
    {
        return sr(enc_p(v, z, q), r);
    }
    __device__ static S st_u_sr(R v, const Quant &, uint32_t r) { return sr(v, r); }
    // the value a pressure stores before rounding
    __host__ __device__ static float enc_p(R v, int z, const Quant &q)
    {
        return (v - (q.pbase + q.pslope * (float)z)) * q.inv_sp;
    }
    // Two encoded values rounded into the low and high halves of one
    // 32-bit word (a pair of the 16-bit state layout), each as sr() rounds
    // it alone.
    __device__ static uint32_t sr2(float lo, uint32_t rlo, float hi, uint32_t rhi)
    {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        // both halves are written below: the initial bits are any value
        // at hand (no register needs clearing first)
        h2 t = __builtin_bit_cast(h2, lo);
        t = __builtin_amdgcn_cvt_sr_f16_f32(t, lo, rlo, false);
        t = __builtin_amdgcn_cvt_sr_f16_f32(t, hi, rhi, true);
        return __builtin_bit_cast(uint32_t, t);
    }
};

// bfloat16 storage of the same fields (bfloat16 has the exponent range of
// float, so the velocity is stored unscaled too).
template <>
struct Store<FSLBM_BF16S> {
    static constexpr int mode = FSLBM_BF16S;
    using S = uint16_t;
    using R = float;
    __host__ __device__ static R ld_p(S v, int z, const Quant &q)
    {
        return (q.pbase + q.pslope * (float)z) + q.sp * bf16_bits_to_f32(v);
    }
    __host__ __device__ static R ld_u(S v, const Quant &) { return bf16_bits_to_f32(v); }
    __host__ __device__ static S st_p(R v, int z, const Quant &q)
    {
        return f32_to_bf16_bits((v - (q.pbase + q.pslope * (float)z)) *
                                q.inv_sp);
    }
    __host__ __device__ static S st_u(R v, const Quant &) { return f32_to_bf16_bits(v); }
    static constexpr bool SR = true;
    // Stochastic rounding to bfloat16 as integer arithmetic: the top 16
    // bits of r added to the float's bits below the bfloat16 mantissa carry
    // into it with probability (dropped bits) / 2^16, then the top half is
    // kept (a carry out of the mantissa moves to the next binade, as it
    // should; finite values stay finite short of the largest binade).
    __device__ static S sr(float x, uint32_t r)
    {
        return (S)((__builtin_bit_cast(uint32_t, x) + (r >> 16)) >> 16);
    }
    __device__ static S st_p_sr(R v, int z, const Quant &q, uint32_t r)
    {
        return sr(enc_p(v, z, q), r);
    }
    __device__ static S st_u_sr(R v, const Quant &, uint32_t r) { return sr(v, r); }
    __host__ __device__ static float enc_p(R v, int z, const Quant &q)
    {
        return (v - (q.pbase + q.pslope * (float)z)) * q.inv_sp;
    }
    __device__ static uint32_t sr2(float lo, uint32_t rlo, float hi, uint32_t rhi)
    {
        const uint32_t a = __builtin_bit_cast(uint32_t, lo) + (rlo >> 16);
        const uint32_t b = __builtin_bit_cast(uint32_t, hi) + (rhi >> 16);
        return __builtin_amdgcn_perm(b, a, 0x07060302u);  // (b.hi << 16) | a.hi
    }
};

// uz of a stored value of column `col` (device code): the 16-bit modes
// add the column's streamwise reference (Quant::uref), once, in float.
template <int MODE>
__device__ inline typename Store<MODE>::R ld_uz(typename Store<MODE>::S v, const Quant &q,
                                                long long col)
{
    const typename Store<MODE>::R w = Store<MODE>::ld_u(v, q);
    if constexpr (Store<MODE>::SR) return q.uref ? w + q.uref[col] : w;
    else return w;
}

// Layout of one stored state of N cells.  FP64 and FP32 hold four planes
// of N values, p, ux, uy, uz.  The 16-bit modes hold two planes of N
// pairs, (p, ux) and (uy, uz), each pair one 32-bit word, so that one
// 32-bit access moves two fields of a cell (a wave reads 256 bytes per
// instruction, as with FP32 fields).  at(f, i, N) is the element index of
// field f (0..3) of cell i.
template <int MODE>
struct StateLayout {
    static constexpr bool PAIRS = sizeof(typename Store<MODE>::S) == 2;
    __host__ __device__ static long long at(int f, long long i, long long N)
    {
        if constexpr (PAIRS) return (long long)(f >> 1) * 2 * N + 2 * i + (f & 1);
        else return (long long)f * N + i;
    }
};
// The same for a precision mode known at run time.
inline long long state_at(int precision, int f, long long i, long long N)
{
    return precision == FSLBM_F16S || precision == FSLBM_BF16S
               ? StateLayout<FSLBM_F16S>::at(f, i, N)
               : StateLayout<FSLBM_F32>::at(f, i, N);
}

}  // namespace fslbm
