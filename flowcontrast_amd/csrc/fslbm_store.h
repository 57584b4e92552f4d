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

// Random words of one cell in one step.  The SR conversions take their
// random bits from the top of the word, so a and a << 16 (b and b << 16)
// give the four fields disjoint bits.
struct SrWords {
    uint32_t a, b;
};
__host__ __device__ inline SrWords sr_words(uint32_t cell, uint32_t key)
{
    const uint32_t a = mix32(cell ^ key);
    return {a, mix32(a + 0x9e3779b9u)};
}

template <int MODE>
struct Store;

template <>
struct Store<FSLBM_F64> {
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

// IEEE binary16 storage of p - (pbase + pslope*z) and of u, scaled.
template <>
struct Store<FSLBM_F16S> {
    using S = _Float16;
    using R = float;
    __host__ __device__ static R ld_p(S v, int z, const Quant &q)
    {
        return (q.pbase + q.pslope * (float)z) + q.sp * (float)v;
    }
    __host__ __device__ static R ld_u(S v, const Quant &q)
    {
        return q.su * (float)v;
    }
    __host__ __device__ static S st_p(R v, int z, const Quant &q)
    {
        return (S)((v - (q.pbase + q.pslope * (float)z)) * q.inv_sp);
    }
    __host__ __device__ static S st_u(R v, const Quant &q)
    {
        return (S)(v * q.inv_su);
    }
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
        return sr((v - (q.pbase + q.pslope * (float)z)) * q.inv_sp, r);
    }
    __device__ static S st_u_sr(R v, const Quant &q, uint32_t r)
    {
        return sr(v * q.inv_su, r);
    }
};

// bfloat16 storage of the same shifted, scaled fields.
template <>
struct Store<FSLBM_BF16S> {
    using S = uint16_t;
    using R = float;
    __host__ __device__ static R ld_p(S v, int z, const Quant &q)
    {
        return (q.pbase + q.pslope * (float)z) + q.sp * bf16_bits_to_f32(v);
    }
    __host__ __device__ static R ld_u(S v, const Quant &q)
    {
        return q.su * bf16_bits_to_f32(v);
    }
    __host__ __device__ static S st_p(R v, int z, const Quant &q)
    {
        return f32_to_bf16_bits((v - (q.pbase + q.pslope * (float)z)) *
                                q.inv_sp);
    }
    __host__ __device__ static S st_u(R v, const Quant &q)
    {
        return f32_to_bf16_bits(v * q.inv_su);
    }
    static constexpr bool SR = true;
    __device__ static S sr(float x, uint32_t r)
    {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        b2 t = {0, 0};
        t = __builtin_amdgcn_cvt_sr_bf16_f32(t, x, r, false);
        return __builtin_bit_cast(uint16_t, t.x);
    }
    __device__ static S st_p_sr(R v, int z, const Quant &q, uint32_t r)
    {
        return sr((v - (q.pbase + q.pslope * (float)z)) * q.inv_sp, r);
    }
    __device__ static S st_u_sr(R v, const Quant &q, uint32_t r)
    {
        return sr(v * q.inv_su, r);
    }
};

}  // namespace fslbm
