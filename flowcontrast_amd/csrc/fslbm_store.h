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
};

template <>
struct Store<FSLBM_F32> {
    using S = float;
    using R = float;
    __host__ __device__ static R ld_p(S v, int, const Quant &) { return v; }
    __host__ __device__ static R ld_u(S v, const Quant &) { return v; }
    __host__ __device__ static S st_p(R v, int, const Quant &) { return v; }
    __host__ __device__ static S st_u(R v, const Quant &) { return v; }
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
};

}  // namespace fslbm
