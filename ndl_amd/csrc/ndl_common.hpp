// ndl_common.hpp -- device/host helpers shared by every kernel of libndl_gpu.
//
// Element codecs, the reference's border maps and output narrowing, and fast
// 32-bit division by run-time constants.  gfx950 only; compiled with
// -ffp-contract=off so that NDL_ACCUM_REF arithmetic rounds every product and
// sum separately, as the reference's FP64 loop does (convolution.h:64-70).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ndl_gpu.h"

namespace ndl {

// ---------------------------------------------------------------- storage types
struct bf16_t {
    uint16_t bits;
};
struct f16_t {
    uint16_t bits;
};

template <class T> struct is_int_storage { static constexpr bool value = false; };
template <> struct is_int_storage<uint8_t> { static constexpr bool value = true; };
template <> struct is_int_storage<int8_t> { static constexpr bool value = true; };
template <> struct is_int_storage<uint16_t> { static constexpr bool value = true; };
template <> struct is_int_storage<int16_t> { static constexpr bool value = true; };
template <> struct is_int_storage<int32_t> { static constexpr bool value = true; };

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// round-to-nearest-even float -> bfloat16 (NaN kept quiet)
__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float f16_to_f32(uint16_t h)
{
    _Float16 v;
    __builtin_memcpy(&v, &h, 2);
    return (float)v;
}
__device__ __forceinline__ uint16_t f32_to_f16(float f)
{
    _Float16 v = (_Float16)f; // v_cvt_f16_f32: round to nearest even
    uint16_t h;
    __builtin_memcpy(&h, &v, 2);
    return h;
}
__device__ __forceinline__ uint16_t f64_to_f16(double d)
{
    _Float16 v = (_Float16)d; // direct, single rounding
    uint16_t h;
    __builtin_memcpy(&h, &v, 2);
    return h;
}
__device__ __forceinline__ uint16_t f64_to_bf16(double d) { return f32_to_bf16((float)d); }

// static_cast<integral DstT>(double) as the reference's x86-64 code performs
// it (convolution.h:69): truncate into int32 (NaN/out of range -> INT32_MIN),
// keep the low bits.  std::round first for gaussian_blur (convolution.h:184).
__device__ __forceinline__ int32_t cvtt_i32(double v)
{
    return (v > -2147483649.0 && v < 2147483648.0) ? (int32_t)v : INT32_MIN;
}

// ------------------------------------------------------------------ loaders
template <class T> __device__ __forceinline__ double load_f64(const T* p) { return (double)*p; }
template <> __device__ __forceinline__ double load_f64<bf16_t>(const bf16_t* p) { return (double)bf16_to_f32(p->bits); }
template <> __device__ __forceinline__ double load_f64<f16_t>(const f16_t* p) { return (double)f16_to_f32(p->bits); }

template <class T> __device__ __forceinline__ float load_f32(const T* p) { return (float)*p; }
template <> __device__ __forceinline__ float load_f32<bf16_t>(const bf16_t* p) { return bf16_to_f32(p->bits); }
template <> __device__ __forceinline__ float load_f32<f16_t>(const f16_t* p) { return f16_to_f32(p->bits); }

template <class Acc, class T> __device__ __forceinline__ Acc load_acc(const T* p)
{
    if constexpr (sizeof(Acc) == 8) return load_f64(p);
    else return load_f32(p);
}

// ------------------------------------------------------------------- stores
template <class T> __device__ __forceinline__ void store_from_f64(T* p, double v, int rounding)
{
    if constexpr (is_int_storage<T>::value) {
        if (rounding == NDL_ROUND_NEAREST) v = round(v); // half away from zero, as std::round
        *p = (T)(uint32_t)cvtt_i32(v);
    } else {
        *p = (T)v;
    }
}
template <> __device__ __forceinline__ void store_from_f64<bf16_t>(bf16_t* p, double v, int)
{
    p->bits = f64_to_bf16(v);
}
template <> __device__ __forceinline__ void store_from_f64<f16_t>(f16_t* p, double v, int)
{
    p->bits = f64_to_f16(v);
}

template <class T> __device__ __forceinline__ void store_from_f32(T* p, float v)
{
    *p = (T)v;
}
template <> __device__ __forceinline__ void store_from_f32<bf16_t>(bf16_t* p, float v) { p->bits = f32_to_bf16(v); }
template <> __device__ __forceinline__ void store_from_f32<f16_t>(f16_t* p, float v) { p->bits = f32_to_f16(v); }

// ------------------------------------------------------------- border maps
// _clamp/_wrap/_reflect (mathHelpers.h:23-25), single fold each; callers
// guarantee -m <= x < 2m for wrap/reflect (host-checked tap reach).
__host__ __device__ __forceinline__ int64_t border_map(int border, int64_t m, int64_t x)
{
    if (border == NDL_WRAP) return x < 0 ? x + m : (x >= m ? x - m : x);
    if (border == NDL_REFLECT) return x < 0 ? -x - 1 : (x >= m ? 2 * m - x - 1 : x);
    return x < 0 ? 0 : (x >= m ? m - 1 : x);
}
__host__ __device__ __forceinline__ int border_map32(int border, int m, int x)
{
    if (border == NDL_WRAP) return x < 0 ? x + m : (x >= m ? x - m : x);
    if (border == NDL_REFLECT) return x < 0 ? -x - 1 : (x >= m ? 2 * m - x - 1 : x);
    return x < 0 ? 0 : (x >= m ? m - 1 : x);
}
// border map followed by a clamp into [0, m): positions that only feed
// discarded outputs of partial tiles stay in bounds whatever their reach.
__device__ __forceinline__ int safe_map32(int border, int m, int x)
{
    int y = border_map32(border, m, x);
    return y < 0 ? 0 : (y >= m ? m - 1 : y);
}

// -------------------------------------------- division by run-time constants
// q = floor(n / d) for n < 2^32 via a 33-bit magic multiplier (round-up
// method): l = ceil(log2 d), mult = floor(2^32 (2^l - d) / d) + 1,
// q = (umulhi(n, mult) + n) >> l computed in 64 bits.
struct FastDiv {
    uint32_t d, mult, shift;
    __host__ void init(uint32_t dd)
    {
        d = dd;
        uint32_t l = 0;
        while ((1ull << l) < dd) l++;
        shift = l;
        mult = (uint32_t)((((1ull << l) - dd) << 32) / dd + 1);
    }
    __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
        uint32_t hi = (uint32_t)(((uint64_t)n * mult) >> 32);
        return (uint32_t)(((uint64_t)hi + n) >> shift);
    }
};

} // namespace ndl
