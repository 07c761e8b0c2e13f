// kernels_wide.hip -- one convolution pass along any axis with a run-time
// tap count (wide Gaussians: sigma 5 gives 31 taps, the scale stack of
// detect_keypoints reaches 41), where the one-axis kernels' compile-time
// tap counts (up to 25) stop.
//
// A thread computes 8 consecutive outputs j0 .. j0+7 of one line.  Its
// window of 15 elements (positions j0 - c + t0 .. + 14, border folded per
// position, mathHelpers.h:23-25) sits in registers; the taps run in groups
// of 8 with compile-time register indices inside a group, 8 new elements
// loaded per group and 7 carried over, then the remaining taps.  Each
// output therefore sums its taps in ascending order, as the reference
// (processing/convolution.h:52-71): NDL_ACCUM_REF total = total + x * w in
// FP64 with every product and sum rounded separately (this file is built
// with -ffp-contract=off), one rounding into the output; NDL_ACCUM_F32
// fmaf.  Callers send no zero tap.  Along a unit-stride axis consecutive
// threads take consecutive chunks of the same line; along other axes
// consecutive lines (fastest input dimension first), so loads coalesce.
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <class Acc> __device__ __forceinline__ Acc wide_mac(Acc acc, Acc x, Acc w);
template <> __device__ __forceinline__ double wide_mac<double>(double acc, double x, double w) { return acc + x * w; }
template <> __device__ __forceinline__ float wide_mac<float>(float acc, float x, float w)
{
    return __builtin_fmaf(x, w, acc);
}
template <class Acc> __device__ __forceinline__ Acc wide_tap(const WideArgs& a, int t);
template <> __device__ __forceinline__ double wide_tap<double>(const WideArgs& a, int t) { return a.wd[t]; }
template <> __device__ __forceinline__ float wide_tap<float>(const WideArgs& a, int t) { return a.wf[t]; }

template <class TIn, class TOut, class Acc>
__global__ __launch_bounds__(256) void wide_axis_kernel(const WideArgs a)
{
    const uint32_t item = blockIdx.x * 256u + threadIdx.x;
    if (item >= a.items) return;
    uint32_t line, chunk;
    if (a.chunk_fast) {
        line = a.fsplit.div(item);
        chunk = item - line * a.chunks;
    } else {
        chunk = a.fsplit.div(item);
        line = item - chunk * a.lines;
    }
    int64_t ioff = 0, ooff = 0;
    uint32_t rem = line;
    for (int d = 0; d < a.nd; d++) {
        const uint32_t q = a.fdiv[d].div(rem);
        const int64_t c = (int64_t)(rem - q * (uint32_t)a.ext[d]);
        rem = q;
        ioff += c * a.ist[d];
        ooff += c * a.ost[d];
    }
    const TIn* in = (const TIn*)a.in + ioff;
    const int j0 = (int)chunk * 8;
    const int base = j0 - a.center; // position of window slot 0 at t0 = 0
    const bool inside = base >= 0 && base + a.K + 6 < a.m;
    auto load = [&](int e) -> Acc {
        int p = base + e;
        if (!inside) p = safe_map32(a.border, a.m, p);
        return load_acc<Acc>(in + (int64_t)p * a.in_as);
    };
    Acc win[15], acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = 0;
#pragma unroll
    for (int e = 0; e < 7; e++) win[e] = load(e);
    int t0 = 0;
    for (; t0 + 8 <= a.K; t0 += 8) {
#pragma unroll
        for (int e = 0; e < 8; e++) win[7 + e] = load(t0 + 7 + e);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const Acc w = wide_tap<Acc>(a, t0 + t);
#pragma unroll
            for (int j = 0; j < 8; j++) acc[j] = wide_mac<Acc>(acc[j], win[j + t], w);
        }
#pragma unroll
        for (int e = 0; e < 7; e++) win[e] = win[8 + e];
    }
    const int r = a.K - t0; // 0..7 taps left
#pragma unroll
    for (int e = 0; e < 7; e++)
        if (e < r) win[7 + e] = load(t0 + 7 + e);
#pragma unroll
    for (int t = 0; t < 7; t++) {
        if (t < r) {
            const Acc w = wide_tap<Acc>(a, t0 + t);
#pragma unroll
            for (int j = 0; j < 8; j++) acc[j] = wide_mac<Acc>(acc[j], win[j + t], w);
        }
    }
    TOut* out = (TOut*)a.out + ooff;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        if (j0 + j < a.m) {
            if constexpr (sizeof(Acc) == 8) store_from_f64(out + (int64_t)(j0 + j) * a.out_as, acc[j], a.rounding);
            else store_from_f32(out + (int64_t)(j0 + j) * a.out_as, acc[j]);
        }
    }
}

template <class F> hipError_t with_wide_dtype(int dt, F&& f)
{
    switch (dt) {
    case NDL_U8: return f(uint8_t{});
    case NDL_I8: return f(int8_t{});
    case NDL_U16: return f(uint16_t{});
    case NDL_I16: return f(int16_t{});
    case NDL_I32: return f(int32_t{});
    case NDL_F32: return f(float{});
    case NDL_F64: return f(double{});
    case NDL_F16: return f(f16_t{});
    case NDL_BF16: return f(bf16_t{});
    }
    return hipErrorInvalidValue;
}

template <class TIn, class TOut, class Acc> hipError_t go_wide(const WideArgs& a, hipStream_t s)
{
    const uint32_t blocks = (a.items + 255u) / 256u;
    hipLaunchKernelGGL((wide_axis_kernel<TIn, TOut, Acc>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace

hipError_t launch_wide_axis(const WideArgs& a, int in_dt, int out_dt, int accum, hipStream_t s)
{
    if (a.items == 0) return hipSuccess;
    if (a.K < 1) return hipErrorInvalidValue;
    if (accum == NDL_ACCUM_F32) {
        if (in_dt != out_dt) return hipErrorInvalidValue;
        switch (in_dt) {
        case NDL_F32: return go_wide<float, float, float>(a, s);
        case NDL_F16: return go_wide<f16_t, f16_t, float>(a, s);
        case NDL_BF16: return go_wide<bf16_t, bf16_t, float>(a, s);
        }
        return hipErrorInvalidValue;
    }
    if (accum != NDL_ACCUM_REF) return hipErrorInvalidValue;
    if (out_dt == NDL_F64)
        return with_wide_dtype(in_dt, [&](auto tin) { return go_wide<decltype(tin), double, double>(a, s); });
    if (in_dt == NDL_F64)
        return with_wide_dtype(out_dt, [&](auto tout) { return go_wide<double, decltype(tout), double>(a, s); });
    return hipErrorInvalidValue;
}

} // namespace ndl
