// kernels_short.hip -- one convolution pass along a short axis (at most 16
// elements: the channel axis of a {c, x, y} colour image, which
// gaussian_blur() blurs like any other, convolution.h:151-169).
//
// A thread owns one line (every index but the axis fixed) and writes all m
// outputs; the line's elements are read from memory once and from L1 for
// every further tap.  Output j is, in
// the reference's order (processing/convolution.h:52-71), total = 0 then
// total = total + x[e] * w for the nonzero taps ascending (zero taps are
// skipped, detail.h:164-174), e being the tap's element after the border
// fold (resolved on the host, the same for every line).  NDL_ACCUM_REF:
// FP64 with every product and sum rounded separately (this file is built
// with -ffp-contract=off), one rounding into the output type; NDL_ACCUM_F32:
// fmaf.  Lines are numbered fastest input dimension first, so a wave reads
// neighbouring lines together.
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <class Acc> __device__ __forceinline__ Acc short_mac(Acc acc, Acc x, Acc w);
template <> __device__ __forceinline__ double short_mac<double>(double acc, double x, double w)
{
    return acc + x * w;
}
template <> __device__ __forceinline__ float short_mac<float>(float acc, float x, float w)
{
    return __builtin_fmaf(x, w, acc);
}

template <class Acc> __device__ __forceinline__ Acc short_tap(const ShortAxisArgs& a, int n);
template <> __device__ __forceinline__ double short_tap<double>(const ShortAxisArgs& a, int n) { return a.wd[n]; }
template <> __device__ __forceinline__ float short_tap<float>(const ShortAxisArgs& a, int n) { return a.wf[n]; }

template <class TIn, class TOut, class Acc>
__global__ __launch_bounds__(256) void short_axis_kernel(const ShortAxisArgs a)
{
    const uint32_t line = blockIdx.x * 256u + threadIdx.x;
    if (line >= a.lines) return;
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
    TOut* out = (TOut*)a.out + ooff;
    for (int j = 0; j < a.m; j++) {
        Acc total = 0;
        // the line's elements stay in L1 after their first read
        for (int n = 0; n < a.nz; n++)
            total = short_mac<Acc>(total, load_acc<Acc>(in + a.idx[j][n] * a.in_as), short_tap<Acc>(a, n));
        if constexpr (sizeof(Acc) == 8) store_from_f64(out + j * a.out_as, total, a.rounding);
        else store_from_f32(out + j * a.out_as, total);
    }
}

template <class F> hipError_t with_any_dtype(int dt, F&& f)
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

template <class TIn, class TOut, class Acc>
hipError_t launch_m(const ShortAxisArgs& a, hipStream_t s)
{
    const uint32_t blocks = (a.lines + 255u) / 256u;
    hipLaunchKernelGGL((short_axis_kernel<TIn, TOut, Acc>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace

hipError_t launch_short_axis(const ShortAxisArgs& a, int in_dt, int out_dt, int accum, hipStream_t s)
{
    if (a.lines == 0) return hipSuccess;
    if (a.m < 1 || a.m > kShortMax || a.nz < 0 || a.nz > kArgTaps) return hipErrorInvalidValue;
    if (accum == NDL_ACCUM_F32) {
        if (in_dt != out_dt) return hipErrorInvalidValue;
        switch (in_dt) {
        case NDL_F32: return launch_m<float, float, float>(a, s);
        case NDL_F16: return launch_m<f16_t, f16_t, float>(a, s);
        case NDL_BF16: return launch_m<bf16_t, bf16_t, float>(a, s);
        }
        return hipErrorInvalidValue;
    }
    if (accum != NDL_ACCUM_REF) return hipErrorInvalidValue;
    // the chain's passes: into the FP64 intermediate, out of it, or between two
    if (out_dt == NDL_F64)
        return with_any_dtype(in_dt, [&](auto tin) { return launch_m<decltype(tin), double, double>(a, s); });
    if (in_dt == NDL_F64)
        return with_any_dtype(out_dt, [&](auto tout) { return launch_m<double, decltype(tout), double>(a, s); });
    return hipErrorInvalidValue;
}

} // namespace ndl
