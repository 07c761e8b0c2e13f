// kernels_short.hip -- one convolution pass along a short axis (at most 16
// elements: the channel axis of a {c, x, y} colour image, which
// gaussian_blur() blurs like any other, convolution.h:151-169).
//
// A thread owns one line (every index but the axis fixed), reads its m
// elements once into registers and writes all m outputs.  Output j is, in
// the reference's order (processing/convolution.h:52-71), total = 0 then
// total = total + x[e] * w for the nonzero taps ascending (zero taps are
// skipped, detail.h:164-174), e being the tap's element after the border
// fold: the same for every lane, so it is computed in scalar registers and
// selects among the lane's registers.  The taps run outermost and every
// output of the line takes each tap in turn.  NDL_ACCUM_REF:
// FP64 with every product and sum rounded separately (this file is built
// with -ffp-contract=off), one rounding into the output type; NDL_ACCUM_F32:
// fmaf.  Lines are numbered fastest input dimension first, so a wave reads
// neighbouring lines together.
#include <cstdlib>

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

template <class TIn, class TOut, class Acc, int M>
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
    Acc x[M], total[M];
#pragma unroll
    for (int e = 0; e < M; e++) {
        total[e] = 0;
        if (e < a.m) x[e] = load_acc<Acc>(in + e * a.in_as);
    }
    // taps outermost (each output still sums its taps in ascending order);
    // the folded element of tap t for output j is a scalar computation
#pragma unroll 2
    for (int t = 0; t < a.k; t++) {
        const Acc w = short_tap<Acc>(a, t);
        if (w == (Acc)0) continue;
#pragma unroll
        for (int j = 0; j < M; j++) {
            if (j >= a.m) break;
            const int e = border_map32(a.border, a.m, j + t - a.center);
            Acc v = x[0];
#pragma unroll
            for (int u = 1; u < M; u++)
                if (e == u) v = x[u];
            total[j] = short_mac<Acc>(total[j], v, w);
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) {
        if (j >= a.m) break;
        if constexpr (sizeof(Acc) == 8) store_from_f64(out + j * a.out_as, total[j], a.rounding);
        else store_from_f32(out + j * a.out_as, total[j]);
    }
}

// Axes of at most 4 elements with K odd 3..25 and no zero tap (every
// Gaussian): the line's elements are laid out once along the folded window,
// pad[i] = x[B(i - c)] for i < m + K - 1 (one fold and select per element,
// not per tap and output), and the tap loops are unrolled with compile-time
// indices into pad.
template <class TIn, class TOut, class Acc, int K>
__global__ __launch_bounds__(256) void short_axis4_kernel(const ShortAxisArgs a)
{
    constexpr int M = 4, NP = M + K - 1;
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
    Acc x[M];
#pragma unroll
    for (int e = 0; e < M; e++) x[e] = e < a.m ? load_acc<Acc>(in + e * a.in_as) : (Acc)0;
    Acc pad[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) {
        const int e = border_map32(a.border, a.m, i - a.center);
        pad[i] = e == 1 ? x[1] : (e == 2 ? x[2] : (e == 3 ? x[3] : x[0]));
    }
#pragma unroll
    for (int j = 0; j < M; j++) {
        if (j < a.m) {
            Acc total = 0;
#pragma unroll
            for (int t = 0; t < K; t++) total = short_mac<Acc>(total, pad[j + t], short_tap<Acc>(a, t));
            if constexpr (sizeof(Acc) == 8) store_from_f64(out + j * a.out_as, total, a.rounding);
            else store_from_f32(out + j * a.out_as, total);
        }
    }
}

template <int V> using IC = std::integral_constant<int, V>;

template <class F> hipError_t with_short_k(int k, F&& f)
{
    switch (k) {
#define NDL_K(KK) case KK: return f(IC<KK>{});
        NDL_K(3) NDL_K(5) NDL_K(7) NDL_K(9) NDL_K(11) NDL_K(13) NDL_K(15) NDL_K(17) NDL_K(19) NDL_K(21)
        NDL_K(23) NDL_K(25)
#undef NDL_K
    }
    return hipErrorInvalidValue;
}

// the compile-time-tap form for uint8 / uint16 / int16 / float32 / float64
// lines of at most 4 elements whose taps are all nonzero; -1 otherwise
template <class TIn, class TOut, class Acc> hipError_t launch_short4(const ShortAxisArgs& a, hipStream_t s)
{
    if (a.m > 4) return hipErrorInvalidValue;
    for (int t = 0; t < a.k; t++)
        if (a.wd[t] == 0.0) return hipErrorInvalidValue;
    const uint32_t blocks = (a.lines + 255u) / 256u;
    return with_short_k(a.k, [&](auto kk) {
        hipLaunchKernelGGL((short_axis4_kernel<TIn, TOut, Acc, decltype(kk)::value>), dim3(blocks), dim3(256), 0,
                           s, a);
        return hipGetLastError();
    });
}

// NDL_SHORT4=0: the run-time-tap kernel only (for comparison)
inline bool short4_off()
{
    const char* v = getenv("NDL_SHORT4");
    return v && v[0] == '0';
}

template <class T> constexpr bool short4_type()
{
    return std::is_same_v<T, uint8_t> || std::is_same_v<T, uint16_t> || std::is_same_v<T, int16_t> ||
           std::is_same_v<T, float> || std::is_same_v<T, double>;
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
    if constexpr (short4_type<TIn>() && short4_type<TOut>()) {
        if (!short4_off()) {
            const hipError_t e = launch_short4<TIn, TOut, Acc>(a, s);
            if (e != hipErrorInvalidValue) return e;
        }
    }
    const uint32_t blocks = (a.lines + 255u) / 256u;
    if (a.m <= 4) hipLaunchKernelGGL((short_axis_kernel<TIn, TOut, Acc, 4>), dim3(blocks), dim3(256), 0, s, a);
    else if (a.m <= 8) hipLaunchKernelGGL((short_axis_kernel<TIn, TOut, Acc, 8>), dim3(blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((short_axis_kernel<TIn, TOut, Acc, 16>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace

hipError_t launch_short_axis(const ShortAxisArgs& a, int in_dt, int out_dt, int accum, hipStream_t s)
{
    if (a.lines == 0) return hipSuccess;
    if (a.m < 1 || a.m > kShortMax || a.k < 1 || a.k > kArgTaps) return hipErrorInvalidValue;
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
