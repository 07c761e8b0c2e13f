// kernels_generic.hip -- the always-correct kernels: one convolved axis per
// launch over any strides, element types and borders, plus the dense N-D
// kernel and the tap-gradient reduction.
//
// Semantics follow processing/convolution.h:52-71 per output: taps ascending,
// exactly-zero taps skipped (detail.h:164-174), border resolved per tap
// (detail.h:178-194).  NDL_ACCUM_REF accumulates double(src)*double(tap) with
// every product and sum rounded separately (this file is built with
// -ffp-contract=off), so results are bit-identical to the reference.
#include <type_traits>

#include "ndl_internal.hpp"

namespace ndl {

template <class Acc> __device__ __forceinline__ Acc mac(Acc acc, Acc v, Acc w);
template <> __device__ __forceinline__ double mac<double>(double acc, double v, double w)
{
    return acc + v * w; // two roundings, as the reference's `total += src * k`
}
template <> __device__ __forceinline__ float mac<float>(float acc, float v, float w)
{
    return __builtin_fmaf(v, w, acc);
}

template <class Acc> __device__ __forceinline__ Acc tap_of(const AxisPassArgs& a, int t);
template <> __device__ __forceinline__ double tap_of<double>(const AxisPassArgs& a, int t)
{
    return a.wide_d ? a.wide_d[t] : a.taps_d[t];
}
template <> __device__ __forceinline__ float tap_of<float>(const AxisPassArgs& a, int t)
{
    return a.wide_f ? a.wide_f[t] : a.taps_f[t];
}

template <class T, class Acc> __device__ __forceinline__ void store_acc(T* p, Acc v, int rounding)
{
    if constexpr (sizeof(Acc) == 8) store_from_f64(p, v, rounding);
    else store_from_f32(p, v);
}

// ---------------------------------------------------------------- axis pass
template <class TIn, class TOut, class Acc>
__global__ __launch_bounds__(256) void axis_pass_kernel(const AxisPassArgs a)
{
    uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t ioff = 0, ooff = 0, j = 0;
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        uint32_t q = a.fdiv[d].div(rem);
        int64_t c = (int64_t)(rem - q * (uint32_t)a.ext[d]);
        rem = q;
        ioff += c * a.ist[d];
        ooff += c * a.ost[d];
        if (d == a.axis_pos) j = c;
    }
    const TIn* in = (const TIn*)a.in + (ioff - j * a.in_as);
    Acc total = 0;
    if (a.mode == PASS_GATHER) {
        for (int t = 0; t < a.K; t++) {
            Acc w = tap_of<Acc>(a, t);
            if (w == (Acc)0) continue; // uniform: zero taps are not visited
            int64_t x = j + t - a.center;
            if (x < 0 || x >= a.m) x = border_map(a.border, a.m, x);
            total = mac<Acc>(total, load_acc<Acc>(in + x * a.in_as), w);
        }
    } else {
        const Acc* row = nullptr;
        if (j < a.EL) row = (const Acc*)a.table + j * a.W;
        else if (j >= a.m - a.ER) row = (const Acc*)a.table + (a.EL + (j - (a.m - a.ER))) * a.W;
        if (row) {
            for (int s = 0; s < a.W; s++) {
                const Acc w = row[s];
                const int64_t x = j + s - a.L;
                if (x < 0 || x >= a.m || w == (Acc)0) continue; // zero outside the axis
                total = mac<Acc>(total, load_acc<Acc>(in + x * a.in_as), w);
            }
        } else {
            for (int t = 0; t < a.K; t++) {
                const Acc w = tap_of<Acc>(a, t);
                const int64_t x = j + t - a.center;
                if (x < 0 || x >= a.m || w == (Acc)0) continue;
                total = mac<Acc>(total, load_acc<Acc>(in + x * a.in_as), w);
            }
        }
    }
    store_acc<TOut, Acc>((TOut*)a.out + ooff, total, a.rounding);
}

// --------------------------------------------------------- dtype dispatch
// Calls f(T{}) with the storage type of dtype code dt.
template <class F> hipError_t with_dtype(int dt, F&& f)
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

hipError_t launch_axis_pass(const AxisPassArgs& a, int in_dt, int out_dt, int accum, hipStream_t s)
{
    if (a.total == 0) return hipSuccess;
    uint32_t blocks = (a.total + 255u) / 256u;
    auto go = [&](auto acc_tag, auto tin) {
        using Acc = decltype(acc_tag);
        using TIn = decltype(tin);
        return with_dtype(out_dt, [&](auto tout) {
            using TOut = decltype(tout);
            hipLaunchKernelGGL((axis_pass_kernel<TIn, TOut, Acc>), dim3(blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        });
    };
    if (accum == NDL_ACCUM_REF) return with_dtype(in_dt, [&](auto tin) { return go(double{}, tin); });
    // FP32 accumulation is offered for float storage only
    switch (in_dt) {
    case NDL_F32: return go(float{}, float{});
    case NDL_F16: return go(float{}, f16_t{});
    case NDL_BF16: return go(float{}, bf16_t{});
    }
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------ dense N-D
template <class TIn, class TOut>
__global__ __launch_bounds__(256) void convolve_nd_kernel(const NdArgs a)
{
    uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t coord[NDL_MAX_DIM];
    int64_t ooff = 0;
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        uint32_t q = a.fdiv[d].div(rem);
        coord[d] = (int64_t)(rem - q * (uint32_t)a.ext[d]);
        rem = q;
        ooff += coord[d] * a.ost[d];
    }
    double total = 0;
    for (int t = 0; t < a.ntaps; t++) {
        int64_t off = 0;
        for (int d = 0; d < a.nd; d++) {
            int64_t x = coord[d] + a.tap_off[t * a.nd + d];
            if (x < 0 || x >= a.ext[d]) x = border_map(a.border, a.ext[d], x);
            off += x * a.ist[d];
        }
        total = total + load_f64((const TIn*)a.in + off) * a.tap_val[t];
    }
    store_from_f64((TOut*)a.out + ooff, total, NDL_ROUND_TRUNC);
}

hipError_t launch_convolve_nd(const NdArgs& a, int in_dt, int out_dt, hipStream_t s)
{
    if (a.total == 0) return hipSuccess;
    uint32_t blocks = (a.total + 255u) / 256u;
    return with_dtype(in_dt, [&](auto tin) {
        using TIn = decltype(tin);
        return with_dtype(out_dt, [&](auto tout) {
            using TOut = decltype(tout);
            hipLaunchKernelGGL((convolve_nd_kernel<TIn, TOut>), dim3(blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        });
    });
}

// ---------------------------------------------------------- morphology
// Comparison key of a storage type: integers as themselves, half formats as
// float (exact and order-preserving); the winner is kept in its storage type
// so a same-type output copies it bit for bit (NaN payloads, signed zeros).
template <class T> struct MorphKey { using type = T; };
template <> struct MorphKey<f16_t> { using type = float; };
template <> struct MorphKey<bf16_t> { using type = float; };
template <class T> __device__ __forceinline__ typename MorphKey<T>::type morph_key(T v) { return v; }
template <> __device__ __forceinline__ float morph_key<f16_t>(f16_t v) { return f16_to_f32(v.bits); }
template <> __device__ __forceinline__ float morph_key<bf16_t>(bf16_t v) { return bf16_to_f32(v.bits); }
template <class K> __device__ __forceinline__ bool key_nan(K v)
{
    if constexpr (std::is_floating_point<K>::value) return v != v;
    else return false;
}

template <class TIn, class TOut> __device__ __forceinline__ void morph_store(TOut* p, TIn v)
{
    if constexpr (std::is_same<TIn, TOut>::value) *p = v;
    else store_from_f64(p, load_f64(&v), NDL_ROUND_TRUNC); // static_cast<DstT>(best)
}

template <class TIn, class TOut, bool MAX>
__global__ __launch_bounds__(256) void morph_pass_kernel(const MorphPassArgs a)
{
    using Key = typename MorphKey<TIn>::type;
    uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t ioff = 0, ooff = 0, coff = 0, j = 0;
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        uint32_t q = a.fdiv[d].div(rem);
        int64_t c = (int64_t)(rem - q * (uint32_t)a.ext[d]);
        rem = q;
        ioff += c * a.ist[d];
        ooff += c * a.ost[d];
        if (d == a.axis_pos) j = c;
        if (a.corner) {
            int64_t x = c + a.coff[d];
            if (x < 0 || x >= a.ext[d]) x = border_map(a.cbord[d], a.ext[d], x);
            coff += x * a.cst[d];
        }
    }
    const TIn* in = (const TIn*)a.in + (ioff - j * a.in_as);
    bool have = false;
    TIn best{};
    Key bk{};
    for (int t = 0; t < a.K; t++) {
        if (!(a.wide_mask ? a.wide_mask[t] : a.mask[t])) continue; // uniform
        int64_t x = j + t - a.center;
        if (x < 0 || x >= a.m) x = border_map(a.border, a.m, x);
        const TIn v = in[x * a.in_as];
        const Key k = morph_key(v);
        // strict comparison, first tap wins ties; NaN never wins over a number
        const bool take = !have || (key_nan(bk) ? !key_nan(k) : (MAX ? k > bk : k < bk));
        if (take) {
            best = v;
            bk = k;
            have = true;
        }
    }
    if constexpr (std::is_floating_point<Key>::value) {
        if (a.corner) {
            const TIn cv = ((const TIn*)a.corner)[coff];
            if (key_nan(morph_key(cv))) best = cv;
        }
    }
    morph_store<TIn, TOut>((TOut*)a.out + ooff, best);
}

// The same pass with a compile-time tap count K: NO outputs per thread
// (thread t takes outputs t, t + T, ..., T = all threads), every window value
// of all NO outputs loaded before the first comparison, border maps only for
// windows that reach past an end of the axis.  Comparison order, ties, NaN
// and the corner rule are those of morph_pass_kernel.
template <class TIn, class TOut, bool MAX, int K, int NO>
__global__ __launch_bounds__(256) void morph_pass_k_kernel(const MorphPassArgs a)
{
    using Key = typename MorphKey<TIn>::type;
    const uint32_t nthreads = gridDim.x * 256u;
    const uint32_t first = blockIdx.x * 256u + threadIdx.x;
    TIn v[NO][K];
    int64_t ooff[NO], coff[NO];
    bool live[NO];
#pragma unroll
    for (int o = 0; o < NO; o++) {
        const uint32_t idx = first + (uint32_t)o * nthreads;
        live[o] = idx < a.total;
        ooff[o] = 0;
        coff[o] = 0;
        if (!live[o]) continue;
        int64_t ioff = 0, j = 0;
        uint32_t rem = idx;
        for (int d = 0; d < a.nd; d++) {
            uint32_t q = a.fdiv[d].div(rem);
            int64_t c = (int64_t)(rem - q * (uint32_t)a.ext[d]);
            rem = q;
            ioff += c * a.ist[d];
            ooff[o] += c * a.ost[d];
            if (d == a.axis_pos) j = c;
            if (a.corner) {
                int64_t x = c + a.coff[d];
                if (x < 0 || x >= a.ext[d]) x = border_map(a.cbord[d], a.ext[d], x);
                coff[o] += x * a.cst[d];
            }
        }
        const TIn* in = (const TIn*)a.in + (ioff - j * a.in_as);
        const bool inside = j - a.center >= 0 && j + K - 1 - a.center < a.m;
#pragma unroll
        for (int t = 0; t < K; t++) {
            int64_t x = j + t - a.center;
            if (!inside) x = border_map(a.border, a.m, x);
            v[o][t] = in[x * a.in_as];
        }
    }
#pragma unroll
    for (int o = 0; o < NO; o++) {
        if (!live[o]) continue;
        TIn best = v[o][0];
        Key bk = morph_key(best);
#pragma unroll
        for (int t = 1; t < K; t++) {
            const Key k = morph_key(v[o][t]);
            // strict comparison, first tap wins ties; NaN never wins over a number
            const bool take = key_nan(bk) ? !key_nan(k) : (MAX ? k > bk : k < bk);
            if (take) {
                best = v[o][t];
                bk = k;
            }
        }
        if constexpr (std::is_floating_point<Key>::value) {
            if (a.corner) {
                const TIn cv = ((const TIn*)a.corner)[coff[o]];
                if (key_nan(morph_key(cv))) best = cv;
            }
        }
        morph_store<TIn, TOut>((TOut*)a.out + ooff[o], best);
    }
}

hipError_t launch_morph_pass(const MorphPassArgs& a, int in_dt, int out_dt, hipStream_t s)
{
    if (a.total == 0) return hipSuccess;
    // every tap included (box elements): compile-time tap counts
    bool all = a.K >= 1 && !a.wide_mask;
    for (int t = 0; t < a.K && !a.wide_mask; t++) all = all && a.mask[t];
    // (same-type passes and the common box sizes; the rest take the
    // run-time-tap kernel below)
    if (all && in_dt == out_dt && (a.K == 3 || a.K == 5 || a.K == 7)) {
        constexpr int NO = 4;
        const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, ((uint64_t)a.total + 256 * NO - 1) / (256 * NO));
        return with_dtype(in_dt, [&](auto tin) {
            using TIn = decltype(tin);
            {
                using TOut = TIn;
                switch (a.K) {
#define NDL_MK(K)                                                                                                   \
    case K:                                                                                                         \
        if (a.dilate) hipLaunchKernelGGL((morph_pass_k_kernel<TIn, TOut, true, K, NO>), dim3(blocks), dim3(256), 0, s, a); \
        else hipLaunchKernelGGL((morph_pass_k_kernel<TIn, TOut, false, K, NO>), dim3(blocks), dim3(256), 0, s, a); \
        return hipGetLastError();
                    NDL_MK(3) NDL_MK(5) NDL_MK(7)
#undef NDL_MK
                }
                return hipErrorInvalidValue;
            }
        });
    }
    uint32_t blocks = (a.total + 255u) / 256u;
    return with_dtype(in_dt, [&](auto tin) {
        using TIn = decltype(tin);
        return with_dtype(out_dt, [&](auto tout) {
            using TOut = decltype(tout);
            if (a.dilate) hipLaunchKernelGGL((morph_pass_kernel<TIn, TOut, true>), dim3(blocks), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((morph_pass_kernel<TIn, TOut, false>), dim3(blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        });
    });
}

// the reference's own loop: first included tap, then strictly smaller
// (erode) / larger (dilate) values replace it; SrcT{} without taps
template <class TIn, class TOut, bool MAX>
__global__ __launch_bounds__(256) void morph_nd_kernel(const NdArgs a)
{
    using Key = typename MorphKey<TIn>::type;
    uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t coord[NDL_MAX_DIM];
    int64_t ooff = 0;
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        uint32_t q = a.fdiv[d].div(rem);
        coord[d] = (int64_t)(rem - q * (uint32_t)a.ext[d]);
        rem = q;
        ooff += coord[d] * a.ost[d];
    }
    TIn best{};
    Key bk{};
    for (int t = 0; t < a.ntaps; t++) {
        int64_t off = 0;
        for (int d = 0; d < a.nd; d++) {
            int64_t x = coord[d] + a.tap_off[t * a.nd + d];
            if (x < 0 || x >= a.ext[d]) x = border_map(a.border, a.ext[d], x);
            off += x * a.ist[d];
        }
        const TIn v = ((const TIn*)a.in)[off];
        const Key k = morph_key(v);
        if (t == 0 || (MAX ? k > bk : k < bk)) {
            best = v;
            bk = k;
        }
    }
    morph_store<TIn, TOut>((TOut*)a.out + ooff, best);
}

hipError_t launch_morph_nd(const NdArgs& a, int dilate, int in_dt, int out_dt, hipStream_t s)
{
    if (a.total == 0) return hipSuccess;
    uint32_t blocks = (a.total + 255u) / 256u;
    return with_dtype(in_dt, [&](auto tin) {
        using TIn = decltype(tin);
        return with_dtype(out_dt, [&](auto tout) {
            using TOut = decltype(tout);
            if (dilate) hipLaunchKernelGGL((morph_nd_kernel<TIn, TOut, true>), dim3(blocks), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((morph_nd_kernel<TIn, TOut, false>), dim3(blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        });
    });
}

// ------------------------------------------------------- scale stacks
template <class TIn> __global__ __launch_bounds__(256) void axis_pass_multi_kernel(const MultiAxisArgs a)
{
    uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.p.total) return;
    int64_t ioff = 0, ooff = 0, j = 0;
    uint32_t rem = idx;
    for (int d = 0; d < a.p.nd; d++) {
        uint32_t q = a.p.fdiv[d].div(rem);
        int64_t c = (int64_t)(rem - q * (uint32_t)a.p.ext[d]);
        rem = q;
        ioff += c * a.p.ist[d];
        ooff += c * a.p.ost[d];
        if (d == a.p.axis_pos) j = c;
    }
    const TIn* in = (const TIn*)a.p.in + (ioff - j * a.p.in_as);
    for (int l = 0; l < a.nlev; l++) {
        const double* w = a.taps + (size_t)l * NDL_MAX_TAPS;
        double total = 0;
        for (int t = 0; t < a.K[l]; t++) {
            const double wt = w[t];
            if (wt == 0.0) continue;
            int64_t x = j + t - a.center[l];
            if (x < 0 || x >= a.p.m) x = border_map(a.p.border, a.p.m, x);
            total = total + load_f64(in + x * a.p.in_as) * wt; // the level's line stays in cache
        }
        ((double*)a.outs[l])[ooff] = total;
    }
}

hipError_t launch_axis_pass_multi(const MultiAxisArgs& a, int in_dt, hipStream_t s)
{
    if (a.p.total == 0) return hipSuccess;
    uint32_t blocks = (a.p.total + 255u) / 256u;
    return with_dtype(in_dt, [&](auto tin) {
        using TIn = decltype(tin);
        hipLaunchKernelGGL((axis_pass_multi_kernel<TIn>), dim3(blocks), dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

template <class TIn, class TOut> __global__ __launch_bounds__(256) void difference_kernel(const DiffArgs a)
{
    uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t ao = 0, bo = 0, oo = 0;
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        uint32_t q = a.fdiv[d].div(rem);
        int64_t c = (int64_t)(rem - q * (uint32_t)a.ext[d]);
        rem = q;
        ao += c * a.ast[d];
        bo += c * a.bst[d];
        oo += c * a.ost[d];
    }
    store_from_f64((TOut*)a.out + oo, load_f64((const TIn*)a.b + bo) - load_f64((const TIn*)a.a + ao),
                   NDL_ROUND_TRUNC);
}

hipError_t launch_difference(const DiffArgs& a, int in_dt, int out_dt, hipStream_t s)
{
    if (a.total == 0) return hipSuccess;
    uint32_t blocks = (a.total + 255u) / 256u;
    return with_dtype(in_dt, [&](auto tin) {
        using TIn = decltype(tin);
        return with_dtype(out_dt, [&](auto tout) {
            using TOut = decltype(tout);
            hipLaunchKernelGGL((difference_kernel<TIn, TOut>), dim3(blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        });
    });
}

// ---------------------------------------------------------- downsample
template <class T> __global__ __launch_bounds__(256) void downsample_kernel(const DownArgs a)
{
    uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t lo[NDL_MAX_DIM], hi[NDL_MAX_DIM], cur[NDL_MAX_DIM];
    int64_t ooff = 0, count = 1;
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        uint32_t q = a.fdiv[d].div(rem);
        const int64_t c = (int64_t)(rem - q * (uint32_t)a.oext[d]);
        rem = q;
        ooff += c * a.ost[d];
        if (d == a.channel_axis) {
            lo[d] = c;
            hi[d] = c + 1;
        } else {
            lo[d] = c * a.factor;
            hi[d] = min(lo[d] + a.factor, a.ext[d]);
        }
        cur[d] = lo[d];
        count *= hi[d] - lo[d];
    }
    double sum = 0.0;
    for (int64_t k = 0; k < count; k++) {
        int64_t off = 0;
        for (int d = 0; d < a.nd; d++) off += cur[d] * a.ist[d];
        sum = sum + load_f64((const T*)a.in + off);
        for (int d = 0; d < a.nd; d++) { // block odometer, dimension 0 fastest
            if (d == a.channel_axis) continue;
            if (++cur[d] < hi[d]) break;
            cur[d] = lo[d];
        }
    }
    store_from_f64((T*)a.out + ooff, sum / (double)count, NDL_ROUND_NEAREST);
}

hipError_t launch_downsample(const DownArgs& a, int dt, hipStream_t s)
{
    if (a.total == 0) return hipSuccess;
    uint32_t blocks = (a.total + 255u) / 256u;
    return with_dtype(dt, [&](auto t) {
        using T = decltype(t);
        hipLaunchKernelGGL((downsample_kernel<T>), dim3(blocks), dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

// ------------------------------------------------- blur then downsample
// gaussian_blur's last pass and final store (convolution.h:171-185) at every
// source position of an output's block, then downsample's block mean
// (convolution.h:270-300): the blurred image is never written.
template <class Acc> __device__ __forceinline__ Acc bd_tap(const BlurDownArgs& a, int t);
template <> __device__ __forceinline__ double bd_tap<double>(const BlurDownArgs& a, int t)
{
    return a.wide_d ? a.wide_d[t] : a.taps_d[t];
}
template <> __device__ __forceinline__ float bd_tap<float>(const BlurDownArgs& a, int t)
{
    return a.wide_f ? a.wide_f[t] : a.taps_f[t];
}

// Each lane marches one column of outputs along the pass axis (a chunk of
// tz of them), so the inputs of consecutive outputs, which overlap by all but
// `factor` slices, are re-read from the lane's own recent loads in cache.
template <class TIn, class TOut, class Acc>
__global__ __launch_bounds__(256) void blur_downsample_kernel(const BlurDownArgs a)
{
    const uint32_t chunk = blockIdx.x % (uint32_t)a.tiles_z;
    const uint32_t lane = blockIdx.x / (uint32_t)a.tiles_z * 256u + threadIdx.x;
    if (lane >= a.lanes) return;
    int64_t oc[NDL_MAX_DIM];
    uint32_t rem = lane;
    for (int i = 0; i < a.nl; i++) {
        const int d = a.ldim[i];
        const uint32_t q = a.lfd[i].div(rem);
        oc[d] = (int64_t)(rem - q * (uint32_t)a.oext[d]);
        rem = q;
    }
    const int64_t m = a.ext[a.axis], as = a.ist[a.axis];
    const int64_t oz0 = (int64_t)chunk * a.tz, oz1 = min(oz0 + a.tz, a.oext[a.axis]);
    for (int64_t oz = oz0; oz < oz1; oz++) {
        oc[a.axis] = oz;
        int64_t lo[NDL_MAX_DIM], hi[NDL_MAX_DIM], cur[NDL_MAX_DIM];
        int64_t ooff = 0, count = 1;
        for (int d = 0; d < a.nd; d++) {
            const int64_t c = oc[d];
            ooff += c * a.ost[d];
            if (d == a.channel_axis) {
                lo[d] = c;
                hi[d] = c + 1;
            } else {
                lo[d] = c * a.factor;
                hi[d] = min(lo[d] + a.factor, a.ext[d]);
            }
            cur[d] = lo[d];
            count *= hi[d] - lo[d];
        }
        double sum = 0.0;
        for (int64_t n = 0; n < count; n++) {
            int64_t off = 0;
            for (int d = 0; d < a.nd; d++)
                if (d != a.axis) off += cur[d] * a.ist[d];
            const TIn* line = (const TIn*)a.in + off;
            const int64_t j = cur[a.axis];
            Acc total = 0;
            for (int t = 0; t < a.K; t++) {
                const Acc w = bd_tap<Acc>(a, t);
                if (w == (Acc)0) continue;
                int64_t x = j + t - a.center;
                if (x < 0 || x >= m) x = border_map(a.border, m, x);
                total = mac<Acc>(total, load_acc<Acc>(line + x * as), w);
            }
            TOut blurred; // the value gaussian_blur would have stored
            store_acc<TOut, Acc>(&blurred, total, a.rounding);
            sum = sum + load_f64(&blurred);
            for (int d = 0; d < a.nd; d++) { // block odometer, dimension 0 fastest
                if (d == a.channel_axis) continue;
                if (++cur[d] < hi[d]) break;
                cur[d] = lo[d];
            }
        }
        store_from_f64((TOut*)a.out + ooff, sum / (double)count, NDL_ROUND_NEAREST);
    }
}

hipError_t launch_blur_downsample(const BlurDownArgs& a, int in_dt, int out_dt, int accum, hipStream_t s)
{
    const uint64_t blocks = (uint64_t)a.tiles_z * ((a.lanes + 255u) / 256u);
    if (blocks == 0) return hipSuccess;
    if (blocks > 0xffffffffull) return hipErrorInvalidValue;
    auto go = [&](auto acc_tag, auto tin) {
        using Acc = decltype(acc_tag);
        using TIn = decltype(tin);
        return with_dtype(out_dt, [&](auto tout) {
            using TOut = decltype(tout);
            hipLaunchKernelGGL((blur_downsample_kernel<TIn, TOut, Acc>), dim3((uint32_t)blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        });
    };
    if (accum == NDL_ACCUM_REF) return with_dtype(in_dt, [&](auto tin) { return go(double{}, tin); });
    switch (in_dt) {
    case NDL_F32: return go(float{}, float{});
    case NDL_F16: return go(float{}, f16_t{});
    case NDL_BF16: return go(float{}, bf16_t{});
    }
    return hipErrorInvalidValue;
}

// ------------------------------------------------------ tap gradients
// Tap gradient of one axis: grad[t] = sum_p g[p] * U[p with its coordinate
// j along the axis replaced by B(j + t - c)].  Work item = (segment of
// kTapSeg consecutive positions along the axis, line); with SEG, lanes take
// consecutive lines (coalesced along the fastest other dimension) and share
// the segment, so the border test is uniform.  A thread loads the U window
// of its segment once (kTapSeg + K - 1 values) and its kTapSeg g values, and
// keeps KB >= K FP64 accumulators in registers (taps t >= K are skipped by a
// uniform test).  g and U convert exactly; one partial per (workgroup, tap),
// reduced by wave shuffles then across the four waves in LDS;
// tapgrad_final_kernel sums the partials in a fixed tree order, so the result
// does not depend on scheduling.
// SEG = false (the axis is the fastest dimension of g): lanes take
// consecutive segments of one line instead, so each lane reads a contiguous
// run and the wave a contiguous block.
template <class TG, class TU, int KB, bool SEG>
__global__ __launch_bounds__(256) void tapgrad_partial_kernel(const TapGradArgs a)
{
    constexpr int NW = kTapSeg + KB - 1;
    double acc[KB];
#pragma unroll
    for (int t = 0; t < KB; t++) acc[t] = 0.0;
    const uint32_t step = 256u * gridDim.x;
    for (uint32_t item = blockIdx.x * 256u + threadIdx.x; item < a.items; item += step) {
        uint32_t seg, rem;
        if constexpr (SEG) {
            seg = a.ldiv.div(item);
            rem = item - seg * a.lines;
        } else {
            rem = a.ldiv.div(item);
            seg = item - rem * a.nseg;
        }
        int64_t goff = 0, uoff = 0;
        for (int d = 0; d < a.nl; d++) {
            uint32_t q = a.fdiv[d].div(rem);
            int64_t c = (int64_t)(rem - q * (uint32_t)a.ext[d]);
            rem = q;
            goff += c * a.gst[d];
            uoff += c * a.ust[d];
        }
        const int64_t j0 = (int64_t)seg * kTapSeg;
        const TU* ub = (const TU*)a.u + uoff;
        const TG* gb = (const TG*)a.g + goff;
        // window element i is U at j0 + i - c, border-mapped near the ends;
        // positions past m - 1 + (K-1-c) are read only by outputs past the
        // end of the axis (never accumulated) and are clamped to stay in
        // bounds
        using W = std::conditional_t<sizeof(TU) == 8, double, float>;
        W win[NW];
        const bool inside = j0 - a.center >= 0 && j0 + NW - 1 - a.center < a.m;
        const int64_t last = a.m - 1 + (a.K - 1 - a.center);
#pragma unroll
        for (int i = 0; i < NW; i++) {
            win[i] = W(0);
            if (i < kTapSeg + a.K - 1) {
                int64_t x = j0 + i - a.center;
                if (!inside) x = x > last ? a.m - 1 : border_map(a.border, a.m, x);
                win[i] = ub[x * a.u_as];
            }
        }
#pragma unroll
        for (int r = 0; r < kTapSeg; r++) {
            if (j0 + r < a.m) {
                const double gv = load_f64(gb + (j0 + r) * a.g_as);
#pragma unroll
                for (int t = 0; t < KB; t++)
                    if (t < a.K) acc[t] = acc[t] + gv * (double)win[r + t];
            }
        }
    }
    __shared__ double red[4][KB];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < KB; t++) {
        if (t < a.K) {
            double v = acc[t];
            for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
            if (lane == 0) red[wv][t] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < a.K) {
        const int t = threadIdx.x;
        a.partial[(size_t)blockIdx.x * a.K + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    }
}

// Orders one wave's LDS writes before its later reads by other lanes (and
// those reads before the next writes): the fences keep the compiler from
// moving LDS accesses across; the hardware runs one wave's LDS instructions
// in order.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tap gradient with a compile-time tap count K (odd K <= 25, the common
// filters); same contract and result layout as tapgrad_partial_kernel.
// SEG: as there, a thread walks a segment of kTapSeg positions along the
// axis with its U window in registers, lanes on consecutive lines (each
// window load coalesced across the wave); element offsets inside an item are
// 32-bit (the host checks the span).  !SEG (the axis is g's fastest
// dimension): a wave takes 64 consecutive positions of one line, one per
// lane, and each tap's U load is a contiguous run shifted by the tap (the
// K loads of a lane hit the same cache lines as its neighbours').  FP64
// fused multiply-adds, taps ascending per position.
template <class TG, class TU, int K, bool SEG>
__global__ __launch_bounds__(256) void tapgrad_exact_kernel(const TapGradArgs a)
{
    double acc[K];
#pragma unroll
    for (int t = 0; t < K; t++) acc[t] = 0.0;
    const int32_t u_as = (int32_t)a.u_as, g_as = (int32_t)a.g_as;
    const int32_t m = (int32_t)a.m, c = a.center;
    if constexpr (SEG) {
        constexpr int NW = kTapSeg + K - 1;
        const uint32_t step = 256u * gridDim.x;
        for (uint32_t item = blockIdx.x * 256u + threadIdx.x; item < a.items; item += step) {
            const uint32_t seg = a.ldiv.div(item);
            uint32_t rem = item - seg * a.lines;
            int64_t goff = 0, uoff = 0;
            for (int d = 0; d < a.nl; d++) {
                uint32_t q = a.fdiv[d].div(rem);
                int64_t cc = (int64_t)(rem - q * (uint32_t)a.ext[d]);
                rem = q;
                goff += cc * a.gst[d];
                uoff += cc * a.ust[d];
            }
            const int32_t j0 = (int32_t)seg * kTapSeg;
            const TU* ub = (const TU*)a.u + uoff;
            const TG* gb = (const TG*)a.g + goff;
            const bool inside = j0 - c >= 0 && j0 + NW - 1 - c < m;
            const int32_t last = m - 1 + (K - 1 - c);
            float win[NW];
#pragma unroll
            for (int i = 0; i < NW; i++) {
                int32_t x = j0 + i - c;
                // positions past last are read only by outputs past the end
                // of the axis (never accumulated): clamped to stay in bounds
                if (!inside) x = x > last ? m - 1 : border_map32(a.border, m, x);
                win[i] = (float)ub[x * u_as];
            }
#pragma unroll
            for (int r = 0; r < kTapSeg; r++) {
                if (j0 + r < m) {
                    const double gv = load_f64(gb + (j0 + r) * g_as);
#pragma unroll
                    for (int t = 0; t < K; t++) acc[t] = __builtin_fma(gv, (double)win[r + t], acc[t]);
                }
            }
        }
    } else {
        // wave items: (line, 64 consecutive positions).  The wave stages the
        // item's U window (64 + K - 1 positions, border-mapped) in its own
        // LDS row with one coalesced load per position; each lane then reads
        // its K window values from LDS (wave_lds_sync between the two).
        __shared__ float wrow[4][64 + 32];
        const uint32_t nchunk = (uint32_t)((m + 63) >> 6);
        const uint32_t witems = a.lines * nchunk;
        const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6), nwaves = gridDim.x * 4u;
        const int32_t lane = threadIdx.x & 63;
        float* row = wrow[threadIdx.x >> 6];
        const int32_t last = m - 1 + (K - 1 - c);
        for (uint32_t it = wave; it < witems; it += nwaves) {
            const uint32_t line = it / nchunk, chunk = it - line * nchunk;
            uint32_t rem = line;
            int64_t goff = 0, uoff = 0;
            for (int d = 0; d < a.nl; d++) {
                uint32_t q = a.fdiv[d].div(rem);
                int64_t cc = (int64_t)(rem - q * (uint32_t)a.ext[d]);
                rem = q;
                goff += cc * a.gst[d];
                uoff += cc * a.ust[d];
            }
            const int32_t j0 = (int32_t)chunk * 64;
            const TU* ub = (const TU*)a.u + uoff;
            wave_lds_sync();
#pragma unroll
            for (int i0 = 0; i0 < 64 + K - 1; i0 += 64) {
                const int32_t i = i0 + lane;
                if (i < 64 + K - 1) {
                    int32_t x = j0 + i - c;
                    // positions past last feed only outputs past the end of
                    // the axis (never accumulated): clamped to stay in bounds
                    if (x < 0 || x >= m) x = x > last ? m - 1 : border_map32(a.border, m, x);
                    row[i] = (float)ub[x * u_as];
                }
            }
            wave_lds_sync();
            const int32_t j = j0 + lane;
            if (j < m) {
                const double gv = load_f64((const TG*)a.g + goff + (int64_t)j * g_as);
#pragma unroll
                for (int t = 0; t < K; t++) acc[t] = __builtin_fma(gv, (double)row[lane + t], acc[t]);
            }
        }
    }
    __shared__ double red[4][K];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < K; t++) {
        double v = acc[t];
        for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
        if (lane == 0) red[wv][t] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < K) {
        const int t = threadIdx.x;
        a.partial[(size_t)blockIdx.x * K + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    }
}

// one workgroup per tap: fixed-order tree over the workgroup partials
__global__ __launch_bounds__(256) void tapgrad_final_kernel(const double* partial, int nblocks, int K, double* out)
{
    __shared__ double red[256];
    const int t = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) s = s + partial[(size_t)b * K + t];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = red[0];
}

hipError_t launch_tapgrad_partial(const TapGradArgs& a, int g_dt, int u_is_f64, hipStream_t s)
{
    if (a.K < 1 || a.K > 64) return hipErrorInvalidValue;
    // float U (FP32 accumulation), float g: exact-K instances when the
    // 32-bit item offsets cannot overflow
    const int64_t span = (a.m + kTapSeg + a.K) * (std::max(std::llabs(a.u_as), std::llabs(a.g_as)) + 1);
    if (!u_is_f64 && g_dt == NDL_F32 && (a.K & 1) && a.K <= 25 && span < 0x7fffffffLL) {
        switch (a.K) {
#define NDL_TGX(K)                                                                                           \
    case K:                                                                                                  \
        if (a.seg) hipLaunchKernelGGL((tapgrad_exact_kernel<float, float, K, true>), dim3(a.nblocks), dim3(256), 0, s, a); \
        else hipLaunchKernelGGL((tapgrad_exact_kernel<float, float, K, false>), dim3(a.nblocks), dim3(256), 0, s, a); \
        return hipGetLastError();
            NDL_TGX(1) NDL_TGX(3) NDL_TGX(5) NDL_TGX(7) NDL_TGX(9) NDL_TGX(11) NDL_TGX(13) NDL_TGX(15) NDL_TGX(17)
            NDL_TGX(19) NDL_TGX(21) NDL_TGX(23) NDL_TGX(25)
#undef NDL_TGX
        }
    }
    auto go = [&](auto tu) {
        using TU = decltype(tu);
        return with_dtype(g_dt, [&](auto tg) {
            using TG = decltype(tg);
#define NDL_TG(KB)                                                                                          \
    if (a.K <= KB) {                                                                                        \
        if (a.seg)                                                                                          \
            hipLaunchKernelGGL((tapgrad_partial_kernel<TG, TU, KB, true>), dim3(a.nblocks), dim3(256), 0, s, a); \
        else                                                                                                \
            hipLaunchKernelGGL((tapgrad_partial_kernel<TG, TU, KB, false>), dim3(a.nblocks), dim3(256), 0, s, a); \
        return hipGetLastError();                                                                           \
    }
            NDL_TG(4) NDL_TG(8) NDL_TG(16) NDL_TG(32) NDL_TG(64)
#undef NDL_TG
            return hipErrorInvalidValue;
        });
    };
    return u_is_f64 ? go(double{}) : go(float{});
}

hipError_t launch_tapgrad_final(const double* partial, int nblocks, int K, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(tapgrad_final_kernel, dim3(K), dim3(256), 0, s, partial, nblocks, K, out);
    return hipGetLastError();
}

} // namespace ndl
