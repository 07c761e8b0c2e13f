// kernels_xyref.hip -- two consecutive passes of gaussian_blur()'s FP64 chain
// (processing/convolution.h:151-185: axis 0 into an FP64 buffer, the next
// axis from it) in one launch, the intermediate kept in LDS instead of HBM.
//
// Arithmetic (this file is built with -ffp-contract=off): per output and
// pass, total = 0, then total = total + double(x) * tap for the taps in
// ascending order, every product and sum rounded separately (the reference
// skips exactly-zero taps, detail.h:164-174; kernels with one are not sent
// here); the x pass result is an FP64 value, as the reference's buffer
// holds; the y pass result is rounded once into the output type.
// Bit-identical to the two passes of the chain.
//
// A workgroup (4 waves) owns a 64 x 32 tile of outputs (x by y) of nsl
// consecutive indices of the other dimensions, one after the other (the
// window of the next index is requested before the current one is filtered):
//  1. the input window, (32 + K - 1) rows of (64 + K - 1) elements with the
//     border folded per coordinate, is read into LDS in the input type (wave
//     w takes rows w, w + 4, ...; a lane one element, two when the row is
//     wider than 64; every load of the wave issued before the first store);
//  2. the x pass over every window row, four consecutive outputs per task
//     from a register window of K + 3 values, into an FP64 LDS plane;
//  3. the y pass, four consecutive rows of one column per task from a
//     register window of K + 3 values of that plane, stored to the output,
//     or (a blur followed by downsample(), fx / fy > 1) rounded as the blur
//     stores it and reduced to block means in the downsample's order.
// The tap count K (odd, 3..25, the same on both axes, as every
// gaussian_blur() has) is a template parameter: the windows are registers
// and the tap loops carry no branch.
#include "ndl_internal.hpp"

namespace ndl {
namespace {

constexpr int TXW = kXYTileX, TYH = kXYTileY;

// downsample()'s block means (convolution.h:270-300) of the y pass results
// of one task: rows 4s .. 4s+3 of column `lane` hold acc[0..3]; blocks of FY
// rows (FY divides 4) by fx columns (adjacent lanes), each value rounded as
// gaussian_blur stores it, summed in block order (x fastest, then y) by the
// block's first lane, which stores the mean.
template <class TOut, int FY>
__device__ __forceinline__ void store_block_means(const XYRefArgs& a, const double* acc, int x0, int ybase, int lane,
                                                  int64_t ob)
{
    double vr[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        TOut b;
        store_from_f64(&b, acc[j], a.rounding);
        vr[j] = load_f64(&b);
    }
    const int fx = a.fx;
    const int xb = lane - lane % fx;
    const int nx = min(fx, a.X - (x0 + xb));
#pragma unroll
    for (int b = 0; b < 4; b += FY) {
        double sum = 0.0;
#pragma unroll
        for (int j = b; j < b + FY; j++) {
            const bool yok = ybase + j < a.Y;
            for (int dx = 0; dx < fx; dx++) {
                const double v = __shfl(vr[j], xb + dx);
                if (yok && dx < nx) sum = sum + v;
            }
        }
        const int ny = min(FY, a.Y - (ybase + b));
        if (lane == xb && nx > 0 && ny > 0)
            store_from_f64((TOut*)a.out + ob + (int64_t)((x0 + xb) / fx) * a.osx + (int64_t)((ybase + b) / FY) * a.osy,
                           sum / (double)(nx * ny), NDL_ROUND_NEAREST);
    }
}

template <class TIn, class TOut, int K, bool DEC>
__global__ __launch_bounds__(256) void xy_ref_kernel(const XYRefArgs a)
{
    extern __shared__ __align__(16) unsigned char xy_lds[];
    constexpr int WA = TXW + K - 1, HA = TYH + K - 1;
    double* plane = (double*)xy_lds;                                 // HA x TXW, x pass results
    TIn* win = (TIn*)(xy_lds + (size_t)HA * TXW * sizeof(double));   // HA x WA, input window
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    uint32_t t = blockIdx.x;
    uint32_t q = a.ftx.div(t);
    const int x0 = (int)(t - q * (uint32_t)a.tiles_x) * TXW;
    t = q;
    q = a.fty.div(t);
    const int y0 = (int)(t - q * (uint32_t)a.tiles_y) * TYH;
    // this workgroup's indices of the other dimensions: b0 .. b1-1 of
    // their flattened range (dimension order as listed by the host)
    const uint32_t b0 = q * a.nsl, b1 = min(b0 + a.nsl, a.nbatch);
    auto batch_off = [&](uint32_t bb, int64_t& ib, int64_t& ob) {
        ib = 0;
        ob = 0;
        for (int i = 0; i < a.nb; i++) {
            const uint32_t qq = a.bext[i].div(bb);
            const uint32_t r = bb - qq * a.bext[i].d;
            ib += (int64_t)r * a.ib[i];
            ob += (int64_t)r * a.ob[i];
            bb = qq;
        }
    };

    // 1. input window: lanes past the window's right end repeat its last
    //    element (same address, same value), so no load or store is
    //    predicated per lane; rows fold only in tiles that reach past the
    //    first or last row.  Column and row offsets are the same for every
    //    index of the other dimensions: computed once, and the window of the
    //    next index is requested while the current one is filtered.
    constexpr int RW = (HA + 3) / 4; // window rows per wave, at most
    const int64_t xo0 = (int64_t)safe_map32(a.bx, a.X, x0 - a.cx + lane) * a.isx;
    const int c1 = min(lane + 64, WA - 1);
    const int64_t xo1 = (int64_t)safe_map32(a.bx, a.X, x0 - a.cx + c1) * a.isx;
    int64_t roff[RW];
    {
        const int ybase = y0 - a.cy;
        const bool yin = ybase >= 0 && ybase + HA <= a.Y;
#pragma unroll
        for (int j = 0; j < RW; j++) {
            const int r = min(wv + 4 * j, HA - 1);
            roff[j] = (int64_t)(yin ? ybase + r : safe_map32(a.by, a.Y, ybase + r)) * a.isy;
        }
    }
    TIn e0[RW], e1[RW];
    auto issue = [&](const TIn* in) {
#pragma unroll
        for (int j = 0; j < RW; j++) {
            if (wv + 4 * j < HA) {
                e0[j] = in[roff[j] + xo0];
                e1[j] = in[roff[j] + xo1];
            }
        }
    };
    int64_t ib, ob;
    batch_off(b0, ib, ob);
    issue((const TIn*)a.in + ib);
    for (uint32_t bb = b0; bb < b1; bb++) {
#pragma unroll
        for (int j = 0; j < RW; j++) {
            const int r = wv + 4 * j;
            if (r < HA) {
                win[r * WA + lane] = e0[j];
                win[r * WA + c1] = e1[j];
            }
        }
        __syncthreads();
        const int64_t obc = ob;
        if (bb + 1 < b1) {
            batch_off(bb + 1, ib, ob);
            issue((const TIn*)a.in + ib);
        }

        // 2. x pass: task (window row r, outputs 4s .. 4s+3)
        for (int r = threadIdx.x >> 4; r < HA; r += 16) {
            const int s = threadIdx.x & 15;
            const TIn* src = win + r * WA + 4 * s;
            double v[K + 3];
#pragma unroll
            for (int i = 0; i < K + 3; i++) v[i] = load_f64(src + i);
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double w = a.wx[k];
#pragma unroll
                for (int j = 0; j < 4; j++) acc[j] = acc[j] + v[k + j] * w;
            }
#pragma unroll
            for (int j = 0; j < 4; j++) plane[r * TXW + 4 * s + j] = acc[j];
        }
        __syncthreads();

        // 3. y pass: task (column lane, rows 4s .. 4s+3)
        const int x = x0 + lane;
        for (int s = wv; s < TYH / 4; s += 4) {
            const double* src = plane + 4 * s * TXW + lane;
            double v[K + 3];
#pragma unroll
            for (int i = 0; i < K + 3; i++) v[i] = src[i * TXW];
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double w = a.wy[k];
#pragma unroll
                for (int j = 0; j < 4; j++) acc[j] = acc[j] + v[k + j] * w;
            }
            if constexpr (DEC) {
                if (a.fy == 1) store_block_means<TOut, 1>(a, acc, x0, y0 + 4 * s, lane, obc);
                else if (a.fy == 2) store_block_means<TOut, 2>(a, acc, x0, y0 + 4 * s, lane, obc);
                else store_block_means<TOut, 4>(a, acc, x0, y0 + 4 * s, lane, obc);
            } else if (x < a.X) {
                TOut* col = (TOut*)a.out + obc + (int64_t)x * a.osx;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int y = y0 + 4 * s + j;
                    if (y < a.Y) store_from_f64(col + (int64_t)y * a.osy, acc[j], a.rounding);
                }
            }
        }
    }
}

// The same two passes preceded by a pass along a short axis c (at most four
// elements: the channels of a colour image), all three in the reference's
// order and arithmetic (the chain's c pass writes an FP64 image, x and y
// read it).  The tile loads the raw window of every channel once; for each
// output channel it forms the c-pass values of the window in an FP64 LDS
// window (taps ascending, the short-axis kernel's border rule: position
// ch + t - cc mapped once, a position past the axis reading element 0, a
// mapped position 1..3 beyond the extent reading zero), then runs the x and
// y passes of xy_ref_kernel on it.  No FP64 intermediate reaches HBM.  DEC:
// the y pass results go to downsample()'s block means over x and y as in
// xy_ref_kernel (channels are not decimated).
template <class TIn, class TOut, int K, bool DEC>
__global__ __launch_bounds__(256) void cxy_ref_kernel(const CXYRefArgs ca)
{
    extern __shared__ __align__(16) unsigned char cxy_lds[];
    const XYRefArgs& a = ca.xy;
    constexpr int WA = TXW + K - 1, HA = TYH + K - 1;
    double* plane = (double*)cxy_lds;                                       // HA x TXW
    double* cw = plane + HA * TXW;                                          // HA x WA, c-pass values
    TIn* win = (TIn*)(cxy_lds + (size_t)(HA * TXW + HA * WA) * sizeof(double)); // m x HA x WA, raw
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = ca.m;

    uint32_t t = blockIdx.x;
    uint32_t q = a.ftx.div(t);
    const int x0 = (int)(t - q * (uint32_t)a.tiles_x) * TXW;
    t = q;
    q = a.fty.div(t);
    const int y0 = (int)(t - q * (uint32_t)a.tiles_y) * TYH;
    t = q;
    int64_t ib = 0, ob = 0;
    for (int i = 0; i < a.nb; i++) {
        q = a.bext[i].div(t);
        const uint32_t r = t - q * a.bext[i].d;
        ib += (int64_t)r * a.ib[i];
        ob += (int64_t)r * a.ob[i];
        t = q;
    }

    // 1. raw windows of every channel (as xy_ref_kernel: lanes past the right
    //    end repeat the last element, rows fold only near the first or last row)
    {
        constexpr int RW = (HA + 3) / 4;
        const int64_t xo0 = (int64_t)safe_map32(a.bx, a.X, x0 - a.cx + lane) * a.isx;
        const int c1 = min(lane + 64, WA - 1);
        const int64_t xo1 = (int64_t)safe_map32(a.bx, a.X, x0 - a.cx + c1) * a.isx;
        const int ybase = y0 - a.cy;
        const bool yin = ybase >= 0 && ybase + HA <= a.Y;
        // every channel's loads are issued before the first LDS store
        TIn e0[kCXYMaxC][RW], e1[kCXYMaxC][RW];
#pragma unroll
        for (int ch = 0; ch < kCXYMaxC; ch++) {
            if (ch < m) {
                const TIn* in = (const TIn*)a.in + ib + (int64_t)ch * ca.ics;
#pragma unroll
                for (int j = 0; j < RW; j++) {
                    const int r = wv + 4 * j;
                    if (r < HA) {
                        const int yy = yin ? ybase + r : safe_map32(a.by, a.Y, ybase + r);
                        const TIn* row = in + (int64_t)yy * a.isy;
                        e0[ch][j] = row[xo0];
                        e1[ch][j] = row[xo1];
                    }
                }
            }
        }
#pragma unroll
        for (int ch = 0; ch < kCXYMaxC; ch++) {
            if (ch < m) {
                TIn* w = win + ch * (HA * WA);
#pragma unroll
                for (int j = 0; j < RW; j++) {
                    const int r = wv + 4 * j;
                    if (r < HA) {
                        w[r * WA + lane] = e0[ch][j];
                        w[r * WA + c1] = e1[ch][j];
                    }
                }
            }
        }
    }
    __syncthreads();

    const int x = x0 + lane;
    for (int ch = 0; ch < m; ch++) {
        // 1b. c pass of output channel ch over the window; the source channel
        //     of every tap is the same for the whole workgroup
        int soff[K];
        bool zero[K], any_zero = false;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = border_map32(ca.cb, m, ch + k - ca.cc);
            const bool mid = e >= 1 && e <= 3;
            soff[k] = mid && e < m ? e * (HA * WA) : 0;
            zero[k] = mid && e >= m;
            any_zero = any_zero || zero[k];
        }
        auto cpass = [&](auto with_zero) {
            for (int idx = threadIdx.x; idx < HA * WA; idx += 256) {
                TIn raw[K];
#pragma unroll
                for (int k = 0; k < K; k++) raw[k] = win[soff[k] + idx];
                double total = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    double v = load_f64(&raw[k]);
                    if constexpr (decltype(with_zero)::value) v = zero[k] ? 0.0 : v;
                    total = total + v * ca.wc[k];
                }
                cw[idx] = total;
            }
        };
        // a tap past the axis reads zero only when its reach exceeds the
        // extent, which the reference's argument check rejects: the common
        // loop carries no select
        if (any_zero) cpass(std::true_type{});
        else cpass(std::false_type{});
        __syncthreads();

        // 2. x pass: task (window row r, outputs 4s .. 4s+3)
        for (int r = threadIdx.x >> 4; r < HA; r += 16) {
            const int s = threadIdx.x & 15;
            const double* src = cw + r * WA + 4 * s;
            double v[K + 3];
#pragma unroll
            for (int i = 0; i < K + 3; i++) v[i] = src[i];
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double w = a.wx[k];
#pragma unroll
                for (int j = 0; j < 4; j++) acc[j] = acc[j] + v[k + j] * w;
            }
#pragma unroll
            for (int j = 0; j < 4; j++) plane[r * TXW + 4 * s + j] = acc[j];
        }
        __syncthreads();

        // 3. y pass: task (column lane, rows 4s .. 4s+3), stored into channel ch
        for (int s = wv; s < TYH / 4; s += 4) {
            const double* src = plane + 4 * s * TXW + lane;
            double v[K + 3];
#pragma unroll
            for (int i = 0; i < K + 3; i++) v[i] = src[i * TXW];
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double w = a.wy[k];
#pragma unroll
                for (int j = 0; j < 4; j++) acc[j] = acc[j] + v[k + j] * w;
            }
            const int64_t obch = ob + (int64_t)ch * ca.ocs;
            if constexpr (DEC) {
                if (a.fy == 1) store_block_means<TOut, 1>(a, acc, x0, y0 + 4 * s, lane, obch);
                else if (a.fy == 2) store_block_means<TOut, 2>(a, acc, x0, y0 + 4 * s, lane, obch);
                else store_block_means<TOut, 4>(a, acc, x0, y0 + 4 * s, lane, obch);
            } else if (x < a.X) {
                TOut* col = (TOut*)a.out + obch + (int64_t)x * a.osx;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int y = y0 + 4 * s + j;
                    if (y < a.Y) store_from_f64(col + (int64_t)y * a.osy, acc[j], a.rounding);
                }
            }
        }
        // the next channel's c pass rewrites cw only after every wave has
        // passed the barrier behind the x pass; its x pass rewrites the
        // plane only after the barrier behind that c pass, which every wave
        // reaches after its y pass
    }
}

template <int V> using IC = std::integral_constant<int, V>;

// f(TIn, TOut, K): uint8 / uint16 / int16 / float32 / float64 input, the
// output FP64 or the input's type (from float64: also uint8 / uint16 /
// int16 / float32), odd K 3..25
template <class F> hipError_t with_xy_instance(int in_dt, int out_dt, int k, F&& f)
{
    auto pick_k = [&](auto tin, auto tout) -> hipError_t {
        switch (k) {
#define NDL_K(KK) case KK: return f(tin, tout, IC<KK>{});
            NDL_K(3) NDL_K(5) NDL_K(7) NDL_K(9) NDL_K(11) NDL_K(13) NDL_K(15) NDL_K(17) NDL_K(19) NDL_K(21)
            NDL_K(23) NDL_K(25)
#undef NDL_K
        }
        return hipErrorInvalidValue;
    };
    auto pick = [&](auto tin) -> hipError_t {
        if (out_dt == NDL_F64) return pick_k(tin, double{});
        if (out_dt == in_dt) return pick_k(tin, tin);
        return hipErrorInvalidValue;
    };
    switch (in_dt) {
    case NDL_U8: return pick(uint8_t{});
    case NDL_U16: return pick(uint16_t{});
    case NDL_I16: return pick(int16_t{});
    case NDL_F32: return pick(float{});
    case NDL_F64:
        // an FP64 intermediate into the chain's final image
        switch (out_dt) {
        case NDL_U8: return pick_k(double{}, uint8_t{});
        case NDL_U16: return pick_k(double{}, uint16_t{});
        case NDL_I16: return pick_k(double{}, int16_t{});
        case NDL_F32: return pick_k(double{}, float{});
        }
        return pick(double{});
    }
    return hipErrorInvalidValue;
}

} // namespace

size_t xy_ref_lds_bytes(int kx, int ky, int in_size)
{
    const size_t ha = (size_t)(TYH + ky - 1);
    return ha * TXW * sizeof(double) + ha * (size_t)(TXW + kx - 1) * (size_t)in_size;
}

size_t cxy_ref_lds_bytes(int k, int in_size)
{
    const size_t ha = (size_t)(TYH + k - 1), wa = (size_t)(TXW + k - 1);
    return ha * TXW * sizeof(double) + ha * wa * sizeof(double) + (size_t)kCXYMaxC * ha * wa * (size_t)in_size;
}

hipError_t launch_cxy_ref(const CXYRefArgs& a, int in_dt, int out_dt, hipStream_t s)
{
    if (a.xy.ntiles == 0) return hipSuccess;
    if (a.xy.kx != a.xy.ky || a.m < 1 || a.m > kCXYMaxC || a.xy.kx > 9) return hipErrorInvalidValue;
    if (out_dt != in_dt && out_dt != NDL_F64) return hipErrorInvalidValue;
    const bool dec = a.xy.fx * a.xy.fy > 1;
    if (dec && out_dt != in_dt) return hipErrorInvalidValue;
    auto go = [&](auto tin, auto tout) -> hipError_t {
        using TIn = decltype(tin);
        using TOut = decltype(tout);
        const size_t lds = cxy_ref_lds_bytes(a.xy.kx, (int)sizeof(TIn));
        if (lds > 65536) return hipErrorInvalidValue;
        switch (a.xy.kx) {
#define NDL_K(KK)                                                                                              \
    case KK:                                                                                                   \
        if (dec)                                                                                               \
            hipLaunchKernelGGL((cxy_ref_kernel<TIn, TOut, KK, true>), dim3(a.xy.ntiles), dim3(256), lds, s, a);  \
        else                                                                                                   \
            hipLaunchKernelGGL((cxy_ref_kernel<TIn, TOut, KK, false>), dim3(a.xy.ntiles), dim3(256), lds, s, a); \
        return hipGetLastError();
            NDL_K(3) NDL_K(5) NDL_K(7) NDL_K(9)
#undef NDL_K
        }
        return hipErrorInvalidValue;
    };
    auto pick = [&](auto tin) -> hipError_t {
        if (out_dt == NDL_F64) return go(tin, double{});
        return go(tin, tin);
    };
    switch (in_dt) {
    case NDL_U8: return pick(uint8_t{});
    case NDL_U16: return pick(uint16_t{});
    case NDL_I16: return pick(int16_t{});
    case NDL_F32: return pick(float{});
    }
    return hipErrorInvalidValue;
}

hipError_t launch_xy_ref(const XYRefArgs& a, int in_dt, int out_dt, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    if (a.kx != a.ky) return hipErrorInvalidValue;
    return with_xy_instance(in_dt, out_dt, a.kx, [&](auto tin, auto tout, auto kk) {
        using TIn = decltype(tin);
        const size_t lds = xy_ref_lds_bytes(a.kx, a.ky, (int)sizeof(TIn));
        if (lds > 65536) return hipErrorInvalidValue;
        if (a.fx * a.fy > 1)
            hipLaunchKernelGGL((xy_ref_kernel<TIn, decltype(tout), decltype(kk)::value, true>), dim3(a.ntiles),
                               dim3(256), lds, s, a);
        else
            hipLaunchKernelGGL((xy_ref_kernel<TIn, decltype(tout), decltype(kk)::value, false>), dim3(a.ntiles),
                               dim3(256), lds, s, a);
        return hipGetLastError();
    });
}

} // namespace ndl
