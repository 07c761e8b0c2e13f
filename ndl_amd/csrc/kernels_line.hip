// kernels_line.hip -- one-axis passes over float32 / float64 storage (and
// bf16 / f16 storage with FP32 accumulation): a 1-D
// kernel image {k, 1, 1}, {1, k, 1}, ... given to convolve()
// (processing/convolution.h:52-71), or one per-axis pass of gaussian_blur()
// (convolution.h:107-186, through FP64 intermediates).  With OP = 1 / 2 the
// same kernels run the per-axis min / max passes of erode / dilate over a
// box (processing/morphology.h:57-101) on float32 and uint8 images.
//
// Arithmetic (this file is built with -ffp-contract=off):
//  * Acc = double (NDL_ACCUM_REF): total = 0, then total = total +
//    double(x) * tap for the taps in ascending order, every product and sum
//    rounded separately, one rounding into the output type -- the
//    reference's loop exactly.  Callers send no zero tap here (the reference
//    skips them, detail.h:164-174).
//  * Acc = float (NDL_ACCUM_F32): acc = fmaf(tap, x, acc), taps ascending.
//
// line_march_kernel (an axis other than the unit-stride one): every lane owns
// V consecutive elements of the unit-stride dimension x of one line (a fixed
// index of every other dimension) and marches along the convolved axis z.
// The input slices it has read sit in a ring of K + D registers per element:
// slice i+D is requested while output i-(K-1) is computed, so D loads per
// lane stay in flight.  z positions, their border fold and the taps are the
// same for the whole workgroup (scalar registers).  Lanes are numbered
// x-fastest over (x / V, line), so a workgroup reads 256 x V contiguous
// elements per slice where x rows are dense.  No LDS and no barrier.
//
// row_pass_kernel (the unit-stride axis): see below.
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <int V> using IC = std::integral_constant<int, V>;

// V consecutive elements in one access (V * sizeof(T) bytes, aligned to it)
template <class T, int V> struct alignas(sizeof(T) * V) Vec {
    T v[V];
};
template <class T, int V> __device__ __forceinline__ void load_v(const T* p, T* v)
{
    const Vec<T, V> q = *reinterpret_cast<const Vec<T, V>*>(p);
#pragma unroll
    for (int e = 0; e < V; e++) v[e] = q.v[e];
}
template <class T, int V> __device__ __forceinline__ void store_v(T* p, const T* v)
{
    Vec<T, V> q;
#pragma unroll
    for (int e = 0; e < V; e++) q.v[e] = v[e];
    *reinterpret_cast<Vec<T, V>*>(p) = q;
}

template <class Acc> __device__ __forceinline__ Acc mac1(Acc acc, Acc w, Acc x);
template <> __device__ __forceinline__ double mac1<double>(double acc, double w, double x)
{
    return acc + x * w; // two roundings, as the reference's `total += src * k`
}
template <> __device__ __forceinline__ float mac1<float>(float acc, float w, float x) { return fmaf(w, x, acc); }

// storage element <-> accumulator (bf16 / f16 through their bit patterns)
template <class Acc, class T> __device__ __forceinline__ Acc to_acc(const T& x)
{
    if constexpr (sizeof(Acc) == 8) return load_f64(&x);
    else return load_f32(&x);
}
template <class Acc, class T> __device__ __forceinline__ void from_acc(T& o, Acc v, int rounding)
{
    if constexpr (sizeof(Acc) == 8) store_from_f64(&o, v, rounding);
    else store_from_f32(&o, v);
}

// erode (OP 1) / dilate (OP 2): whether v replaces the current best; taps
// run in ascending order, strict comparisons (the first tap wins ties, so a
// -0 / +0 tie keeps the earlier one) and a NaN never wins over a number
// (processing/morphology.h:57-101 with the NaN rule of the separable passes)
template <int OP, class T> __device__ __forceinline__ bool morph_take(T best, T v)
{
    if constexpr (std::is_floating_point<T>::value)
        if (best != best) return v == v;
    return OP == 2 ? v > best : v < best;
}

template <class Acc> __device__ __forceinline__ Acc tap(const float* w, const double* wd, int t);
template <> __device__ __forceinline__ double tap<double>(const float*, const double* wd, int t) { return wd[t]; }
template <> __device__ __forceinline__ float tap<float>(const float* w, const double*, int t) { return w[t]; }

template <class TIn, class TOut, class Acc, int OP, int K, int D, int V>
__global__ __launch_bounds__(256) void line_march_kernel(const LineArgs a)
{
    constexpr int R = K + D;  // ring: K inputs of the current output plus D in flight
    // block -> (z chunk, lane block); chunks fastest so that the chunks of a
    // lane block run together and share their halo slices in cache
    const uint32_t chunk = blockIdx.x % (uint32_t)a.tiles_z;
    const uint32_t lb = blockIdx.x / (uint32_t)a.tiles_z;
    const uint32_t lane = lb * 256u + threadIdx.x;
    const bool live = lane < a.lanes;
    // surplus lanes of the last block read the last lane's elements and store nothing
    const uint32_t l = live ? lane : a.lanes - 1;
    uint32_t b = a.xv.div(l);
    const uint32_t xv = l - b * a.xv.d;
    int64_t so = (int64_t)xv * V, dof = (int64_t)xv * V, co = 0;
    for (int i = 0; i < a.nb; i++) {
        const uint32_t q = a.bext[i].div(b);
        const uint32_t r = b - q * a.bext[i].d;
        so += (int64_t)r * a.sb[i];
        dof += (int64_t)r * a.db[i];
        if constexpr (OP != 0) co += (int64_t)safe_map32(a.cbb[i], (int)a.bext[i].d, (int)r + a.cob[i]) * a.cbs[i];
        b = q;
    }
    // corner offsets of this lane's x elements (erode / dilate, last float pass)
    int64_t cxo[OP != 0 ? V : 1];
    if constexpr (OP != 0) {
#pragma unroll
        for (int e = 0; e < V; e++)
            cxo[e] = co + (int64_t)safe_map32(a.cbx, a.X, (int)xv * V + e + a.cox) * a.cxs;
    }
    const TIn* scol = (const TIn*)a.src + so;
    TOut* dcol = (TOut*)a.dst + dof;
    const int z0 = (int)chunk * a.tz;
    const int n_out = min(a.tz, a.Z - z0);
    const int n_in = n_out + K - 1;
    const int p0 = z0 - a.c;  // position of input 0

    TIn ring[R][V];
#pragma unroll
    for (int j = 0; j < R; j++)
#pragma unroll
        for (int e = 0; e < V; e++) ring[j][e] = TIn{};
    // Every step issues its load and runs its taps unconditionally (steps
    // past the chunk re-read input n_in - 1, warm-up and trailing steps
    // compute values nobody stores; the host sizes chunks for whole blocks
    // of R steps): with no load or product under a branch the compiler
    // waits for exactly the ring slot it needs, and the straight-line steps
    // overlap the next loads with the current taps.
    // Inputs i in [lo, hi) lie inside the axis and are read at the running
    // offset `off` = (p0 + i) * sz; the border map runs only outside it.
    const int lo = max(0, -p0), hi = min(n_in, a.Z - p0);
    auto load = [&](int i, int64_t off, TIn* v) {
        int64_t o = off;
        if (i < lo || i >= hi) o = (int64_t)safe_map32(a.border, a.Z, p0 + min(i, n_in - 1)) * a.sz;
        load_v<TIn, V>(scol + o, v);
    };
    int64_t off = (int64_t)p0 * a.sz;
#pragma unroll
    for (int j = 0; j < D; j++) {
        load(j, off, ring[j]);
        off += a.sz;
    }
    int64_t doff = (int64_t)(z0 - (K - 1)) * a.dz; // output of step i: i - (K - 1)
    for (int i0 = 0; i0 < n_in; i0 += R) {
#pragma unroll
        for (int u = 0; u < R; u++) {
            const int i = i0 + u;
            load(i + D, off, ring[(u + D) % R]);
            off += a.sz;
            TOut o[V];
            if constexpr (OP == 0) {
                Acc acc[V];
#pragma unroll
                for (int e = 0; e < V; e++) acc[e] = 0;
#pragma unroll
                for (int t = 0; t < K; t++)
#pragma unroll
                    for (int e = 0; e < V; e++)
                        acc[e] = mac1<Acc>(acc[e], tap<Acc>(a.w, a.wd, t),
                                           to_acc<Acc>(ring[(u - (K - 1) + t + R) % R][e]));
#pragma unroll
                for (int e = 0; e < V; e++) from_acc<Acc>(o[e], acc[e], a.rounding);
            } else {
#pragma unroll
                for (int e = 0; e < V; e++) {
                    TIn best = ring[(u - (K - 1) + R) % R][e];
#pragma unroll
                    for (int t = 1; t < K; t++) {
                        const TIn v = ring[(u - (K - 1) + t + R) % R][e];
                        if (morph_take<OP>(best, v)) best = v;
                    }
                    o[e] = best;
                }
                if constexpr (std::is_floating_point<TIn>::value) {
                    if (a.corner && i >= K - 1 && i < n_in) {
                        const int zc = safe_map32(a.cbz, a.Z, z0 + i - (K - 1) + a.coz);
#pragma unroll
                        for (int e = 0; e < V; e++) {
                            const TIn cv = ((const TIn*)a.corner)[cxo[e] + (int64_t)zc * a.czs];
                            if (cv != cv) o[e] = cv;
                        }
                    }
                }
            }
            if (live && i >= K - 1 && i < n_in) store_v<TOut, V>(dcol + doff, o);
            doff += a.dz;
        }
    }
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows: every wave takes segments of L = 64 V consecutive outputs of one row
// in turn (segment s, s + waves, ...).  A segment's inputs -- V elements per
// lane plus the c left and K-1-c right halo elements, border-folded near the
// row ends (single-fold maps, mathHelpers.h:23-25) -- go to the wave's own
// LDS window; each lane then reads its K+V-1 window values and writes V
// outputs.  The next segment's loads are issued before the current one is
// computed.  Only wave-level ordering is needed (no workgroup barrier).
template <class TIn, class TOut, class Acc, int OP, int K, int V>
__global__ __launch_bounds__(256) void row_pass_kernel(const RowArgs a)
{
    constexpr int L = 64 * V;
    constexpr int HLMAX = (K - 1 + 3) & ~3;            // >= the aligned left halo for any centre
    constexpr int W = (HLMAX + L + K + 3) & ~3;         // window elements per wave
    __shared__ __align__(32) TIn win[4][W];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    TIn* buf = win[wv];
    const int c = a.c;
    const int hl = (c + 3) & ~3;                        // main part starts on a 4-element boundary
    const uint32_t nwaves = gridDim.x * 4u;

    auto row_of = [&](uint32_t s, int64_t& so, int64_t& dof, int& x0) {
        uint32_t r = a.segs.div(s);
        x0 = (int)(s - r * a.segs.d) * L;
        so = 0;
        dof = 0;
        for (int i = 0; i < a.nb; i++) {
            const uint32_t q = a.bext[i].div(r);
            const uint32_t rr = r - q * a.bext[i].d;
            so += (int64_t)rr * a.sb[i];
            dof += (int64_t)rr * a.db[i];
            r = q;
        }
    };
    auto issue = [&](uint32_t s, TIn* m, TIn& h) {
        int64_t so, dof;
        int x0;
        row_of(s, so, dof, x0);
        const TIn* row = (const TIn*)a.src + so;
        const int xm = x0 + V * lane;
        if (xm + V - 1 < a.X) {
            load_v<TIn, V>(row + xm, m);
        } else {
#pragma unroll
            for (int e = 0; e < V; e++) m[e] = row[safe_map32(a.border, a.X, xm + e)];
        }
        if (lane < K - 1) {
            const int p = lane < c ? x0 - c + lane : x0 + L + (lane - c);
            h = row[safe_map32(a.border, a.X, p)];
        }
    };

    uint32_t seg = blockIdx.x * 4u + (uint32_t)wv;
    TIn m[V], h{};
    if (seg < a.total) issue(seg, m, h);
    for (; seg < a.total; seg += nwaves) {
        wave_lds_sync();  // the previous segment's window reads are done
        store_v<TIn, V>(buf + hl + V * lane, m);
        if (lane < K - 1) buf[lane < c ? hl - c + lane : hl + L + (lane - c)] = h;
        wave_lds_sync();
        if (seg + nwaves < a.total) issue(seg + nwaves, m, h);
        TIn wv_[K + V - 1];
#pragma unroll
        for (int j = 0; j < K + V - 1; j++) wv_[j] = buf[hl - c + V * lane + j];
        TOut o[V];
        if constexpr (OP == 0) {
            Acc acc[V];
#pragma unroll
            for (int e = 0; e < V; e++) acc[e] = 0;
#pragma unroll
            for (int t = 0; t < K; t++)
#pragma unroll
                for (int e = 0; e < V; e++)
                    acc[e] = mac1<Acc>(acc[e], tap<Acc>(a.w, a.wd, t), to_acc<Acc>(wv_[e + t]));
#pragma unroll
            for (int e = 0; e < V; e++) from_acc<Acc>(o[e], acc[e], a.rounding);
        } else {
#pragma unroll
            for (int e = 0; e < V; e++) {
                TIn best = wv_[e];
#pragma unroll
                for (int t = 1; t < K; t++)
                    if (morph_take<OP>(best, wv_[e + t])) best = wv_[e + t];
                o[e] = best;
            }
        }
        int64_t so, dof;
        int x0;
        row_of(seg, so, dof, x0);
        TOut* drow = (TOut*)a.dst + dof;
        const int xo = x0 + V * lane;
        if (xo + V - 1 < a.X) {
            store_v<TOut, V>(drow + xo, o);
        } else {
#pragma unroll
            for (int e = 0; e < V; e++)
                if (xo + e < a.X) drow[xo + e] = o[e];
        }
    }
}

// f(TIn, TOut, Acc, OP) for the storage / accumulation / operation
// combinations (erode and dilate: float32 and uint8, the same type in and out)
template <class F> hipError_t with_types(const LineKey& key, F&& f)
{
    if (key.op == 1 || key.op == 2) {
        if (key.in_dt != key.out_dt) return hipErrorInvalidValue;
        if (key.in_dt == NDL_F32) return key.op == 1 ? f(float{}, float{}, float{}, IC<1>{}) : f(float{}, float{}, float{}, IC<2>{});
        if (key.in_dt == NDL_U8)
            return key.op == 1 ? f(uint8_t{}, uint8_t{}, float{}, IC<1>{}) : f(uint8_t{}, uint8_t{}, float{}, IC<2>{});
        return hipErrorInvalidValue;
    }
    if (key.op != 0) return hipErrorInvalidValue;
    if (key.accum == NDL_ACCUM_F32) {
        if (key.in_dt == NDL_F32 && key.out_dt == NDL_F32) return f(float{}, float{}, float{}, IC<0>{});
        if (key.in_dt == NDL_BF16 && key.out_dt == NDL_BF16) return f(bf16_t{}, bf16_t{}, float{}, IC<0>{});
        if (key.in_dt == NDL_F16 && key.out_dt == NDL_F16) return f(f16_t{}, f16_t{}, float{}, IC<0>{});
        return hipErrorInvalidValue;
    }
    if (key.accum != NDL_ACCUM_REF) return hipErrorInvalidValue;
    if (key.in_dt == NDL_F32 && key.out_dt == NDL_F32) return f(float{}, float{}, double{}, IC<0>{});
    if (key.in_dt == NDL_F32 && key.out_dt == NDL_F64) return f(float{}, double{}, double{}, IC<0>{});
    if (key.in_dt == NDL_F64 && key.out_dt == NDL_F64) return f(double{}, double{}, double{}, IC<0>{});
    if (key.in_dt == NDL_F64 && key.out_dt == NDL_F32) return f(double{}, float{}, double{}, IC<0>{});
    // a single pass over a uint8 image (convolve() with a one-axis kernel)
    if (key.in_dt == NDL_U8 && key.out_dt == NDL_U8) return f(uint8_t{}, uint8_t{}, double{}, IC<0>{});
    // the last pass of a chain into an integral image (rounded per a.rounding)
    if (key.in_dt == NDL_F64 && key.out_dt == NDL_U8) return f(double{}, uint8_t{}, double{}, IC<0>{});
    if (key.in_dt == NDL_F64 && key.out_dt == NDL_U16) return f(double{}, uint16_t{}, double{}, IC<0>{});
    if (key.in_dt == NDL_F64 && key.out_dt == NDL_I16) return f(double{}, int16_t{}, double{}, IC<0>{});
    return hipErrorInvalidValue;
}

template <class F> hipError_t with_k(int k, F&& f)
{
    switch (k) {
#define NDL_K(KK) case KK: return f(IC<KK>{});
        NDL_K(3) NDL_K(5) NDL_K(7) NDL_K(9) NDL_K(11) NDL_K(13) NDL_K(15) NDL_K(17) NDL_K(19) NDL_K(21) NDL_K(23)
        NDL_K(25)
#undef NDL_K
    }
    return hipErrorInvalidValue;
}

// f(TIn, TOut, Acc, OP, K, D, V): line march instances.  Float input: V 4,
// 2, 1 with 8 loads in flight (and V 4 with 4 or 12 for FP32 accumulation;
// no V 4 with FP64 accumulators beyond 21 taps, whose ring and sums would
// spill); double input: V 2, 1 (16 bytes per lane at most); bf16 / f16 and
// uint8: V 4, 1.  Erode / dilate: box sizes up to 11.
template <class F> hipError_t with_line_instance(const LineKey& key, F&& f)
{
    return with_types(key, [&](auto ti, auto to, auto ac, auto op) {
        using TIn = decltype(ti);
        constexpr int OP = decltype(op)::value;
        return with_k(key.k, [&](auto kk) {
            if constexpr (OP != 0 && decltype(kk)::value > 11) {
                return hipErrorInvalidValue;
            } else if constexpr (sizeof(TIn) <= 2) {
                if (key.v == 4 && key.d == 8) return f(ti, to, ac, op, kk, IC<8>{}, IC<4>{});
                if (key.v == 1 && key.d == 8) return f(ti, to, ac, op, kk, IC<8>{}, IC<1>{});
                return hipErrorInvalidValue;
            } else {
                if constexpr (sizeof(TIn) == 4 && (sizeof(decltype(ac)) == 4 || decltype(kk)::value <= 21)) {
                    if (key.v == 4 && key.d == 8) return f(ti, to, ac, op, kk, IC<8>{}, IC<4>{});
                    if constexpr (std::is_same_v<decltype(ac), float> && OP == 0) {
                        if (key.v == 4 && key.d == 4) return f(ti, to, ac, op, kk, IC<4>{}, IC<4>{});
                        if (key.v == 4 && key.d == 12) return f(ti, to, ac, op, kk, IC<12>{}, IC<4>{});
                    }
                }
                if constexpr (OP == 0) {
                    if (key.v == 2 && key.d == 8) return f(ti, to, ac, op, kk, IC<8>{}, IC<2>{});
                }
                if (key.v == 1 && key.d == 8) return f(ti, to, ac, op, kk, IC<8>{}, IC<1>{});
                return hipErrorInvalidValue;
            }
        });
    });
}

// f(TIn, TOut, Acc, OP, K, V): row instances, V 4 or 1 (erode / dilate: box
// sizes up to 11)
template <class F> hipError_t with_row_instance(const LineKey& key, F&& f)
{
    return with_types(key, [&](auto ti, auto to, auto ac, auto op) {
        return with_k(key.k, [&](auto kk) {
            if constexpr (decltype(op)::value != 0 && decltype(kk)::value > 11) {
                return hipErrorInvalidValue;
            } else {
                if (key.v == 4) return f(ti, to, ac, op, kk, IC<4>{});
                if (key.v == 1) return f(ti, to, ac, op, kk, IC<1>{});
                return hipErrorInvalidValue;
            }
        });
    });
}

// ------------------------------------------------- blur then downsample
// A blurred value as gaussian_blur would store it in the output type (code
// dt), read back as double; and the block mean stored as downsample() does.
template <class Acc> __device__ __forceinline__ double stored_value(Acc v, int dt, int rounding)
{
    auto rt = [&](auto tag) {
        using T = decltype(tag);
        T b;
        if constexpr (sizeof(Acc) == 8) store_from_f64(&b, v, rounding);
        else store_from_f32(&b, v);
        return load_f64(&b);
    };
    switch (dt) {
    case NDL_U8: return rt(uint8_t{});
    case NDL_I8: return rt(int8_t{});
    case NDL_U16: return rt(uint16_t{});
    case NDL_I16: return rt(int16_t{});
    case NDL_I32: return rt(int32_t{});
    case NDL_F32: return rt(float{});
    case NDL_F16: return rt(f16_t{});
    case NDL_BF16: return rt(bf16_t{});
    default: return (double)v;
    }
}

__device__ __forceinline__ void store_mean(void* out, int64_t off, double v, int dt)
{
    switch (dt) {
    case NDL_U8: store_from_f64((uint8_t*)out + off, v, NDL_ROUND_NEAREST); break;
    case NDL_I8: store_from_f64((int8_t*)out + off, v, NDL_ROUND_NEAREST); break;
    case NDL_U16: store_from_f64((uint16_t*)out + off, v, NDL_ROUND_NEAREST); break;
    case NDL_I16: store_from_f64((int16_t*)out + off, v, NDL_ROUND_NEAREST); break;
    case NDL_I32: store_from_f64((int32_t*)out + off, v, NDL_ROUND_NEAREST); break;
    case NDL_F32: store_from_f64((float*)out + off, v, NDL_ROUND_NEAREST); break;
    case NDL_F16: store_from_f64((f16_t*)out + off, v, NDL_ROUND_NEAREST); break;
    case NDL_BF16: store_from_f64((bf16_t*)out + off, v, NDL_ROUND_NEAREST); break;
    default: store_from_f64((double*)out + off, v, NDL_ROUND_NEAREST); break;
    }
}

template <class Acc> __device__ __forceinline__ Acc bd_tap(const BlurDownArgs& a, int t);
template <> __device__ __forceinline__ double bd_tap<double>(const BlurDownArgs& a, int t) { return a.taps_d[t]; }
template <> __device__ __forceinline__ float bd_tap<float>(const BlurDownArgs& a, int t) { return a.taps_f[t]; }

// Lanes of a wave: position in the block p slowest, output column fastest
// (ncol columns, P positions: p * ncol + column), so each position reads
// ncol consecutive columns.  Every lane marches its source column along the
// axis with a ring of K + D registers (as line_march_kernel), rounds each
// pass value into the output type, and the p = 0 lane of each column adds
// the P values of a slice in position order (the other decimated dimensions,
// dimension 0 fastest), slice after slice: downsample()'s odometer with the
// axis slowest.  At a block's last slice it stores the mean.  Lanes past P,
// past the columns or on a position beyond the ragged edge read column 0
// and add nothing.
template <int K, class Acc>
__global__ __launch_bounds__(256) void blur_down_march_kernel(const BlurDownArgs a)
{
    constexpr int D = 8, R = K + D;
    const uint32_t chunk = blockIdx.x % (uint32_t)a.tiles_z;
    const uint32_t wave = blockIdx.x / (uint32_t)a.tiles_z * 4u + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p = lane / a.ncol;
    const int ci = lane - p * a.ncol;
    const uint32_t col = wave * (uint32_t)a.ncol + (uint32_t)ci;
    bool valid = p < a.P && col < a.lanes;
    int64_t so = 0, dof = 0;
    int pcount = 1; // positions of this column's block across the other dimensions
    {
        uint32_t rem = col < a.lanes ? col : 0;
        int pr = p;
        for (int i = 0; i < a.nl; i++) {
            const int d = a.ldim[i];
            const uint32_t q = a.lfd[i].div(rem);
            const int64_t oc = (int64_t)(rem - q * (uint32_t)a.oext[d]);
            rem = q;
            dof += oc * a.ost[d];
            if (d == a.channel_axis) {
                so += oc * a.ist[d];
                continue;
            }
            const int dig = pr % a.factor;
            pr /= a.factor;
            const int64_t sc = oc * a.factor + dig;
            pcount *= (int)min((int64_t)a.factor, a.ext[d] - oc * a.factor);
            if (sc >= a.ext[d]) valid = false;
            so += sc * a.ist[d];
        }
    }
    if (!valid) so = 0;
    const int Z = (int)a.ext[a.axis];
    const int f = a.axis == a.channel_axis ? 1 : a.factor; // slices per block
    const int ob0 = (int)chunk * a.tz;
    const int ob1 = min(ob0 + a.tz, (int)a.oext[a.axis]);
    const int zs0 = ob0 * f;
    const int n_src = min(ob1 * f, Z) - zs0;
    const int n_in = n_src + K - 1;
    const int p0 = zs0 - a.center;
    const Acc* scol = (const Acc*)a.in + so;
    const int64_t sz = a.ist[a.axis], dz = a.ost[a.axis];
    const int src_lane = ci;

    // as line_march_kernel: every step loads and runs its taps
    // unconditionally (past the chunk it re-reads input n_in - 1), so the
    // compiler waits for one ring slot at a time; inputs in [lo, hi) are
    // read at a running offset, the border map runs outside it
    Acc ring[R];
#pragma unroll
    for (int j = 0; j < R; j++) ring[j] = Acc(0);
    const int lo = max(0, -p0), hi = min(n_in, Z - p0);
    auto load = [&](int i, int64_t off) {
        int64_t o = off;
        if (i < lo || i >= hi) o = (int64_t)safe_map32(a.border, Z, p0 + min(i, n_in - 1)) * sz;
        return scol[o];
    };
    int64_t off = (int64_t)p0 * sz;
#pragma unroll
    for (int j = 0; j < D; j++) {
        ring[j] = load(j, off);
        off += sz;
    }
    double sum = 0.0;
    int nz = 0, ob = ob0;
    for (int i0 = 0; i0 < n_in; i0 += R) {
#pragma unroll
        for (int u = 0; u < R; u++) {
            const int i = i0 + u;
            ring[(u + D) % R] = load(i + D, off);
            off += sz;
            Acc acc = 0;
#pragma unroll
            for (int t = 0; t < K; t++) acc = mac1<Acc>(acc, bd_tap<Acc>(a, t), ring[(u - (K - 1) + t + R) % R]);
            if (i >= K - 1 && i < n_in) {
                const double v = stored_value<Acc>(acc, a.out_dt, a.rounding);
                const int ok = valid ? 1 : 0;
                for (int q = 0; q < a.P; q++) {
                    const double vq = __shfl(v, src_lane + q * a.ncol);
                    if (__shfl(ok, src_lane + q * a.ncol)) sum = sum + vq;
                }
                nz++;
                if (nz == f || zs0 + i - (K - 1) == Z - 1) {
                    if (p == 0 && valid) store_mean(a.out, dof + (int64_t)ob * dz, sum / (double)(pcount * nz), a.out_dt);
                    sum = 0.0;
                    nz = 0;
                    ob++;
                }
            }
        }
    }
}

template <class F> hipError_t with_blur_down_instance(int K, int accum, F&& f)
{
    return with_k(K, [&](auto kk) {
        if (accum == NDL_ACCUM_REF) return f(kk, double{});
        if (accum == NDL_ACCUM_F32) return f(kk, float{});
        return hipErrorInvalidValue;
    });
}

} // namespace

hipError_t launch_blur_down_march(const BlurDownArgs& a, int accum, hipStream_t s)
{
    if (a.P < 1 || a.P > 64 || a.ncol != 64 / a.P) return hipErrorInvalidValue;
    const uint64_t waves = ((uint64_t)a.lanes + a.ncol - 1) / a.ncol;
    const uint64_t blocks = (uint64_t)a.tiles_z * ((waves + 3) / 4);
    if (blocks == 0) return hipSuccess;
    if (blocks > 0xffffffffull) return hipErrorInvalidValue;
    return with_blur_down_instance(a.K, accum, [&](auto kk, auto ac) {
        hipLaunchKernelGGL((blur_down_march_kernel<decltype(kk)::value, decltype(ac)>), dim3((uint32_t)blocks),
                           dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

bool row_pass_supported(const LineKey& key)
{
    return with_row_instance(key, [](auto, auto, auto, auto, auto, auto) { return hipSuccess; }) == hipSuccess;
}

int row_pass_blocks_per_cu(const LineKey& key)
{
    int n = 0;
    hipError_t e = with_row_instance(key, [&](auto ti, auto to, auto ac, auto op, auto kk, auto vv) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n,
            reinterpret_cast<const void*>(&row_pass_kernel<decltype(ti), decltype(to), decltype(ac),
                                                           decltype(op)::value, decltype(kk)::value,
                                                           decltype(vv)::value>),
            256, 0);
    });
    return e == hipSuccess ? n : 0;
}

hipError_t launch_row_pass(const RowArgs& a, const LineKey& key, uint32_t blocks, hipStream_t s)
{
    if (a.total == 0 || blocks == 0) return hipSuccess;
    return with_row_instance(key, [&](auto ti, auto to, auto ac, auto op, auto kk, auto vv) {
        hipLaunchKernelGGL((row_pass_kernel<decltype(ti), decltype(to), decltype(ac), decltype(op)::value,
                                            decltype(kk)::value, decltype(vv)::value>),
                           dim3(blocks), dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

bool line_march_supported(const LineKey& key)
{
    return with_line_instance(key, [](auto, auto, auto, auto, auto, auto, auto) { return hipSuccess; }) ==
           hipSuccess;
}

int line_march_blocks_per_cu(const LineKey& key)
{
    int n = 0;
    hipError_t e = with_line_instance(key, [&](auto ti, auto to, auto ac, auto op, auto kk, auto dd, auto vv) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n,
            reinterpret_cast<const void*>(&line_march_kernel<decltype(ti), decltype(to), decltype(ac),
                                                             decltype(op)::value, decltype(kk)::value,
                                                             decltype(dd)::value, decltype(vv)::value>),
            256, 0);
    });
    return e == hipSuccess ? n : 0;
}

hipError_t launch_line_march(const LineArgs& a, const LineKey& key, hipStream_t s)
{
    const uint64_t blocks = (uint64_t)a.tiles_z * ((a.lanes + 255u) / 256u);
    if (blocks == 0) return hipSuccess;
    return with_line_instance(key, [&](auto ti, auto to, auto ac, auto op, auto kk, auto dd, auto vv) {
        hipLaunchKernelGGL((line_march_kernel<decltype(ti), decltype(to), decltype(ac), decltype(op)::value,
                                              decltype(kk)::value, decltype(dd)::value, decltype(vv)::value>),
                           dim3((uint32_t)blocks), dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

} // namespace ndl
