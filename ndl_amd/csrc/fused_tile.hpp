// fused_tile.hpp -- device code of the fused separable kernel, shared by the
// forward instances (kernels_fused.hip) and the adjoint instances with edge
// weight rows (kernels_fused_edge.hip).
// Up to three convolved axes in one pass over HBM.
//
// A workgroup of TX*TY/CPT threads owns a TX x TY tile of output columns
// (x = the unit-stride dimension, y = a second axis) and marches along a third
// axis z, one input slice at a time:
//   1. the slice's (TY+ky-1) x (TX+kx-1) input window, border indices
//      resolved once per workgroup (detail.h:178-194 semantics), is requested
//      two steps ahead into registers (two buffers) and published to LDS;
//   2. y-pass: LDS -> LDS, 4 rows per task;
//   3. x-pass: LDS -> registers, CPT consecutive x outputs per thread;
//   steps are software-pipelined (y-pass of slice s beside the x/z-pass of
//   slice s-1, double-buffered LDS) so each step needs one barrier;
//   4. z-pass: each thread keeps the last P = kz x-filtered slices of its
//      columns in a register ring; once kz slices are in, one output slice is
//      written.  The slice loop is unrolled by 2P so every ring slot and
//      prefetch buffer is a compile-time register choice.
// Every input element is read from HBM once per tile and every output written
// once; only the y/x halos are re-read (from L2) and re-computed.
//
// Tap counts are compile-time for the common filters (everything unrolled,
// taps in scalar registers, LDS offsets as immediates); a generic
// instantiation per ring period takes kx/ky at run time (4-tap chunks).
//
// Accumulation: FP32 FMA.  Integer (uint8) storage reaches this kernel only
// with integer taps and sums below 2^24, so every float value is an exact
// integer and the result is bit-identical to the reference's FP64 sum
// (convolution.h:64-70).  The host routes filters with an exactly-zero tap
// to the per-axis chain, which skips them (detail.h:164-174).
#pragma once
#include <type_traits>
#include <utility>

#include "ndl_internal.hpp"

namespace ndl {
namespace fused_detail {

template <class TS> __device__ __forceinline__ float load_s(const TS* p)
{
    if constexpr (is_int_storage<TS>::value) return (float)*p;
    else return load_f32(p);
}
template <class TS> __device__ __forceinline__ void store_s(TS* p, float v)
{
    if constexpr (is_int_storage<TS>::value) *p = (TS)(uint32_t)(int32_t)v; // exact integer, low bits (convolution.h:69)
    else store_from_f32(p, v);
}

template <int V> using IC = std::integral_constant<int, V>;

// lambdas of the tile body must inline: arrays they capture by reference
// (ring, prefetch buffers) would otherwise live in scratch memory
#define NDL_INLINE __attribute__((always_inline))

template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(IC<I>{}), ...);
}
// f(IC<0>{}), ..., f(IC<N-1>{}) in order
template <int N, class F> __device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// acc[o] += sum_{j<NTAP} w[j] * win[o + j], o < NO, taps ascending
// (taps IL window elements apart: interleaved channels, IL = channel count)
template <int NO, int NTAP, int NW, int IL = 1>
__device__ __forceinline__ void fma_block(float (&acc)[NO], const float (&win)[NW], const float* w)
{
    static_assert(NO + (NTAP - 1) * IL <= NW, "window too short");
#pragma unroll
    for (int j = 0; j < NTAP; j++) {
        const float wj = w[j];
#pragma unroll
        for (int o = 0; o < NO; o++) acc[o] = __builtin_fmaf(win[o + j * IL], wj, acc[o]);
    }
}

// 1-D correlation of NO consecutive outputs read through get(i) (window
// element i).  KS > 0: KS taps, fully unrolled.  KS == 0: K taps at run time
// in 4-tap chunks over a sliding window: chunk [t0, t0+4) needs elements
// [t0, t0 + NO + 3), of which the first NO - 1 are carried over from the
// previous chunk, so every element is read once and nothing past the window
// (NO + K - 1 elements) is touched.
template <int NO, int KS, int IL = 1, class Get>
__device__ __forceinline__ void correlate(float (&acc)[NO], int K, const float* w, Get get)
{
    static_assert(KS > 0 || IL == 1, "interleaved taps need a compile-time tap count");
    if constexpr (KS > 0) {
        constexpr int NW = NO + (KS - 1) * IL;
        float win[NW];
#pragma unroll
        for (int i = 0; i < NW; i++) win[i] = get(i);
        fma_block<NO, KS, NW, IL>(acc, win, w);
    } else {
        constexpr int NW = NO + 3;
        float win[NW];
#pragma unroll
        for (int i = 0; i < NO - 1; i++) win[i] = get(i);
        int t0 = 0;
        for (; t0 + 4 <= K; t0 += 4) {
#pragma unroll
            for (int i = NO - 1; i < NW; i++) win[i] = get(t0 + i);
            fma_block<NO, 4, NW>(acc, win, w + t0);
#pragma unroll
            for (int i = 0; i < NO - 1; i++) win[i] = win[i + 4];
        }
        const int rem = K - t0;
        if (rem > 0) {
#pragma unroll
            for (int i = NO - 1; i < NO + 2; i++)
                if (i < NO - 1 + rem) win[i] = get(t0 + i);
            if (rem == 3) fma_block<NO, 3, NW>(acc, win, w + t0);
            else if (rem == 2) fma_block<NO, 2, NW>(acc, win, w + t0);
            else fma_block<NO, 1, NW>(acc, win, w + t0);
        }
    }
}

// Row index of position i in an axis's edge weight table, or -1 for a row
// that takes the band (i in [eL, M - eR)).
__device__ __forceinline__ int edge_row(int i, int M, int eL, int eR)
{
    return i < eL ? i : (i >= M - eR ? eL + (i - (M - eR)) : -1);
}

// Ring step U of period P (= kz): this slice's x-filtered values go to slot U;
// once P slices are in, output z = z0 + s - (P-1) is
// sum_t w[t] * in[z + t - cz], t ascending = lag l = P-1-t descending.
// wedge (uniform, EDGE only): this output slice's own weight row, window
// order t, in place of the band.
template <int U, int P, int CPT, bool EDGE, class TS>
__device__ __forceinline__ void ring_step(float (&ring)[CPT][P], const float (&xv)[CPT], int s, const FusedArgs& a,
                                          TS* out, bool row_ok, int x_first, const float* wedge, bool full)
{
#pragma unroll
    for (int o = 0; o < CPT; o++) ring[o][U] = xv[o];
    if (s < P - 1) return;
    float acc[CPT];
#pragma unroll
    for (int o = 0; o < CPT; o++) acc[o] = 0.f;
    // two loops, not a select of weight pointers: a pointer into the
    // by-value argument block would make the compiler copy it to scratch
    if (EDGE && wedge) {
#pragma unroll
        for (int l = P - 1; l >= 0; l--) {
            const float w = wedge[P - 1 - l];
#pragma unroll
            for (int o = 0; o < CPT; o++) acc[o] = __builtin_fmaf(ring[o][(U - l + P) % P], w, acc[o]);
        }
    } else {
#pragma unroll
        for (int l = P - 1; l >= 0; l--) {
            const float w = a.wzlag[l];
#pragma unroll
            for (int o = 0; o < CPT; o++) acc[o] = __builtin_fmaf(ring[o][(U - l + P) % P], w, acc[o]);
        }
    }
    if (full) { // uniform: the whole tile is inside x and y
#pragma unroll
        for (int o = 0; o < CPT; o++) store_s<TS>(out + o, acc[o]);
    } else if (row_ok) {
#pragma unroll
        for (int o = 0; o < CPT; o++)
            if (x_first + o < a.X) store_s<TS>(out + o, acc[o]);
    }
}

// Tap-gradient step (GRAD): the ring keeps U = x through the x/y taps, and
// once P slices are in, output z = z0 + s - (P-1) adds g[z] * U[B(z + t - c)]
// for every tap t = P-1-l (lag l) to FP64 accumulators.  Positions outside
// the image read g as 0.
template <int U, int P, int CPT, class TS>
__device__ __forceinline__ void grad_step(float (&ring)[CPT][P], const float (&xv)[CPT], int s, const FusedArgs& a,
                                          const TS* g, bool row_ok, int x_first, bool full, double (&gacc)[P])
{
#pragma unroll
    for (int o = 0; o < CPT; o++) ring[o][U] = xv[o];
    if (s < P - 1) return;
    float gv[CPT];
#pragma unroll
    for (int o = 0; o < CPT; o++) gv[o] = (full || (row_ok && x_first + o < a.X)) ? load_s<TS>(g + o) : 0.f;
#pragma unroll
    for (int l = P - 1; l >= 0; l--) {
#pragma unroll
        for (int o = 0; o < CPT; o++)
            gacc[P - 1 - l] = __builtin_fma((double)gv[o], (double)ring[o][(U - l + P) % P], gacc[P - 1 - l]);
    }
}

// One tile: `tile` indexes (x tile, y tile, z chunk, batch) with x fastest.
// KXS/KYS: compile-time x/y tap counts, 0 = taken from the arguments.
// LDS (floats): sIn = (NR+3) x NC input window (+3 rows of slack for the
// chunked y reads), sY = TY x NCP y-filtered rows (NCP >= NC+3, a multiple
// of 4, slack for the chunked x reads).
// EDGE: rows within an axis's edge band (a.eL / a.eR rows at each end, one
// weight row each in a.etab, window order) take their own weights in place of
// the band: the transposed clamp / asymmetric-reflect operator.
template <class TS, int TX, int TY, int CPT, int KXS, int KYS, int P, int IL = 1, bool EDGE = false, bool GRAD = false>
__device__ __forceinline__ void fused_tile(const FusedArgs& a, uint32_t tile, unsigned char* smem_raw)
{
    static_assert(!GRAD || (!EDGE && TY > 1), "tap gradients use the plain 2-D tile");
    static_assert(!EDGE || IL == 1, "edge rows are for the unit-interleave tile");
    constexpr int NT = TX * TY / CPT;
    constexpr int XG = TX / CPT;
    static_assert(CPT >= 1 && XG * CPT == TX && XG * TY == NT, "tile must map CPT x-outputs per thread");
    constexpr int RY = 4;
    // TY == 1: row march -- one tile row, no y-pass; the x-pass reads the
    // staged input row and the marched axis (z) is an image's y axis
    constexpr bool NOY = TY == 1;
    static_assert(NOY ? (KYS == 1 && !EDGE) : TY % RY == 0, "TY must be 1 or a multiple of 4");
    constexpr int KMAX = 25;
    constexpr int NCMAX = TX + ((KXS > 0 ? KXS : KMAX) - 1) * IL;
    constexpr int NRMAX = TY + (KYS > 0 ? KYS : KMAX) - 1;
    constexpr int NLMAX = (NCMAX * NRMAX + NT - 1) / NT;
    constexpr int NTASKMAX = NOY ? 1 : (NCMAX * (TY / RY) + NT - 1) / NT;

    const int kx = KXS > 0 ? KXS : a.kx;
    const int ky = KYS > 0 ? KYS : a.ky;
    const int NC = TX + (kx - 1) * IL; // window columns (x taps IL apart)
    const int NR = TY + ky - 1;
    const int NCP = (NC + 3 + 3) & ~3;
    const int NIN = NR * NC;
    // two input windows and two y-filtered planes (pipelined steps)
    // PAD (compile-time taps): every load slot and every lane of the y-pass
    // waves works unconditionally -- surplus load lanes fetch element 0 into
    // LDS past the window, surplus y lanes recompute a task of the previous
    // wave (the same value to the same address) -- so the per-step work has
    // no lane-divergent branches
    constexpr bool PAD = KXS > 0 && KYS > 0;
    const int IN_FL = PAD ? max(((NR + 3) * NC + 3) & ~3, NLMAX * NT) : ((NR + 3) * NC + 3) & ~3;
    const int Y_FL = TY * NCP;
    float* sIn0 = (float*)smem_raw;
    float* sIn1 = sIn0 + IN_FL;
    float* sY0 = sIn1 + IN_FL;
    float* sY1 = sY0 + Y_FL;
    // EDGE: the weight rows, staged once per tile; regions of 2*kx*kx,
    // 2*ky*ky and 2*P*P floats for x, y, z (at most K rows at each end)
    float* sE = sY1 + Y_FL;
    const int eoff_y = 2 * kx * kx, eoff_z = eoff_y + 2 * ky * ky;
    // run-time taps live in LDS (x at [0, 32), y at [32, 64)): indexing the
    // by-value argument block at run-time offsets makes the compiler copy it
    // to scratch.  Compile-time taps are read from the arguments directly.
    float* sW = sE + (EDGE ? eoff_z + 2 * P * P : 0);
    const float* wxp;
    const float* wyp;
    if constexpr (KXS > 0) wxp = a.wx;
    else wxp = sW;
    if constexpr (KYS > 0) wyp = a.wy;
    else wyp = sW + 32;

    uint32_t q = a.fdx.div(tile);
    const int tix = (int)(tile - q * (uint32_t)a.tiles_x);
    uint32_t q2 = a.fdy.div(q);
    const int tiy = (int)(q - q2 * (uint32_t)a.tiles_y);
    uint32_t bi = a.fdz.div(q2);
    const int tiz = (int)(q2 - bi * (uint32_t)a.tiles_z);
    int64_t soff = 0, doff = 0;
    for (int d = 0; d < a.nb; d++) {
        uint32_t qq = a.bdiv[d].div(bi);
        int64_t c = (int64_t)(bi - qq * (uint32_t)a.bext[d]);
        bi = qq;
        soff += c * a.sb[d];
        doff += c * a.db[d];
    }
    const int x0 = tix * TX, y0 = tiy * TY, z0 = tiz * a.tz;
    const int nz = min(a.tz, a.Z - z0);
    const int nsl = nz + P - 1;
    const TS* src = (const TS*)a.src + soff;
    TS* dst = (TS*)a.dst + doff;
    const int tid = threadIdx.x;

    // per-thread load slots: element i = tid + k*NT of the window, border
    // indices resolved once for the whole march.  Offsets are taken from the
    // smallest in-slice offset (row Y-1 when the row stride is negative) and
    // kept in bytes, so they are non-negative 32-bit values added to a scalar
    // base (the host keeps every slice span below 4 GiB).
    const int64_t lmin = a.sy < 0 ? (int64_t)(a.Y - 1) * a.sy : 0;
    uint32_t loff[NLMAX];
#pragma unroll
    for (int k = 0; k < NLMAX; k++) {
        const int i = tid + k * NT;
        loff[k] = 0;
        if (i < NIN) {
            const int r = i / NC;
            const int c = i - r * NC;
            const int yy = safe_map32(a.by, a.Y, y0 - a.cy + r);
            int xx;
            if constexpr (IL == 1) {
                xx = safe_map32(a.bx, a.X, x0 - a.cx + c);
            } else {
                // merged (channel, x) index: map x, keep the channel
                const int e = x0 - a.cx * IL + c;
                const int w = e >= 0 ? e / IL : -((-e + IL - 1) / IL);
                xx = safe_map32(a.bx, a.X / IL, w) * IL + (e - w * IL);
            }
            loff[k] = (uint32_t)(((int64_t)yy * a.sy + xx - lmin) * (int64_t)sizeof(TS));
        }
    }
    // y-pass tasks: column c, rows [r0, r0+RY); with one task per thread the
    // tasks fill the first ceil(ntask / 64) waves
    const int ywave = tid >> 6;
    int tcol[NTASKMAX], trow[NTASKMAX];
#pragma unroll
    for (int k = 0; k < NTASKMAX; k++) {
        int task = tid + k * NT;
        if constexpr (PAD && NTASKMAX == 1) {
            task = ywave * 64 + (tid & 63);
            if (task >= NC * (TY / RY) && NC * (TY / RY) >= 64) task -= 64;
        }
        const int rg = task / NC;
        tcol[k] = task - rg * NC;
        trow[k] = rg * RY;
    }
    const int ntask = NC * (TY / RY);
    // x-pass / ring ownership: CPT consecutive x of row ry
    const int ry = tid / XG;
    const int xs = (tid - ry * XG) * CPT;
    const bool row_ok = y0 + ry < a.Y;
    const bool full = x0 + TX <= a.X && y0 + TY <= a.Y;
    // tiles that reach into an edge band (uniform)
    const bool ex = EDGE && a.etab[0] && (x0 < a.eL[0] || x0 + TX > a.X - a.eR[0]);
    const bool ey = EDGE && a.etab[1] && (y0 < a.eL[1] || y0 + TY > a.Y - a.eR[1]);

    // the run-time-tap path reads up to 3 rows / columns past a window in
    // 4-tap chunks: that slack must hold finite values.  Compile-time taps
    // read exactly the window.
    if constexpr (KXS == 0 || KYS == 0) {
        if (tid == 0) {
#pragma unroll
            for (int j = 0; j < 32; j++) {
                sW[j] = a.wx[j];
                sW[32 + j] = a.wy[j];
            }
        }
        for (int i = tid; i < 3 * NC; i += NT) {
            sIn0[NIN + i] = 0.f;
            sIn1[NIN + i] = 0.f;
        }
        for (int i = tid; i < TY * (NCP - NC); i += NT) {
            const int r = i / (NCP - NC), c = NC + i - r * (NCP - NC);
            sY0[r * NCP + c] = 0.f;
            sY1[r * NCP + c] = 0.f;
        }
    }

    // two register buffers: slice s+1 is in flight while slice s+2 is
    // requested, so a load has two slice steps to arrive
    float pre[NLMAX], pre2[NLMAX];
    // the z border map only matters for chunks that reach past an end
    const bool z_interior = z0 - a.cz >= 0 && z0 - a.cz + nsl <= a.Z;
    const TS* src_l = src + lmin;                          // offsets are relative to this
    const TS* src_z = src_l + (int64_t)(z0 - a.cz) * a.sz; // slice 0 of an interior chunk
    auto load_slice = [&](int s, float (&dstv)[NLMAX]) NDL_INLINE {
        const TS* base;
        if (z_interior) base = src_z + (int64_t)s * a.sz;
        else base = src_l + (int64_t)safe_map32(a.bz, a.Z, z0 - a.cz + s) * a.sz;
#pragma unroll
        for (int k = 0; k < NLMAX; k++)
            if (PAD || tid + k * NT < NIN)
                dstv[k] = load_s<TS>((const TS*)((const unsigned char*)base + loff[k]));
    };

    float ring[CPT][P];
#pragma unroll
    for (int o = 0; o < CPT; o++)
#pragma unroll
        for (int p = 0; p < P; p++) ring[o][p] = 0.f;
    double gacc[GRAD ? P : 1];
#pragma unroll
    for (int t = 0; t < (GRAD ? P : 1); t++) gacc[t] = 0.0;

    // y-pass of one slice: window `in` -> y-filtered plane `yp`
    // one y-pass task: column c of the window, output rows [r0, r0 + RY)
    auto y_task = [&](const float* in, float* yp, int r0, int c) NDL_INLINE {
        const float* col = in + r0 * NC + c;
        float o[RY];
#pragma unroll
        for (int r = 0; r < RY; r++) o[r] = 0.f;
        correlate<RY, KYS>(o, ky, wyp, [&](int i) NDL_INLINE { return col[i * NC]; });
        if constexpr (EDGE) {
            if (ey) {
#pragma unroll
                for (int r = 0; r < RY; r++) {
                    const int yy = y0 + r0 + r;
                    const int ri = yy < a.Y ? edge_row(yy, a.Y, a.eL[1], a.eR[1]) : -1;
                    if (ri >= 0) {
                        const float* w = sE + eoff_y + ri * ky;
                        float acc = 0.f;
                        for (int t = 0; t < ky; t++) acc = __builtin_fmaf(col[(r + t) * NC], w[t], acc);
                        o[r] = acc;
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RY; r++) yp[(r0 + r) * NCP + c] = o[r];
    };
    auto y_pass = [&](const float* in, float* yp) NDL_INLINE {
#pragma unroll
        for (int k = 0; k < NTASKMAX; k++) {
            if constexpr (PAD && NTASKMAX == 1) {
                // whole waves: the first ceil(ntask / 64) take a task each
                if (__builtin_amdgcn_readfirstlane(ywave) * 64 < ntask) y_task(in, yp, trow[k], tcol[k]);
            } else if (tid + k * NT < ntask) {
                y_task(in, yp, trow[k], tcol[k]);
            }
        }
    };
    // x-pass of one y-filtered plane, then ring slot U (slice s)
    auto xz_pass = [&](auto slot, int s, const float* yp) NDL_INLINE {
        constexpr int U = decltype(slot)::value;
        float xv[CPT];
#pragma unroll
        for (int o = 0; o < CPT; o++) xv[o] = 0.f;
        const float* row = NOY ? yp + xs : yp + ry * NCP + xs;
        correlate<CPT, KXS, IL>(xv, kx, wxp, [&](int i) NDL_INLINE { return row[i]; });
        const int zout = z0 + s - (P - 1);
        const float* wedge = nullptr;
        if constexpr (EDGE) {
            if (ex) {
#pragma unroll
                for (int o = 0; o < CPT; o++) {
                    const int xx = x0 + xs + o;
                    const int ri = xx < a.X ? edge_row(xx, a.X, a.eL[0], a.eR[0]) : -1;
                    if (ri >= 0) {
                        const float* w = sE + ri * kx;
                        float acc = 0.f;
                        for (int t = 0; t < kx; t++) acc = __builtin_fmaf(row[o + t], w[t], acc);
                        xv[o] = acc;
                    }
                }
            }
            if (a.etab[2] && s >= P - 1) {
                const int ri = edge_row(zout, a.Z, a.eL[2], a.eR[2]);
                if (ri >= 0) wedge = sE + eoff_z + ri * P;
            }
        }
        TS* out = dst + (int64_t)zout * a.dz + (int64_t)(y0 + ry) * a.dy + (x0 + xs);
        if constexpr (GRAD) grad_step<U, P, CPT, TS>(ring, xv, s, a, out, row_ok, x0 + xs, full, gacc);
        else ring_step<U, P, CPT, EDGE, TS>(ring, xv, s, a, out, row_ok, x0 + xs, wedge, full);
    };
    auto publish = [&](const float (&v)[NLMAX], float* in) NDL_INLINE {
#pragma unroll
        for (int k = 0; k < NLMAX; k++)
            if (PAD || tid + k * NT < NIN) in[tid + k * NT] = v[k];
    };

    // Pipelined schedule, one barrier per step.  Step s: publish slice s+1
    // (its registers then fetch slice s+3), y-pass slice s, x/z-pass slice
    // s-1.  Slice j lives in registers pre[j&1], LDS window sIn[j&1] and
    // y-plane sY[j&1]; the barrier ends every read of the buffers the next
    // step overwrites.
    if constexpr (EDGE) {
        auto stage = [&](const float* tab, int n, float* dstp) NDL_INLINE {
            for (int i = tid; i < n; i += NT) dstp[i] = tab[i];
        };
        if (a.etab[0]) stage(a.etab[0], (a.eL[0] + a.eR[0]) * kx, sE);
        if (a.etab[1]) stage(a.etab[1], (a.eL[1] + a.eR[1]) * ky, sE + eoff_y);
        if (a.etab[2]) stage(a.etab[2], (a.eL[2] + a.eR[2]) * P, sE + eoff_z);
    }
    load_slice(0, pre);
    if (nsl > 1) load_slice(1, pre2);
    __syncthreads(); // slack fill done
    publish(pre, sIn0);
    if (nsl > 2) load_slice(2, pre);
    __syncthreads();
    auto step = [&](auto ic, int s) NDL_INLINE {
        constexpr int u = decltype(ic)::value; // s = s0 + u, s0 a multiple of 2P
        constexpr bool even = (u % 2) == 0;
        if (s + 1 < nsl) {
            if constexpr (even) {
                publish(pre2, sIn1);
                if (s + 3 < nsl) load_slice(s + 3, pre2);
            } else {
                publish(pre, sIn0);
                if (s + 3 < nsl) load_slice(s + 3, pre);
            }
        }
        if constexpr (NOY) {
            // the x-pass of slice s reads its window straight from sIn
            if (s < nsl) xz_pass(IC<u % P>{}, s, even ? sIn0 : sIn1);
        } else {
            if (s < nsl) y_pass(even ? sIn0 : sIn1, even ? sY0 : sY1);
            if (s >= 1) xz_pass(IC<(u + P - 1) % P>{}, s - 1, even ? sY1 : sY0);
        }
        __syncthreads();
    };
    // 2P steps per iteration keep ring slots and buffers compile-time
    for (int s0 = 0; s0 <= nsl; s0 += 2 * P) {
        static_for<2 * P>([&](auto ic) NDL_INLINE {
            const int s = s0 + decltype(ic)::value;
            if (s <= nsl) step(ic, s);
        });
    }
    if constexpr (GRAD) {
        // the tile's partial sums: waves by shuffles, then the NT/64 wave
        // values in a fixed order (the march ended with a barrier, so sIn0
        // is free)
        double* red = (double*)sIn0;
        const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
        for (int t = 0; t < P; t++) {
            double v = gacc[t];
            for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
            if (lane == 0) red[wv * P + t] = v;
        }
        __syncthreads();
        if (tid < P) {
            double sum = 0.0;
            for (int w = 0; w < NT / 64; w++) sum = sum + red[w * P + tid];
            a.gpart[(size_t)tile * P + tid] = sum;
        }
    }
}

// Run-time-tap instances hold up to 153 VGPRs (3 waves per SIMD, one
// 512-thread workgroup per CU); asking for 4 waves per SIMD spills a few
// registers (8 at ring period 19, 31 at 25) but lets two workgroups share a
// CU, which measured faster on the 256-instance batched workload (plan 4.12
// -> 3.48 ms).  Compile-time-tap instances already fit 5-6 waves.
#ifndef NDL_QUEUE_WAVES_PER_EU
#define NDL_QUEUE_WAVES_PER_EU 4
#endif
template <class TS, int TX, int TY, int CPT, int KXS, int KYS, int P, int IL, bool EDGE = false, bool GRAD = false>
__global__ __launch_bounds__(TX * TY / CPT)
__attribute__((amdgpu_waves_per_eu(KXS == 0 || KYS == 0 ? NDL_QUEUE_WAVES_PER_EU : 1)))
void fused_tile_kernel(const FusedArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // each XCD gets a contiguous range of tiles when remapping (workgroups
    // are dealt round-robin over the 8 XCDs)
    uint32_t tile = blockIdx.x;
    if (a.xcd_remap) {
        const uint32_t per = a.ntiles >> 3;
        tile = (tile & 7u) * per + (tile >> 3);
    }
    fused_tile<TS, TX, TY, CPT, KXS, KYS, P, IL, EDGE, GRAD>(a, tile, smem_raw);
}

// Persistent work queue over many images with the same z tap count:
// workgroups take tiles from one counter until all `total` tiles of all jobs
// are done.  tile_start[j] is the first global tile of job j (prefix sums,
// tile_start[njobs] = total).
template <class TS, int TX, int TY, int CPT, int P>
__global__ __launch_bounds__(TX * TY / CPT) __attribute__((amdgpu_waves_per_eu(NDL_QUEUE_WAVES_PER_EU)))
void fused_queue_kernel(const FusedArgs* __restrict__ jobs,
                                                                    const uint32_t* __restrict__ tile_start,
                                                                    int njobs, uint32_t total, uint32_t* counter)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ uint32_t next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (t >= total) break; // every workgroup reaches this once the counter passes total
        int lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (tile_start[mid] <= t) lo = mid;
            else hi = mid;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        fused_tile<TS, TX, TY, CPT, 0, TY == 1 ? 1 : 0, P>(jobs[lo], t - tile_start[lo], smem_raw);
        __syncthreads(); // LDS is reused by the next tile
    }
}


} // namespace fused_detail
} // namespace ndl
