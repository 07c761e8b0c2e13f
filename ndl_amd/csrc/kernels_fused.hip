// kernels_fused.hip -- up to three convolved axes in one pass over HBM.
//
// A workgroup of 256 threads owns a TX x TY tile of output columns
// (x = the unit-stride dimension, y = a second axis) and marches along a third
// axis z, one input slice at a time:
//   1. the slice's (TY+ky-1) x (TX+kx-1) input window, border indices
//      resolved once per workgroup (detail.h:178-194 semantics), is prefetched
//      into registers one slice ahead and published to LDS;
//   2. y-pass: LDS -> LDS, register-blocked RY rows per task;
//   3. x-pass: LDS -> registers, each thread CPT consecutive x outputs;
//   4. z-pass: each thread keeps the last P x-filtered slices of its columns
//      in a register ring; once kz slices are in, one output slice is written.
// Every input element is read from HBM once per tile and every output written
// once; only the y/x halos are re-read (from L2) and re-computed.
//
// Accumulation: FP32 FMA for float storage (f32/f16/bf16), exact int32 for
// integer storage with integer taps (bit-identical to the reference's FP64
// sum, which is exact for such data, convolution.h:64-70).  Exactly-zero taps
// are skipped like detail::kernelIncludedTaps (detail.h:164-174).
#include <type_traits>

#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <class TS> __device__ __forceinline__ float load_s(const TS* p)
{
    if constexpr (is_int_storage<TS>::value) return (float)*p;
    else return load_f32(p);
}
// Integer storage only reaches this kernel with integer taps and sums below
// 2^24, so the float value is an exact integer (convolution.h:69 low bits).
template <class TS> __device__ __forceinline__ void store_s(TS* p, float v)
{
    if constexpr (is_int_storage<TS>::value) *p = (TS)(uint32_t)(int32_t)v;
    else store_from_f32(p, v);
}

// acc[o] += sum_{j<NTAP} w[j] * win[o + j], o < NO, taps ascending
template <int NO, int NTAP, int NW>
__device__ __forceinline__ void fma_block(float (&acc)[NO], const float (&win)[NW], const float* w)
{
    static_assert(NO + NTAP - 1 <= NW, "window too short");
#pragma unroll
    for (int j = 0; j < NTAP; j++) {
        const float wj = w[j];
#pragma unroll
        for (int o = 0; o < NO; o++) acc[o] = __builtin_fmaf(win[o + j], wj, acc[o]);
    }
}

// 1-D correlation of NO consecutive outputs read through `get(i)` (window
// element i), K taps at `w`: 4-tap chunks, then a 1..3-tap tail.
template <int NO, class Get>
__device__ __forceinline__ void correlate(float (&acc)[NO], int K, const float* w, Get get)
{
    constexpr int NW = NO + 3;
    int t0 = 0;
    for (; t0 + 4 <= K; t0 += 4) {
        float win[NW];
#pragma unroll
        for (int i = 0; i < NW; i++) win[i] = get(t0 + i);
        fma_block<NO, 4, NW>(acc, win, w + t0);
    }
    const int rem = K - t0;
    if (rem > 0) {
        float win[NW];
#pragma unroll
        for (int i = 0; i < NW; i++) win[i] = get(t0 + i);
        if (rem == 3) fma_block<NO, 3, NW>(acc, win, w + t0);
        else if (rem == 2) fma_block<NO, 2, NW>(acc, win, w + t0);
        else fma_block<NO, 1, NW>(acc, win, w + t0);
    }
}

// Ring step U of period P (= kz): this slice's x-filtered values go to slot U;
// once P slices are in, output z = z0 + s - (P-1) is
// sum_t w[t] * in[z + t - cz], t ascending = lag l = P-1-t descending.
template <int U, int P, int CPT, class TS>
__device__ __forceinline__ void ring_step(float (&ring)[CPT][P], const float (&xv)[CPT], int s, const FusedArgs& a,
                                          TS* out, bool row_ok, int x_first)
{
#pragma unroll
    for (int o = 0; o < CPT; o++) ring[o][U] = xv[o];
    if (s < P - 1) return;
    float acc[CPT];
#pragma unroll
    for (int o = 0; o < CPT; o++) acc[o] = 0.f;
#pragma unroll
    for (int l = P - 1; l >= 0; l--) {
        const float w = a.wzlag[l];
#pragma unroll
        for (int o = 0; o < CPT; o++) acc[o] = __builtin_fmaf(ring[o][(U - l + P) % P], w, acc[o]);
    }
    if (row_ok) {
#pragma unroll
        for (int o = 0; o < CPT; o++)
            if (x_first + o < a.X) store_s<TS>(out + o, acc[o]);
    }
}

template <int U, int P, int CPT, class TS>
__device__ __forceinline__ void ring_dispatch(int u, float (&ring)[CPT][P], const float (&xv)[CPT], int s,
                                              const FusedArgs& a, TS* out, bool row_ok, int x_first)
{
    if constexpr (U < P) {
        if (u == U) {
            ring_step<U, P, CPT, TS>(ring, xv, s, a, out, row_ok, x_first);
            return;
        }
        ring_dispatch<U + 1, P, CPT, TS>(u, ring, xv, s, a, out, row_ok, x_first);
    }
}

// One tile: `tile` indexes (x tile, y tile, z chunk, batch) with x fastest.
// LDS (floats): sIn = (NR+3) x NC input window (+3 rows of slack for the
// 4-tap chunked y reads), sY = TY x NCP y-filtered rows (NCP >= NC+3, a
// multiple of 4, slack for the chunked x reads).
template <class TS, int TX, int TY, int CPT, int P>
__device__ __forceinline__ void fused_tile(const FusedArgs& a, uint32_t tile, unsigned char* smem_raw)
{
    constexpr int kNT = TX * TY / CPT; // threads per workgroup
    constexpr int XG = TX / CPT;
    static_assert(CPT >= 1 && XG * CPT == TX && XG * TY == kNT, "tile must map CPT x-outputs per thread");
    constexpr int RY = 4;
    static_assert(TY % RY == 0, "TY must be a multiple of 4");
    constexpr int KMAX = 25;
    constexpr int NCMAX = TX + KMAX - 1;
    constexpr int NRMAX = TY + KMAX - 1;
    constexpr int NLMAX = (NCMAX * NRMAX + kNT - 1) / kNT;
    constexpr int NTASKMAX = (NCMAX * (TY / RY) + kNT - 1) / kNT;

    const int NC = TX + a.kx - 1;
    const int NR = TY + a.ky - 1;
    const int NCP = (NC + 3 + 3) & ~3;
    const int NIN = NR * NC;
    float* sIn = (float*)smem_raw;
    float* sY = sIn + (((NR + 3) * NC + 3) & ~3);

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
    // indices resolved once for the whole march
    int32_t loff[NLMAX];
#pragma unroll
    for (int k = 0; k < NLMAX; k++) {
        const int i = tid + k * kNT;
        loff[k] = 0;
        if (i < NIN) {
            const int r = i / NC;
            const int c = i - r * NC;
            const int yy = safe_map32(a.by, a.Y, y0 - a.cy + r);
            const int xx = safe_map32(a.bx, a.X, x0 - a.cx + c);
            loff[k] = (int32_t)((int64_t)yy * a.sy + xx);
        }
    }
    // y-pass tasks: column c, rows [r0, r0+RY)
    int tcol[NTASKMAX], trow[NTASKMAX];
#pragma unroll
    for (int k = 0; k < NTASKMAX; k++) {
        const int task = tid + k * kNT;
        const int rg = task / NC;
        tcol[k] = task - rg * NC;
        trow[k] = rg * RY;
    }
    const int ntask = NC * (TY / RY);
    // x-pass / ring ownership: CPT consecutive x of row ry
    const int ry = tid / XG;
    const int xs = (tid - ry * XG) * CPT;
    const bool row_ok = y0 + ry < a.Y;

    // slack reads must see finite values: zero the whole tile once
    const int lds_floats = (((NR + 3) * NC + 3) & ~3) + TY * NCP;
    for (int i = tid; i < lds_floats; i += kNT) sIn[i] = 0.f;

    float pre[NLMAX];
    auto load_slice = [&](int s) {
        const int zz = safe_map32(a.bz, a.Z, z0 - a.cz + s);
        const TS* base = src + (int64_t)zz * a.sz;
#pragma unroll
        for (int k = 0; k < NLMAX; k++)
            if (tid + k * kNT < NIN) pre[k] = load_s<TS>(base + loff[k]);
    };

    float ring[CPT][P];
#pragma unroll
    for (int o = 0; o < CPT; o++)
#pragma unroll
        for (int p = 0; p < P; p++) ring[o][p] = 0.f;

    load_slice(0);
    __syncthreads(); // zero fill done before the first publish
    for (int s = 0; s < nsl; s++) {
#pragma unroll
        for (int k = 0; k < NLMAX; k++)
            if (tid + k * kNT < NIN) sIn[tid + k * kNT] = pre[k];
        __syncthreads();
        if (s + 1 < nsl) load_slice(s + 1);

        // y-pass: sIn -> sY
#pragma unroll
        for (int k = 0; k < NTASKMAX; k++) {
            if (tid + k * kNT < ntask) {
                const float* col = sIn + trow[k] * NC + tcol[k];
                float o[RY];
#pragma unroll
                for (int r = 0; r < RY; r++) o[r] = 0.f;
                correlate<RY>(o, a.ky, a.wy, [&](int i) { return col[i * NC]; });
#pragma unroll
                for (int r = 0; r < RY; r++) sY[(trow[k] + r) * NCP + tcol[k]] = o[r];
            }
        }
        __syncthreads();

        // x-pass: sY -> registers
        float xv[CPT];
#pragma unroll
        for (int o = 0; o < CPT; o++) xv[o] = 0.f;
        {
            const float* row = sY + ry * NCP + xs;
            correlate<CPT>(xv, a.kx, a.wx, [&](int i) { return row[i]; });
        }

        // z-ring
        const int zout = z0 + s - (P - 1);
        TS* out = dst + (int64_t)zout * a.dz + (int64_t)(y0 + ry) * a.dy + (x0 + xs);
        ring_dispatch<0, P, CPT, TS>(s % P, ring, xv, s, a, out, row_ok, x0 + xs);
    }
}

template <class TS, int TX, int TY, int CPT, int P>
__global__ __launch_bounds__(TX * TY / CPT) void fused_tile_kernel(const FusedArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // each XCD gets a contiguous range of tiles when remapping (workgroups
    // are dealt round-robin over the 8 XCDs)
    uint32_t tile = blockIdx.x;
    if (a.xcd_remap) {
        const uint32_t per = a.ntiles >> 3;
        tile = (tile & 7u) * per + (tile >> 3);
    }
    fused_tile<TS, TX, TY, CPT, P>(a, tile, smem_raw);
}

// Persistent work queue over many images with the same z tap count:
// workgroups take tiles from one counter until all `total` tiles of all jobs
// are done.  tile_start[j] is the first global tile of job j (prefix sums,
// tile_start[njobs] = total).
template <class TS, int TX, int TY, int CPT, int P>
__global__ __launch_bounds__(TX * TY / CPT) void fused_queue_kernel(const FusedArgs* __restrict__ jobs,
                                                          const uint32_t* __restrict__ tile_start, int njobs,
                                                          uint32_t total, uint32_t* counter)
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
        fused_tile<TS, TX, TY, CPT, P>(jobs[lo], t - tile_start[lo], smem_raw);
        __syncthreads(); // LDS is reused by the next tile
    }
}

__global__ void reset_counter_kernel(uint32_t* c) { *c = 0u; }

// ------------------------------------------------------------- dispatch
constexpr int kTX = 64, kTY = 16;

template <class F> hipError_t with_ring(int ring, F&& f)
{
    switch (ring) {
#define NDL_RING(R) case R: return f(std::integral_constant<int, R>{});
        NDL_RING(1) NDL_RING(3) NDL_RING(5) NDL_RING(7) NDL_RING(9) NDL_RING(11) NDL_RING(13)
        NDL_RING(15) NDL_RING(17) NDL_RING(19) NDL_RING(21) NDL_RING(23) NDL_RING(25)
#undef NDL_RING
    }
    return hipErrorInvalidValue;
}

template <class F> hipError_t with_storage(int dt, F&& f)
{
    switch (dt) {
    case NDL_F32: return f(float{});
    case NDL_F16: return f(f16_t{});
    case NDL_BF16: return f(bf16_t{});
    case NDL_U8: return f(uint8_t{});
    }
    return hipErrorInvalidValue;
}

template <class F> hipError_t with_cpt(int cpt, F&& f)
{
    switch (cpt) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    }
    return hipErrorInvalidValue;
}

template <class F> hipError_t with_instance(const FusedConfig& cfg, int dt, F&& f)
{
    if (cfg.tx != kTX || cfg.ty != kTY) return hipErrorInvalidValue;
    return with_storage(dt, [&](auto ts) {
        return with_cpt(cfg.cpt, [&](auto c) { return with_ring(cfg.ring, [&](auto r) { return f(ts, c, r); }); });
    });
}

} // namespace

int fused_ring_for(int kz)
{
    return (kz >= 1 && kz <= 25 && (kz & 1)) ? kz : -1;
}

size_t fused_lds_bytes(const FusedArgs& a, const FusedConfig& cfg, bool)
{
    const int NC = cfg.tx + a.kx - 1;
    const int NR = cfg.ty + a.ky - 1;
    const int NCP = (NC + 3 + 3) & ~3;
    return (size_t)((((NR + 3) * NC + 3) & ~3) + cfg.ty * NCP) * 4;
}

hipError_t launch_fused(const FusedArgs& a, const FusedConfig& cfg, int dt, bool, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    if (cfg.ring != a.kz) return hipErrorInvalidValue;
    const size_t lds = fused_lds_bytes(a, cfg, false);
    return with_instance(cfg, dt, [&](auto ts, auto c, auto r) {
        using TS = decltype(ts);
        constexpr int CPT = decltype(c)::value, NT = kTX * kTY / CPT;
        hipLaunchKernelGGL((fused_tile_kernel<TS, kTX, kTY, CPT, decltype(r)::value>), dim3(a.ntiles), dim3(NT), lds, s,
                           a);
        return hipGetLastError();
    });
}

hipError_t launch_fused_queue(const FusedQueue& q, const FusedConfig& cfg, int dt, bool, size_t lds, hipStream_t s)
{
    if (q.total == 0) return hipSuccess;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, q.counter);
    return with_instance(cfg, dt, [&](auto ts, auto c, auto r) {
        using TS = decltype(ts);
        constexpr int CPT = decltype(c)::value, NT = kTX * kTY / CPT;
        hipLaunchKernelGGL((fused_queue_kernel<TS, kTX, kTY, CPT, decltype(r)::value>), dim3(q.grid), dim3(NT), lds, s,
                           q.jobs, q.tile_start, q.njobs, q.total, q.counter);
        return hipGetLastError();
    });
}

int fused_blocks_per_cu(const FusedConfig& cfg, int dt, size_t lds)
{
    int n = 0;
    hipError_t e = with_instance(cfg, dt, [&](auto ts, auto c, auto r) {
        using TS = decltype(ts);
        constexpr int CPT = decltype(c)::value, NT = kTX * kTY / CPT;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void*>(&fused_tile_kernel<TS, kTX, kTY, CPT, decltype(r)::value>), NT, lds);
    });
    return e == hipSuccess ? n : 0;
}

} // namespace ndl
