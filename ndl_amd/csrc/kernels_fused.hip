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
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <class Acc> __device__ __forceinline__ Acc fmac(Acc acc, Acc v, Acc w);
template <> __device__ __forceinline__ float fmac<float>(float acc, float v, float w) { return __builtin_fmaf(v, w, acc); }
template <> __device__ __forceinline__ int32_t fmac<int32_t>(int32_t acc, int32_t v, int32_t w) { return acc + v * w; }

template <class TS, class Acc> __device__ __forceinline__ Acc load_s(const TS* p)
{
    if constexpr (is_int_storage<TS>::value) return (Acc)*p;
    else return load_f32(p);
}
template <class TS, class Acc> __device__ __forceinline__ void store_s(TS* p, Acc v)
{
    if constexpr (is_int_storage<TS>::value) *p = (TS)(uint32_t)v; // exact sum, low bits (convolution.h:69)
    else store_from_f32(p, v);
}

template <class Acc> struct Taps;
template <> struct Taps<float> {
    static __device__ __forceinline__ float x(const FusedArgs& a, int t) { return a.wx[t]; }
    static __device__ __forceinline__ float y(const FusedArgs& a, int t) { return a.wy[t]; }
    static __device__ __forceinline__ float z(const FusedArgs& a, int l) { return a.wzlag[l]; }
};
template <> struct Taps<int32_t> {
    static __device__ __forceinline__ int32_t x(const FusedArgs& a, int t) { return a.iwx[t]; }
    static __device__ __forceinline__ int32_t y(const FusedArgs& a, int t) { return a.iwy[t]; }
    static __device__ __forceinline__ int32_t z(const FusedArgs& a, int l) { return a.iwzlag[l]; }
};

// Ring step U of period P: store this slice's x-filtered values in slot U; if
// kz slices have arrived, emit output z = z0 + s - (kz-1) as
// sum_t w[t] * in[z + t - cz] with t ascending = lag l = kz-1-t descending.
template <int U, int P, int CPT, class TS, class Acc>
__device__ __forceinline__ void ring_step(Acc (&ring)[CPT][P], const Acc (&xv)[CPT], int s,
                                          const FusedArgs& a, TS* out, bool row_ok, int x_first)
{
#pragma unroll
    for (int o = 0; o < CPT; o++) ring[o][U] = xv[o];
    if (s < a.kz - 1) return;
    Acc acc[CPT];
#pragma unroll
    for (int o = 0; o < CPT; o++) acc[o] = 0;
#pragma unroll
    for (int l = P - 1; l >= 0; l--) {
        if (l < a.kz) {
            const Acc w = Taps<Acc>::z(a, l);
            if (w != (Acc)0) {
#pragma unroll
                for (int o = 0; o < CPT; o++) acc[o] = fmac<Acc>(acc[o], ring[o][(U - l + P) % P], w);
            }
        }
    }
    if (row_ok) {
#pragma unroll
        for (int o = 0; o < CPT; o++)
            if (x_first + o < a.X) store_s<TS, Acc>(out + o, acc[o]);
    }
}

template <int U, int P, int CPT, class TS, class Acc>
__device__ __forceinline__ void ring_dispatch(int u, Acc (&ring)[CPT][P], const Acc (&xv)[CPT], int s,
                                              const FusedArgs& a, TS* out, bool row_ok, int x_first)
{
    if constexpr (U < P) {
        if (u == U) {
            ring_step<U, P, CPT, TS, Acc>(ring, xv, s, a, out, row_ok, x_first);
            return;
        }
        ring_dispatch<U + 1, P, CPT, TS, Acc>(u, ring, xv, s, a, out, row_ok, x_first);
    }
}

// One tile: `tile` indexes (x tile, y tile, z chunk, batch) with x fastest.
template <class TS, class Acc, int TX, int TY, int P, int KMAX>
__device__ __forceinline__ void fused_tile(const FusedArgs& a, uint32_t tile, unsigned char* smem_raw)
{
    constexpr int NT = 256;
    constexpr int CPT = TX * TY / NT;
    constexpr int XG = TX / CPT;
    static_assert(CPT >= 1 && XG * CPT == TX && XG * TY == NT, "tile must map CPT x-outputs per thread");
    constexpr int RY = TY >= 4 ? 4 : TY;
    static_assert(TY % RY == 0, "TY must be a multiple of RY");
    constexpr int NCMAX = TX + KMAX - 1;
    constexpr int NRMAX = TY + KMAX - 1;
    constexpr int NLMAX = (NCMAX * NRMAX + NT - 1) / NT;
    constexpr int NTASKMAX = (NCMAX * (TY / RY) + NT - 1) / NT;

    const int NC = TX + a.kx - 1;
    const int NR = TY + a.ky - 1;
    const int NCP = (NC + 3) & ~3;
    const int NIN = NR * NC;
    Acc* sIn = (Acc*)smem_raw;
    Acc* sY = sIn + ((NIN + 3) & ~3);

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
    const int nsl = nz + a.kz - 1;
    const TS* src = (const TS*)a.src + soff;
    TS* dst = (TS*)a.dst + doff;
    const int tid = threadIdx.x;

    // ---- per-thread load slots: element i = tid + k*NT of the window
    int32_t loff[NLMAX];
#pragma unroll
    for (int k = 0; k < NLMAX; k++) {
        const int i = tid + k * NT;
        loff[k] = 0;
        if (i < NIN) {
            const int r = i / NC;
            const int c = i - r * NC;
            const int yy = safe_map32(a.by, a.Y, y0 - a.cy + r);
            const int xx = safe_map32(a.bx, a.X, x0 - a.cx + c);
            loff[k] = (int32_t)((int64_t)yy * a.sy + xx);
        }
    }
    // ---- y-pass tasks: column c, rows [r0, r0+RY)
    int tcol[NTASKMAX], trow[NTASKMAX];
#pragma unroll
    for (int k = 0; k < NTASKMAX; k++) {
        const int task = tid + k * NT;
        const int rg = task / NC;
        tcol[k] = task - rg * NC;
        trow[k] = rg * RY;
    }
    const int ntask = NC * (TY / RY);
    // ---- x-pass / z-ring ownership
    const int ry = tid / XG;
    const int xs = (tid - ry * XG) * CPT;
    const bool row_ok = y0 + ry < a.Y;

    Acc pre[NLMAX];
    auto load_slice = [&](int s) {
        const int zz = safe_map32(a.bz, a.Z, z0 - a.cz + s);
        const TS* base = src + (int64_t)zz * a.sz;
#pragma unroll
        for (int k = 0; k < NLMAX; k++)
            if (tid + k * NT < NIN) pre[k] = load_s<TS, Acc>(base + loff[k]);
    };

    Acc ring[CPT][P];
#pragma unroll
    for (int o = 0; o < CPT; o++)
#pragma unroll
        for (int p = 0; p < P; p++) ring[o][p] = 0;

    load_slice(0);
    for (int s = 0; s < nsl; s++) {
#pragma unroll
        for (int k = 0; k < NLMAX; k++)
            if (tid + k * NT < NIN) sIn[tid + k * NT] = pre[k];
        __syncthreads();
        if (s + 1 < nsl) load_slice(s + 1);

        // y-pass: sIn -> sY
#pragma unroll
        for (int k = 0; k < NTASKMAX; k++) {
            if (tid + k * NT < ntask) {
                const Acc* col = sIn + trow[k] * NC + tcol[k];
                Acc win[RY + KMAX - 1];
#pragma unroll
                for (int i = 0; i < RY + KMAX - 1; i++) win[i] = (i < RY + a.ky - 1) ? col[i * NC] : (Acc)0;
                Acc o[RY];
#pragma unroll
                for (int r = 0; r < RY; r++) o[r] = 0;
#pragma unroll
                for (int t = 0; t < KMAX; t++) {
                    if (t < a.ky) {
                        const Acc w = Taps<Acc>::y(a, t);
                        if (w != (Acc)0) {
#pragma unroll
                            for (int r = 0; r < RY; r++) o[r] = fmac<Acc>(o[r], win[r + t], w);
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < RY; r++) sY[(trow[k] + r) * NCP + tcol[k]] = o[r];
            }
        }
        __syncthreads();

        // x-pass: sY -> registers
        Acc xv[CPT];
        {
            const Acc* row = sY + ry * NCP + xs;
            Acc win[CPT + KMAX - 1];
#pragma unroll
            for (int i = 0; i < CPT + KMAX - 1; i++) win[i] = (i < CPT + a.kx - 1) ? row[i] : (Acc)0;
#pragma unroll
            for (int o = 0; o < CPT; o++) xv[o] = 0;
#pragma unroll
            for (int t = 0; t < KMAX; t++) {
                if (t < a.kx) {
                    const Acc w = Taps<Acc>::x(a, t);
                    if (w != (Acc)0) {
#pragma unroll
                        for (int o = 0; o < CPT; o++) xv[o] = fmac<Acc>(xv[o], win[o + t], w);
                    }
                }
            }
        }

        // z-ring
        const int zout = z0 + s - (a.kz - 1);
        TS* out = dst + (int64_t)zout * a.dz + (int64_t)(y0 + ry) * a.dy + (x0 + xs);
        ring_dispatch<0, P, CPT, TS, Acc>(s % P, ring, xv, s, a, out, row_ok, x0 + xs);
    }
}

template <class TS, class Acc, int TX, int TY, int P, int KMAX>
__global__ __launch_bounds__(256) void fused_tile_kernel(const FusedArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // each XCD gets a contiguous range of tiles when remapping (workgroups
    // are dealt round-robin over the 8 XCDs)
    uint32_t tile = blockIdx.x;
    if (a.xcd_remap) {
        const uint32_t per = a.ntiles >> 3;
        tile = (tile & 7u) * per + (tile >> 3);
    }
    fused_tile<TS, Acc, TX, TY, P, KMAX>(a, tile, smem_raw);
}

// Persistent work queue over many images: workgroups take tiles from one
// counter until all `total` tiles of all jobs are done.  tile_start[j] is the
// first global tile of job j (prefix sums, tile_start[njobs] = total).
template <class TS, class Acc, int TX, int TY, int P, int KMAX>
__global__ __launch_bounds__(256) void fused_queue_kernel(const FusedArgs* __restrict__ jobs,
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
        fused_tile<TS, Acc, TX, TY, P, KMAX>(jobs[lo], t - tile_start[lo], smem_raw);
        __syncthreads(); // LDS is reused by the next tile
    }
}

__global__ void reset_counter_kernel(uint32_t* c) { *c = 0u; }

// ------------------------------------------------------------- dispatch
constexpr int kRings[] = {1, 5, 9, 13, 19, 25};

template <class TS, class Acc, int TX, int TY, int KMAX>
hipError_t launch_ring(const FusedArgs& a, int ring, size_t lds, hipStream_t s)
{
    dim3 grid(a.ntiles), block(256);
    switch (ring) {
    case 1: hipLaunchKernelGGL((fused_tile_kernel<TS, Acc, TX, TY, 1, KMAX>), grid, block, lds, s, a); break;
    case 5: hipLaunchKernelGGL((fused_tile_kernel<TS, Acc, TX, TY, 5, KMAX>), grid, block, lds, s, a); break;
    case 9: hipLaunchKernelGGL((fused_tile_kernel<TS, Acc, TX, TY, 9, KMAX>), grid, block, lds, s, a); break;
    case 13: hipLaunchKernelGGL((fused_tile_kernel<TS, Acc, TX, TY, 13, KMAX>), grid, block, lds, s, a); break;
    case 19: hipLaunchKernelGGL((fused_tile_kernel<TS, Acc, TX, TY, 19, KMAX>), grid, block, lds, s, a); break;
    case 25: hipLaunchKernelGGL((fused_tile_kernel<TS, Acc, TX, TY, 25, KMAX>), grid, block, lds, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class TS, class Acc>
hipError_t launch_tile(const FusedArgs& a, const FusedConfig& cfg, size_t lds, hipStream_t s)
{
    if (cfg.kmax != 25) return hipErrorInvalidValue;
    if (cfg.tx == 64 && cfg.ty == 16) return launch_ring<TS, Acc, 64, 16, 25>(a, cfg.ring, lds, s);
    return hipErrorInvalidValue;
}

template <class TS, class Acc, int TX, int TY, int KMAX>
hipError_t launch_queue_ring(const FusedQueue& q, int ring, size_t lds, hipStream_t s)
{
    dim3 grid(q.grid), block(256);
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, q.counter);
    switch (ring) {
#define NDL_QCASE(R)                                                                                         \
    case R:                                                                                                  \
        hipLaunchKernelGGL((fused_queue_kernel<TS, Acc, TX, TY, R, KMAX>), grid, block, lds, s, q.jobs,      \
                           q.tile_start, q.njobs, q.total, q.counter);                                      \
        break;
        NDL_QCASE(1) NDL_QCASE(5) NDL_QCASE(9) NDL_QCASE(13) NDL_QCASE(19) NDL_QCASE(25)
#undef NDL_QCASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class TS, class Acc>
hipError_t launch_queue_tile(const FusedQueue& q, const FusedConfig& cfg, size_t lds, hipStream_t s)
{
    if (cfg.kmax != 25) return hipErrorInvalidValue;
    if (cfg.tx == 64 && cfg.ty == 16) return launch_queue_ring<TS, Acc, 64, 16, 25>(q, cfg.ring, lds, s);
    return hipErrorInvalidValue;
}

} // namespace

hipError_t launch_fused_queue(const FusedQueue& q, const FusedConfig& cfg, int dt, bool int_acc, size_t lds,
                              hipStream_t s)
{
    if (q.total == 0) return hipSuccess;
    if (int_acc) {
        if (dt == NDL_U8) return launch_queue_tile<uint8_t, int32_t>(q, cfg, lds, s);
        return hipErrorInvalidValue;
    }
    switch (dt) {
    case NDL_F32: return launch_queue_tile<float, float>(q, cfg, lds, s);
    case NDL_F16: return launch_queue_tile<f16_t, float>(q, cfg, lds, s);
    case NDL_BF16: return launch_queue_tile<bf16_t, float>(q, cfg, lds, s);
    }
    return hipErrorInvalidValue;
}

int fused_ring_for(int kz)
{
    for (int r : kRings)
        if (kz <= r) return r;
    return -1;
}

size_t fused_lds_bytes(const FusedArgs& a, const FusedConfig& cfg, bool acc8)
{
    const int NC = cfg.tx + a.kx - 1;
    const int NR = cfg.ty + a.ky - 1;
    const int NCP = (NC + 3) & ~3;
    return (size_t)(((NR * NC + 3) & ~3) + cfg.ty * NCP) * (acc8 ? 8 : 4);
}

hipError_t launch_fused(const FusedArgs& a, const FusedConfig& cfg, int dt, bool int_acc, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    const size_t lds = fused_lds_bytes(a, cfg, false);
    if (int_acc) {
        switch (dt) {
        case NDL_U8: return launch_tile<uint8_t, int32_t>(a, cfg, lds, s);
        }
        return hipErrorInvalidValue;
    }
    switch (dt) {
    case NDL_F32: return launch_tile<float, float>(a, cfg, lds, s);
    case NDL_F16: return launch_tile<f16_t, float>(a, cfg, lds, s);
    case NDL_BF16: return launch_tile<bf16_t, float>(a, cfg, lds, s);
    }
    return hipErrorInvalidValue;
}

} // namespace ndl
