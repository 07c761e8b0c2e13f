// kernels_line.hip -- one-axis passes along a dimension that is not the
// unit-stride one (a 1-D kernel image {1, k, 1} or {1, 1, k} given to
// convolve(), processing/convolution.h:52-71, or one per-axis pass of
// gaussian_blur(), convolution.h:107-186).
//
// Every lane owns V consecutive elements of the unit-stride dimension x of
// one line (a fixed index of every other dimension) and marches along the
// convolved axis z.  The input slices it has read sit in a ring of K + D
// registers per element: slice i+D is requested while output i-(K-1) is
// computed, so D loads per lane stay in flight.  z positions, their border
// fold and the taps are the same for the whole workgroup (scalar registers).
// Lanes are numbered x-fastest over (x / V, line), so a workgroup reads
// 256 x V x 4 contiguous bytes per slice where x rows are dense.  No LDS and
// no barrier.
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <int V> using IC = std::integral_constant<int, V>;

template <int V> __device__ __forceinline__ void load_v(const float* p, float* v)
{
    if constexpr (V == 4) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else if constexpr (V == 2) {
        const float2 q = *reinterpret_cast<const float2*>(p);
        v[0] = q.x; v[1] = q.y;
    } else {
        v[0] = *p;
    }
}
template <int V> __device__ __forceinline__ void store_v(float* p, const float* v)
{
    if constexpr (V == 4) *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    else if constexpr (V == 2) *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    else *p = v[0];
}

template <int K, int D, int V>
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
    int64_t so = (int64_t)xv * V, dof = (int64_t)xv * V;
    for (int i = 0; i < a.nb; i++) {
        const uint32_t q = a.bext[i].div(b);
        const uint32_t r = b - q * a.bext[i].d;
        so += (int64_t)r * a.sb[i];
        dof += (int64_t)r * a.db[i];
        b = q;
    }
    const float* scol = (const float*)a.src + so;
    float* dcol = (float*)a.dst + dof;
    const int z0 = (int)chunk * a.tz;
    const int n_out = min(a.tz, a.Z - z0);
    const int n_in = n_out + K - 1;
    const int p0 = z0 - a.c;  // position of input 0

    float ring[R][V];
    auto load = [&](int i, float* v) {
        const int p = safe_map32(a.border, a.Z, p0 + i);
        load_v<V>(scol + (int64_t)p * a.sz, v);
    };
#pragma unroll
    for (int j = 0; j < D; j++)
        if (j < n_in) load(j, ring[j]);
    for (int i0 = 0; i0 < n_in; i0 += R) {
#pragma unroll
        for (int u = 0; u < R; u++) {
            const int i = i0 + u;
            if (i + D < n_in) load(i + D, ring[(u + D) % R]);
            if (i >= K - 1 && i < n_in) {
                float acc[V];
#pragma unroll
                for (int e = 0; e < V; e++) acc[e] = 0.f;
#pragma unroll
                for (int t = 0; t < K; t++)
#pragma unroll
                    for (int e = 0; e < V; e++) acc[e] = fmaf(a.w[t], ring[(u - (K - 1) + t + R) % R][e], acc[e]);
                if (live) store_v<V>(dcol + (int64_t)(z0 + i - (K - 1)) * a.dz, acc);
            }
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
template <int K, int V>
__global__ __launch_bounds__(256) void row_pass_kernel(const RowArgs a)
{
    constexpr int L = 64 * V;
    constexpr int HLMAX = (K - 1 + 3) & ~3;            // >= the aligned left halo for any centre
    constexpr int W = (HLMAX + L + K + 3) & ~3;         // window floats per wave
    __shared__ __align__(16) float win[4][W];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    float* buf = win[wv];
    const int c = a.c;
    const int hl = (c + 3) & ~3;                        // main part starts 16-byte aligned
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
    auto issue = [&](uint32_t s, float* m, float& h) {
        int64_t so, dof;
        int x0;
        row_of(s, so, dof, x0);
        const float* row = (const float*)a.src + so;
        const int xm = x0 + V * lane;
        if (xm + V - 1 < a.X) {
            load_v<V>(row + xm, m);
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
    float m[V], h = 0.f;
    if (seg < a.total) issue(seg, m, h);
    for (; seg < a.total; seg += nwaves) {
        wave_lds_sync();  // the previous segment's window reads are done
        store_v<V>(buf + hl + V * lane, m);
        if (lane < K - 1) buf[lane < c ? hl - c + lane : hl + L + (lane - c)] = h;
        wave_lds_sync();
        if (seg + nwaves < a.total) issue(seg + nwaves, m, h);
        float wv_[K + V - 1];
#pragma unroll
        for (int j = 0; j < K + V - 1; j++) wv_[j] = buf[hl - c + V * lane + j];
        float acc[V];
#pragma unroll
        for (int e = 0; e < V; e++) acc[e] = 0.f;
#pragma unroll
        for (int t = 0; t < K; t++)
#pragma unroll
            for (int e = 0; e < V; e++) acc[e] = fmaf(a.w[t], wv_[e + t], acc[e]);
        int64_t so, dof;
        int x0;
        row_of(seg, so, dof, x0);
        float* drow = (float*)a.dst + dof;
        const int xo = x0 + V * lane;
        if (xo + V - 1 < a.X) {
            store_v<V>(drow + xo, acc);
        } else {
#pragma unroll
            for (int e = 0; e < V; e++)
                if (xo + e < a.X) drow[xo + e] = acc[e];
        }
    }
}

template <class F> hipError_t with_row_instance(int dt, int k, int v, F&& f)
{
    if (dt != NDL_F32 || (v != 1 && v != 4)) return hipErrorInvalidValue;
#define NDL_ROWK(KK)                                      \
    if (k == KK) {                                        \
        if (v == 4) return f(IC<KK>{}, IC<4>{});          \
        return f(IC<KK>{}, IC<1>{});                      \
    }
    NDL_ROWK(3) NDL_ROWK(5) NDL_ROWK(7) NDL_ROWK(9) NDL_ROWK(11) NDL_ROWK(13) NDL_ROWK(15) NDL_ROWK(17)
    NDL_ROWK(19) NDL_ROWK(21) NDL_ROWK(23) NDL_ROWK(25)
#undef NDL_ROWK
    return hipErrorInvalidValue;
}

// f(K, D, V) for the supported combinations (float32 storage)
template <class F> hipError_t with_line_instance(int dt, int k, int v, int d, F&& f)
{
    if (dt != NDL_F32) return hipErrorInvalidValue;
#define NDL_LINE(KK)                                                  \
    if (k == KK) {                                                    \
        if (v == 1 && d == 8) return f(IC<KK>{}, IC<8>{}, IC<1>{});   \
        if (v == 2 && d == 8) return f(IC<KK>{}, IC<8>{}, IC<2>{});   \
        if (v == 4 && d == 4) return f(IC<KK>{}, IC<4>{}, IC<4>{});   \
        if (v == 4 && d == 8) return f(IC<KK>{}, IC<8>{}, IC<4>{});   \
        return hipErrorInvalidValue;                                  \
    }
    NDL_LINE(3) NDL_LINE(5) NDL_LINE(7) NDL_LINE(9) NDL_LINE(11) NDL_LINE(13) NDL_LINE(15) NDL_LINE(17)
    NDL_LINE(19) NDL_LINE(21) NDL_LINE(23) NDL_LINE(25)
#undef NDL_LINE
    return hipErrorInvalidValue;
}

} // namespace

bool line_march_supported(int dt, int k, int v, int d)
{
    return with_line_instance(dt, k, v, d, [](auto, auto, auto) { return hipSuccess; }) == hipSuccess;
}

int line_march_blocks_per_cu(int dt, int k, int v, int d)
{
    int n = 0;
    hipError_t e = with_line_instance(dt, k, v, d, [&](auto kk, auto dd, auto vv) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n,
            reinterpret_cast<const void*>(
                &line_march_kernel<decltype(kk)::value, decltype(dd)::value, decltype(vv)::value>),
            256, 0);
    });
    return e == hipSuccess ? n : 0;
}

bool row_pass_supported(int dt, int k, int v)
{
    return with_row_instance(dt, k, v, [](auto, auto) { return hipSuccess; }) == hipSuccess;
}

int row_pass_blocks_per_cu(int dt, int k, int v)
{
    int n = 0;
    hipError_t e = with_row_instance(dt, k, v, [&](auto kk, auto vv) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void*>(&row_pass_kernel<decltype(kk)::value, decltype(vv)::value>), 256, 0);
    });
    return e == hipSuccess ? n : 0;
}

hipError_t launch_row_pass(const RowArgs& a, int dt, int v, uint32_t blocks, hipStream_t s)
{
    if (a.total == 0 || blocks == 0) return hipSuccess;
    return with_row_instance(dt, a.k, v, [&](auto kk, auto vv) {
        hipLaunchKernelGGL((row_pass_kernel<decltype(kk)::value, decltype(vv)::value>), dim3(blocks), dim3(256), 0, s,
                           a);
        return hipGetLastError();
    });
}

hipError_t launch_line_march(const LineArgs& a, int dt, int v, int d, hipStream_t s)
{
    const uint64_t blocks = (uint64_t)a.tiles_z * ((a.lanes + 255u) / 256u);
    if (blocks == 0) return hipSuccess;
    return with_line_instance(dt, a.k, v, d, [&](auto kk, auto dd, auto vv) {
        hipLaunchKernelGGL((line_march_kernel<decltype(kk)::value, decltype(dd)::value, decltype(vv)::value>),
                           dim3((uint32_t)blocks), dim3(256), 0, s, a);
        return hipGetLastError();
    });
}

} // namespace ndl
