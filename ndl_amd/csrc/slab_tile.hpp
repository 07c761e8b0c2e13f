// slab_tile.hpp -- fused separable kernel over whole rows ("slab" tiles),
// up to three convolved axes (x = unit stride, y, z) in one pass over HBM.
//
// A workgroup owns every column of TY = 4 consecutive rows (a row of X <=
// 256 elements) and a chunk of tz output slices along z.  Its waves take one
// of two roles:
//   Y waves: one column and the 4 rows per lane.  The 4 + ky - 1 input rows
//            of a slice are read straight from global memory (buffer loads,
//            consecutive lanes on consecutive columns), D slices ahead into
//            registers; the y-pass result goes to an LDS y-plane whose halo
//            columns are filled by the lanes owning the fold sources (the x
//            border of a whole row is a fold of the row itself, so no x halo
//            is ever loaded or filtered twice).
//   XZ waves: two consecutive columns of one row per lane: the x-pass from
//            8-byte LDS reads of the y-plane window, then a register ring of
//            the last kz x-filtered slices; once kz slices are in, one z
//            output pair per slice, one 8-byte store (512 contiguous bytes
//            per wave).
// The y-planes form a ring of 8 buffers in LDS, two groups of 4 slices.
// The roles synchronise once per group with LDS counters instead of a
// workgroup barrier per slice: Y waves count the groups they wrote, XZ waves
// the groups they read, and each side waits (bounded spin) only for the
// group it needs, so a Y wave whose loads arrive late stalls the XZ waves
// only once they have used up the group ahead.
// Only the z halo of a chunk (kz - 1 slices) is filtered twice; with one
// long chunk per workgroup that is the only recomputation of the pass.
//
// Accumulation: FP32 FMA, taps ascending within each pass, passes y, x, z.
#pragma once
#include "fused_tile.hpp"

namespace ndl {
namespace slab_detail {
using fused_detail::IC;
using fused_detail::static_for;

constexpr int kTY = 4;     // rows per tile
constexpr int kTXW = 256;  // columns per tile row

// s_waitcnt immediate for lgkmcnt(0) (gfx9 encoding: vmcnt in bits 3:0 and
// 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int kWaitLgkm0 = 0xc07f;

// Wait until the LDS counter at cnt reaches target (other waves of the
// workgroup add to it).  The spin is bounded (2^21 polls with a short sleep,
// far beyond any real wait) so no wave can spin forever, and written in asm
// so the register allocator sees straight-line code around it.
__device__ __forceinline__ void wait_ge(const int* cnt, int target)
{
    const uint32_t addr = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const int*)cnt;
    int v, sv, it;
    __asm__ volatile(
        "s_mov_b32 %[it], 0x200000\n"
        "1:\n\t"
        "ds_read_b32 %[v], %[addr]\n\t"
        "s_waitcnt lgkmcnt(0)\n\t"
        "v_readfirstlane_b32 %[s], %[v]\n\t"
        "s_cmp_ge_i32 %[s], %[tgt]\n\t"
        "s_cbranch_scc1 2f\n\t"
        "s_sleep 1\n\t"
        "s_sub_u32 %[it], %[it], 1\n\t"
        "s_cmp_eq_u32 %[it], 0\n\t"
        "s_cbranch_scc0 1b\n"
        "2:"
        : [v] "=&v"(v), [s] "=&s"(sv), [it] "=&s"(it)
        : [addr] "v"(addr), [tgt] "s"(target)
        : "scc", "memory");
}
__device__ __forceinline__ void signal(int* cnt)
{
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0); // this wave's LDS writes are done
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __asm__ volatile("" ::: "memory");
}

// element pair store (8-byte aligned: the host checks)
__device__ __forceinline__ void store_pair(float* p, float v0, float v1, bool ok0, bool ok1)
{
    if (ok0 && ok1) {
        *reinterpret_cast<float2*>(p) = make_float2(v0, v1);
        return;
    }
    if (ok0) p[0] = v0;
    if (ok1) p[1] = v1;
}

// window geometry of the LDS y-plane: column x lives at x + CXS, so the
// x-pass window of four outputs at c4 starts at LDS column c4 (aligned)
// (halo columns on both sides; lanes past the row write their values to
// columns beyond the right halo, which nothing reads)
template <int KX> struct XWin {
    static constexpr int CXS = (KX - 1) / 2;                      // centre (odd, centred kernels)
    static constexpr int PITCH = (kTXW + 2 * (KX - 1) + 3) & ~3; // row pitch, floats
};

typedef unsigned int slab_u2 __attribute__((ext_vector_type(2)));

// Step loops are split into blocks whose steps need no per-step checks
// (the steady state) and short checked prologue / tail blocks: a per-step
// condition costs scalar instructions, and the scalar unit is shared by all
// waves of a CU.
template <int KX, int KY, int KZ, int D>
__device__ __forceinline__ void slab_tile(const FusedArgs& a, uint32_t tile)
{
    constexpr int TY = kTY, TXW = kTXW;
    constexpr int NWY = TXW / 64;         // Y waves (one column per lane)
    constexpr int NRY = TY + KY - 1;      // input rows of a slice window
    using W = XWin<KX>;
    constexpr int YP = W::PITCH;
    constexpr int NXW = KX + 1;           // x window of two outputs
    constexpr int PLANE = TY * YP;        // floats per y-plane
    // the roles synchronise once per group of P slices: NR = 2 P y-planes,
    // group k in slot k % 2
    constexpr int P = 4, G = 2, NR = G * P;
    static_assert(KX % 2 == 1 && KX <= 17, "odd x kernels up to 17 taps");
    static_assert((D - 1) * NRY <= 63, "loads in flight exceed the wait counter");
    __shared__ __align__(16) float ypl[NR * PLANE];
    // per group slot: Y waves that wrote its planes / XZ waves that read
    // them, summed over its uses (use u of slot g is group u * G + g)
    __shared__ int ycnt[G];
    __shared__ int xcnt[G];

    // tile -> (y tile, z chunk, batch), y fastest
    uint32_t q = a.fdy.div(tile);
    const int tiy = (int)(tile - q * (uint32_t)a.tiles_y);
    uint32_t bi = a.fdz.div(q);
    const int tiz = (int)(q - bi * (uint32_t)a.tiles_z);
    int64_t soff = 0, doff = 0;
    for (int d = 0; d < a.nb; d++) {
        uint32_t qq = a.bdiv[d].div(bi);
        int64_t c = (int64_t)(bi - qq * (uint32_t)a.bext[d]);
        bi = qq;
        soff += c * a.sb[d];
        doff += c * a.db[d];
    }
    const int y0 = tiy * TY, z0 = tiz * a.tz;
    const int nz = min(a.tz, a.Z - z0);
    const int nsl = nz + KZ - 1; // input slices of the chunk
    const int X = a.X;
    const float* src = (const float*)a.src + soff;
    float* dst = (float*)a.dst + doff;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const bool z_interior = z0 - a.cz >= 0 && z0 - a.cz + nsl <= a.Z;
    if (threadIdx.x < G) ycnt[threadIdx.x] = 0;
    if (threadIdx.x < G) xcnt[threadIdx.x] = 0;
    __syncthreads();

    if (w < NWY) {
        // ------------------------------------------------------ Y role
        const int c = w * 64 + lane;
        // lanes past the row load a valid column; byte offset of the column
        const uint32_t coff = (uint32_t)min(c, X - 1) * (uint32_t)sizeof(float);
        // byte offsets of the input rows within a slice (uniform; the host
        // keeps a slice's rows within 2 GiB at a non-negative row stride)
        int32_t roff[NRY];
#pragma unroll
        for (int i = 0; i < NRY; i++)
            roff[i] = (int32_t)((int64_t)safe_map32(a.by, a.Y, y0 - a.cy + i) * a.sy * (int64_t)sizeof(float));
        // LDS columns this lane's result goes to: its own, and the halo
        // column it is the fold source of (reflect / wrap; at most one per
        // lane for X >= 2 cx, which the host checks), or its own again.
        // Lanes past the row use columns beyond the right halo.  Every lane
        // writes both, so a step has no lane-dependent branch.
        const int own = c < X ? W::CXS + c : W::CXS + c + KX - 1;
        int hx = own;
        if (c < X) {
            if (a.bx == NDL_REFLECT) {
                if (c < W::CXS) hx = W::CXS - 1 - c;
                if (c >= X - W::CXS) hx = W::CXS + 2 * X - 1 - c;
            } else if (a.bx == NDL_WRAP) {
                if (c >= X - W::CXS) hx = W::CXS + c - X;
                if (c < W::CXS) hx = W::CXS + X + c;
            }
        }
        const bool clamp = a.bx == NDL_CLAMP;
        // slice byte offsets are 32-bit (the host keeps the source span
        // below 2 GiB); slice j of the chunk is input slice z0 - cz + j
        const int32_t szb = (int32_t)(a.sz * (int64_t)sizeof(float));
        // every step loads (past the last slice: the last slice again), so
        // no load sits under a branch and each wait counts exactly the D - 1
        // slices still in flight
        auto load = [&](int j, float (&buf)[NRY]) NDL_INLINE {
            j = min(j, nsl - 1);
            int zz = z0 - a.cz + j;
            if (!z_interior) zz = safe_map32(a.bz, a.Z, zz);
            const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)((const unsigned char*)src + zz * szb),
                                                              (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
            for (int i = 0; i < NRY; i++)
                buf[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, coff, roff[i], 0));
        };
        // y-pass of one slice into y-plane `par` (compile-time parity)
        auto ypass = [&](const float (&in)[NRY], int par) NDL_INLINE {
            float o[TY];
#pragma unroll
            for (int r = 0; r < TY; r++) o[r] = 0.f;
#pragma unroll
            for (int t = 0; t < KY; t++) {
                const float wt = a.wy[t];
#pragma unroll
                for (int r = 0; r < TY; r++) o[r] = __builtin_fmaf(in[r + t], wt, o[r]);
            }
            float* yp = ypl + par * PLANE;
#pragma unroll
            for (int r = 0; r < TY; r++) yp[r * YP + own] = o[r];
#pragma unroll
            for (int r = 0; r < TY; r++) yp[r * YP + hx] = o[r];
            if (clamp) {
                // clamp: every halo column repeats the edge column
                if (c == 0) {
#pragma unroll
                    for (int h = 0; h < W::CXS; h++)
#pragma unroll
                        for (int r = 0; r < TY; r++) yp[r * YP + h] = o[r];
                }
                if (c == X - 1) {
#pragma unroll
                    for (int h = 0; h < KX - 1 - W::CXS; h++)
#pragma unroll
                        for (int r = 0; r < TY; r++) yp[r * YP + W::CXS + X + h] = o[r];
                }
            }
        };
        float pre[D][NRY];
#pragma unroll
        for (int u = 0; u < D; u++) load(u, pre[u]);
        // slice sl: at the start of a group, wait until every XZ wave has
        // read the group's previous planes; the y-pass; at its end, count
        // the group written; then the loads of slice sl + D
        auto yslice = [&](auto ic, int sl) NDL_INLINE {
            constexpr int u = decltype(ic)::value;
            const int k = sl / P; // group
            if (sl % P == 0 && k >= G) wait_ge(&xcnt[k % G], 2 * NWY * (k / G)); // read by every XZ wave
            ypass(pre[u], sl & (NR - 1));
            if (sl % P == P - 1 || sl == nsl - 1) signal(&ycnt[k % G]);
            load(sl + D, pre[u]);
        };
        int s0 = 0;
        for (; s0 + D <= nsl; s0 += D)
            static_for<D>([&](auto ic) NDL_INLINE { yslice(ic, s0 + decltype(ic)::value); });
        static_for<D>([&](auto ic) NDL_INLINE {
            if (s0 + decltype(ic)::value < nsl) yslice(ic, s0 + decltype(ic)::value);
        });
    } else {
        // ------------------------------------------------------ XZ role
        constexpr int QPR = TXW / 2; // lanes per row
        const int qi = (w - NWY) * 64 + lane;
        const int row = __builtin_amdgcn_readfirstlane(qi / QPR);
        const int c2 = (qi - row * QPR) * 2;
        // stores through a buffer resource over one output row of one
        // slice: rows past Y get an empty one, lanes past X (X even) fall
        // outside it
        const bool row_ok = y0 + row < a.Y;
        const uint32_t nrec = row_ok ? (uint32_t)X * 4u : 0u;
        const uint32_t voff = (uint32_t)c2 * 4u;
        float* orow = dst + (int64_t)(row_ok ? y0 + row : 0) * a.dy + (int64_t)z0 * a.dz; // next output slice
        const float* ycol = ypl + row * YP + c2;
        float ring[2][KZ];
#pragma unroll
        for (int p = 0; p < KZ; p++) ring[0][p] = ring[1][p] = 0.f;
        // slice j: its x-pass into ring slot U = j % KZ, then (OUT) the z
        // output of j.  Slices run in blocks of KZ, so every slot is a
        // compile-time register; the first block fills the ring, full blocks
        // run without per-slice checks, a checked block takes the rest.
        auto xslice = [&](auto slot, int j, auto out) NDL_INLINE {
            constexpr int U = decltype(slot)::value;
            const int k = j / P; // group
            if (j % P == 0) wait_ge(&ycnt[k % G], NWY * (k / G + 1)); // written by every Y wave
            const float* yr = ycol + (j & (NR - 1)) * PLANE;
            float wv[NXW + (NXW & 1)];
#pragma unroll
            for (int m = 0; m < (NXW + 1) / 2; m++) {
                const float2 v = *reinterpret_cast<const float2*>(yr + 2 * m);
                wv[2 * m] = v.x;
                wv[2 * m + 1] = v.y;
            }
            float x0 = 0.f, x1 = 0.f;
#pragma unroll
            for (int t = 0; t < KX; t++) {
                const float wt = a.wx[t];
                x0 = __builtin_fmaf(wv[t], wt, x0);
                x1 = __builtin_fmaf(wv[t + 1], wt, x1);
            }
            __asm__ volatile("" :: "v"(x0), "v"(x1));
            if (j % P == P - 1 || j == nsl - 1) signal(&xcnt[k % G]); // the group's windows are in registers
            ring[0][U] = x0;
            ring[1][U] = x1;
            if constexpr (decltype(out)::value) {
                float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
                for (int l = KZ - 1; l >= 0; l--) {
                    const float wl = a.wzlag[l];
                    acc0 = __builtin_fmaf(ring[0][(U - l + KZ) % KZ], wl, acc0);
                    acc1 = __builtin_fmaf(ring[1][(U - l + KZ) % KZ], wl, acc1);
                }
                const auto rd = __builtin_amdgcn_make_buffer_rsrc((void*)orow, (short)0, nrec, 0x00020000);
                slab_u2 v = {__float_as_uint(acc0), __float_as_uint(acc1)};
                __builtin_amdgcn_raw_buffer_store_b64(v, rd, voff, 0, 0);
                orow += a.dz; // outputs leave in slice order
            }
        };
        static_for<KZ>([&](auto ic) NDL_INLINE {
            constexpr int u = decltype(ic)::value;
            xslice(IC<u>{}, u, std::integral_constant<bool, u == KZ - 1>{});
        });
        int j0 = KZ;
        for (; j0 + KZ <= nsl; j0 += KZ)
            static_for<KZ>([&](auto ic) NDL_INLINE { xslice(ic, j0 + decltype(ic)::value, std::true_type{}); });
        static_for<KZ>([&](auto ic) NDL_INLINE {
            const int j = j0 + decltype(ic)::value;
            if (j < nsl) xslice(ic, j, std::true_type{});
        });
    }
}

// threads per workgroup: NWY Y waves + 2 NWY XZ waves
constexpr int kSlabThreads = 3 * kTXW;

template <int KX, int KY, int KZ, int D>
__global__ __launch_bounds__(kSlabThreads) void slab_kernel(const FusedArgs a)
{
    uint32_t tile = blockIdx.x;
    if (a.xcd_remap) {
        const uint32_t per = a.ntiles >> 3;
        tile = (tile & 7u) * per + (tile >> 3);
    }
    slab_tile<KX, KY, KZ, D>(a, tile);
}

} // namespace slab_detail
} // namespace ndl
