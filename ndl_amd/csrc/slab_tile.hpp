// slab_tile.hpp -- fused separable kernel over whole rows ("slab" tiles),
// up to three convolved axes (x = unit stride, y, z) in one pass over HBM.
//
// A workgroup owns every column of TY = 4 consecutive rows (a row of X <=
// 256 elements) and a chunk of tz output slices along z.  Its waves take one
// of two roles and run a two-stage pipeline, one barrier per slice step
// (step s: y-pass of slice s, x- and z-pass of slice s-1):
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
//            output pair per step, one 8-byte store (512 contiguous bytes per
//            wave).
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

// s_waitcnt immediates (gfx9 encoding: vmcnt in bits 3:0 and 15:14,
// expcnt 6:4, lgkmcnt 11:8)
constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | 0x0f70; }
constexpr int kWaitLgkm0 = 0xc07f;

// s_barrier after this wave's LDS operations completed; compiler-level
// memory barriers on both sides keep every LDS access on its side
__device__ __forceinline__ void step_barrier()
{
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    __builtin_amdgcn_s_barrier();
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
    static_assert(KX % 2 == 1 && KX <= 17, "odd x kernels up to 17 taps");
    static_assert((D - 1) * NRY <= 63, "loads in flight exceed the wait counter");
    static_assert(D % 2 == 0, "even prefetch depth: compile-time y-plane parity");
    __shared__ __align__(16) float ypl[2 * PLANE];

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
    // pipeline steps: nsl + 1 (y-pass of slice s, x- and z-pass of s - 1),
    // rounded up to whole groups of D; every wave passes the same NSP barriers
    const int NSP = (nsl + 1 + D - 1) / D * D;
    const int X = a.X;
    const float* src = (const float*)a.src + soff;
    float* dst = (float*)a.dst + doff;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const bool z_interior = z0 - a.cz >= 0 && z0 - a.cz + nsl <= a.Z;

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
        auto ypass = [&](const float (&in)[NRY], auto par) NDL_INLINE {
            float o[TY];
#pragma unroll
            for (int r = 0; r < TY; r++) o[r] = 0.f;
#pragma unroll
            for (int t = 0; t < KY; t++) {
                const float wt = a.wy[t];
#pragma unroll
                for (int r = 0; r < TY; r++) o[r] = __builtin_fmaf(in[r + t], wt, o[r]);
            }
            float* yp = ypl + decltype(par)::value * PLANE;
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
        int s0 = 0;
        // steady state: every step of the block y-filters a slice (D is
        // even, so the y-plane of step s0 + u is u & 1)
        for (; s0 + D <= nsl; s0 += D) {
            static_for<D>([&](auto ic) NDL_INLINE {
                constexpr int u = decltype(ic)::value;
                ypass(pre[u], IC<u & 1>{});
                load(s0 + u + D, pre[u]);
                step_barrier();
            });
        }
        // tail (fewer than 2 D + 2 steps; the buffer of step s stays s % D)
        static_for<2 * D + 2>([&](auto ic) NDL_INLINE {
            constexpr int u = decltype(ic)::value;
            const int s = s0 + u;
            if (s < NSP) {
                if (s < nsl) ypass(pre[u % D], IC<u & 1>{});
                load(s + D, pre[u % D]);
                step_barrier();
            }
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
        // step s x-filters slice j = s - 1 into ring slot j % KZ; once the
        // ring holds KZ slices, the same step emits the z output of j.  KZ
        // steps per block keep every slot a compile-time register; blocks
        // whose steps all do both run without per-step checks.
        auto step = [&](auto ic, int s, auto checked) NDL_INLINE {
            constexpr int u = decltype(ic)::value;
            constexpr int U = (u + KZ - 1) % KZ; // slot of slice s - 1 (blocks start at multiples of KZ)
            constexpr bool CHK = decltype(checked)::value;
            const int j = s - 1;
            if (!CHK || (j >= 0 && j < nsl)) {
                // x-pass: y-plane columns [c2, c2 + KX + 1) hold x in [c2 - cx, c2 - cx + KX + 1)
                const float* yr = ycol + (j & 1) * PLANE;
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
                ring[0][U] = x0;
                ring[1][U] = x1;
                if (!CHK || j >= KZ - 1) {
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
            }
            step_barrier();
        };
        for (int s0 = 0; s0 < NSP; s0 += KZ) {
            if (s0 >= KZ && s0 + KZ - 1 <= nsl) {
                static_for<KZ>([&](auto ic) NDL_INLINE { step(ic, s0 + decltype(ic)::value, std::false_type{}); });
            } else {
                static_for<KZ>([&](auto ic) NDL_INLINE {
                    const int s = s0 + decltype(ic)::value;
                    if (s < NSP) step(ic, s, std::true_type{});
                });
            }
        }
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
