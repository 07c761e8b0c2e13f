// kernels_fused.hip -- forward instances of the fused separable kernel
// (fused_tile.hpp) and the launch / occupancy dispatch.
#include <algorithm>
#include <cstdlib>

#include "fused_tile.hpp"

namespace ndl {
namespace {
using namespace fused_detail;

__global__ void reset_counter_kernel(uint32_t* c) { *c = 0u; }

// ------------------------------------------------------------- dispatch
constexpr int kTX = 64, kTY = 16;
// row march (TY = 1): one wave of 2-output lanes per 128-column tile
constexpr int kRowTX = 128;
template <int TY> constexpr int tile_x() { return TY == 1 ? kRowTX : kTX; }

// Instantiation table.  Specialised (compile-time taps) filters first, then
// generic ring periods per storage type.  f(ts, cpt, kx, ky, ring, il, ty)
// launches.
template <class F> hipError_t with_instance(const FusedConfig& cfg, int dt, int kx, int ky, F&& f)
{
    const int cpt = cfg.cpt, kz = cfg.ring, il = cfg.il;
    if (cfg.ty == 1) {
        if (cfg.tx != kRowTX || cpt != 2 || ky != 1) return hipErrorInvalidValue;
        // uint8 7-tap box on x and y (planar and 3-channel interleaved rows)
        if (dt == NDL_U8 && kx == 7 && kz == 7 && (il == 1 || il == 3)) {
            if (il == 1) return f(uint8_t{}, IC<2>{}, IC<7>{}, IC<1>{}, IC<7>{}, IC<1>{}, IC<1>{});
            return f(uint8_t{}, IC<2>{}, IC<7>{}, IC<1>{}, IC<7>{}, IC<3>{}, IC<1>{});
        }
        // float images: run-time x taps, the y taps are the ring
        if (il != 1 || dt != NDL_F32) return hipErrorInvalidValue;
#define NDL_ROW(KZ) \
    if (kz == KZ) return f(float{}, IC<2>{}, IC<0>{}, IC<1>{}, IC<KZ>{}, IC<1>{}, IC<1>{});
        NDL_ROW(3) NDL_ROW(5) NDL_ROW(7) NDL_ROW(9) NDL_ROW(11) NDL_ROW(13) NDL_ROW(15) NDL_ROW(17) NDL_ROW(19)
        NDL_ROW(21) NDL_ROW(23) NDL_ROW(25)
#undef NDL_ROW
        return hipErrorInvalidValue;
    }
    if (cfg.tx != kTX) return hipErrorInvalidValue;
    if (cfg.ty == 8) {
        // 64 x 8 tiles: twice the tiles per slice, so longer z chunks fill the GPU
        if (dt == NDL_F32 && kx == 11 && ky == 11 && kz == 19 && il == 1) {
            if (cpt == 2) return f(float{}, IC<2>{}, IC<11>{}, IC<11>{}, IC<19>{}, IC<1>{}, IC<8>{});
            if (cpt == 1) return f(float{}, IC<1>{}, IC<11>{}, IC<11>{}, IC<19>{}, IC<1>{}, IC<8>{});
        }
        return hipErrorInvalidValue;
    }
    if (cfg.ty != kTY) return hipErrorInvalidValue;
    // NDL_FUSED_GENERIC=1: run-time-tap instances only (measurement knob)
    static const bool generic_only = getenv("NDL_FUSED_GENERIC") && atoi(getenv("NDL_FUSED_GENERIC")) == 1;
    if (generic_only) goto generic;
#define NDL_SPEC_IL(DT, T, C, KX, KY, KZ, IL)                                                 \
    if (dt == DT && cpt == C && kx == KX && ky == KY && kz == KZ && il == IL)                 \
        return f(T{}, IC<C>{}, IC<KX>{}, IC<KY>{}, IC<KZ>{}, IC<IL>{}, IC<kTY>{});
#define NDL_SPEC(DT, T, C, KX, KY, KZ) NDL_SPEC_IL(DT, T, C, KX, KY, KZ, 1)
    // 3-D Gaussian sigma (1.5, 1.5, 3.0): 11/11/19 taps
    NDL_SPEC(NDL_F32, float, 2, 11, 11, 19) NDL_SPEC(NDL_F32, float, 4, 11, 11, 19)
    // 2-D Gaussian sigma 2: 13 taps
    NDL_SPEC(NDL_F32, float, 2, 13, 13, 1) NDL_SPEC(NDL_F32, float, 4, 13, 13, 1)
    // video: sigma 1 (7 taps) on x, y and a 5-tap temporal binomial
    NDL_SPEC(NDL_F32, float, 2, 7, 7, 5) NDL_SPEC(NDL_F32, float, 4, 7, 7, 5)
    // 9-tap kernels on every axis, float and half-precision storage; the
    // two-axis passes of their tap gradients (x through the other axes)
    NDL_SPEC(NDL_F32, float, 2, 9, 9, 9) NDL_SPEC(NDL_BF16, bf16_t, 2, 9, 9, 9) NDL_SPEC(NDL_F16, f16_t, 2, 9, 9, 9)
    NDL_SPEC(NDL_F32, float, 2, 9, 9, 1) NDL_SPEC(NDL_F32, float, 2, 1, 9, 9)
    // uint8 7-tap box on x and y, planar and 3-channel interleaved
    NDL_SPEC(NDL_U8, uint8_t, 2, 7, 7, 1) NDL_SPEC(NDL_U8, uint8_t, 4, 7, 7, 1)
    NDL_SPEC_IL(NDL_U8, uint8_t, 2, 7, 7, 1, 3)
#undef NDL_SPEC
#undef NDL_SPEC_IL
generic:
    if (cpt != 2 || il != 1) return hipErrorInvalidValue;
#define NDL_GEN(DT, T, KZ) \
    if (dt == DT && kz == KZ) return f(T{}, IC<2>{}, IC<0>{}, IC<0>{}, IC<KZ>{}, IC<1>{}, IC<kTY>{});
    NDL_GEN(NDL_F32, float, 1) NDL_GEN(NDL_F32, float, 3) NDL_GEN(NDL_F32, float, 5) NDL_GEN(NDL_F32, float, 7)
    NDL_GEN(NDL_F32, float, 9) NDL_GEN(NDL_F32, float, 11) NDL_GEN(NDL_F32, float, 13) NDL_GEN(NDL_F32, float, 15)
    NDL_GEN(NDL_F32, float, 17) NDL_GEN(NDL_F32, float, 19) NDL_GEN(NDL_F32, float, 21) NDL_GEN(NDL_F32, float, 23)
    NDL_GEN(NDL_F32, float, 25)
    NDL_GEN(NDL_F16, f16_t, 1) NDL_GEN(NDL_F16, f16_t, 3) NDL_GEN(NDL_F16, f16_t, 5) NDL_GEN(NDL_F16, f16_t, 7)
    NDL_GEN(NDL_F16, f16_t, 9) NDL_GEN(NDL_F16, f16_t, 11) NDL_GEN(NDL_F16, f16_t, 13)
    NDL_GEN(NDL_BF16, bf16_t, 1) NDL_GEN(NDL_BF16, bf16_t, 3) NDL_GEN(NDL_BF16, bf16_t, 5)
    NDL_GEN(NDL_BF16, bf16_t, 7) NDL_GEN(NDL_BF16, bf16_t, 9) NDL_GEN(NDL_BF16, bf16_t, 11)
    NDL_GEN(NDL_BF16, bf16_t, 13)
    NDL_GEN(NDL_U8, uint8_t, 1) NDL_GEN(NDL_U8, uint8_t, 3) NDL_GEN(NDL_U8, uint8_t, 5) NDL_GEN(NDL_U8, uint8_t, 7)
#undef NDL_GEN
    return hipErrorInvalidValue;
}

// queue kernels: float storage, generic taps
template <class F> hipError_t with_queue_instance(const FusedConfig& cfg, int dt, F&& f)
{
    if (cfg.cpt != 2 || cfg.il != 1 || dt != NDL_F32) return hipErrorInvalidValue;
    if (cfg.ty == 1) {
        if (cfg.tx != kRowTX) return hipErrorInvalidValue;
        switch (cfg.ring) {
#define NDL_QR(R) case R: return f(IC<R>{}, IC<1>{});
            NDL_QR(3) NDL_QR(5) NDL_QR(7) NDL_QR(9) NDL_QR(11) NDL_QR(13) NDL_QR(15) NDL_QR(17) NDL_QR(19) NDL_QR(21)
            NDL_QR(23) NDL_QR(25)
#undef NDL_QR
        }
        return hipErrorInvalidValue;
    }
    if (cfg.tx != kTX || cfg.ty != kTY) return hipErrorInvalidValue;
    switch (cfg.ring) {
#define NDL_Q(R) case R: return f(IC<R>{}, IC<kTY>{});
        NDL_Q(1) NDL_Q(3) NDL_Q(5) NDL_Q(7) NDL_Q(9) NDL_Q(11) NDL_Q(13) NDL_Q(15) NDL_Q(17) NDL_Q(19) NDL_Q(21)
        NDL_Q(23) NDL_Q(25)
#undef NDL_Q
    }
    return hipErrorInvalidValue;
}

} // namespace

int fused_ring_for(int kz)
{
    return (kz >= 1 && kz <= 25 && (kz & 1)) ? kz : -1;
}

size_t fused_lds_bytes(const FusedArgs& a, const FusedConfig& cfg, bool)
{
    const int NC = cfg.tx + (a.kx - 1) * cfg.il;
    const int NR = cfg.ty + a.ky - 1;
    const int NCP = (NC + 3 + 3) & ~3;
    // the compile-time-tap instances pad the window to whole load slots
    // (NLMAX * NT floats); a bound that covers every instance:
    const int NT = cfg.tx * cfg.ty / cfg.cpt;
    const size_t in_fl = std::max<size_t>((((NR + 3) * NC + 3) & ~3), (size_t)((NR * NC + NT - 1) / NT) * NT);
    size_t floats = (size_t)2 * (in_fl + cfg.ty * NCP);
    // edge instances: weight rows of x, y and z (at most K rows per end)
    if (cfg.edge) floats += (size_t)2 * (a.kx * a.kx + a.ky * a.ky + a.kz * a.kz);
    floats += 64; // run-time x / y taps
    return floats * 4;
}

bool fused_supported(const FusedConfig& cfg, int dt, int kx, int ky, bool queue)
{
    if (cfg.grad) return !queue && fused_grad_supported(cfg, dt, kx, ky);
    if (cfg.edge) return !queue && fused_edge_supported(cfg, dt, kx, ky);
    if (queue) return with_queue_instance(cfg, dt, [](auto, auto) { return hipSuccess; }) == hipSuccess;
    return with_instance(cfg, dt, kx, ky, [](auto, auto, auto, auto, auto, auto, auto) { return hipSuccess; }) ==
           hipSuccess;
}

hipError_t launch_fused(const FusedArgs& a, const FusedConfig& cfg, int dt, bool, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    if (cfg.ring != a.kz) return hipErrorInvalidValue;
    const size_t lds = fused_lds_bytes(a, cfg, false);
    if (cfg.grad) return launch_fused_grad(a, cfg, dt, lds, s);
    if (cfg.edge) return launch_fused_edge(a, cfg, dt, lds, s);
    return with_instance(cfg, dt, a.kx, a.ky, [&](auto ts, auto c, auto kxs, auto kys, auto r, auto il, auto ty) {
        using TS = decltype(ts);
        constexpr int TY = decltype(ty)::value, CPT = decltype(c)::value, TX = tile_x<TY>(), NT = TX * TY / CPT;
        hipLaunchKernelGGL((fused_tile_kernel<TS, TX, TY, CPT, decltype(kxs)::value, decltype(kys)::value,
                                              decltype(r)::value, decltype(il)::value>),
                           dim3(a.ntiles), dim3(NT), lds, s, a);
        return hipGetLastError();
    });
}

hipError_t launch_fused_queue(const FusedQueue& q, const FusedConfig& cfg, int dt, bool, size_t lds, hipStream_t s)
{
    if (q.total == 0) return hipSuccess;
    if (cfg.edge) return hipErrorInvalidValue;
    if (!fused_supported(cfg, dt, 0, 0, true)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, q.counter);
    return with_queue_instance(cfg, dt, [&](auto r, auto ty) {
        constexpr int TY = decltype(ty)::value, TX = tile_x<TY>(), NT = TX * TY / 2;
        hipLaunchKernelGGL((fused_queue_kernel<float, TX, TY, 2, decltype(r)::value>), dim3(q.grid), dim3(NT), lds,
                           s, q.jobs, q.tile_start, q.njobs, q.total, q.counter);
        return hipGetLastError();
    });
}

int fused_blocks_per_cu(const FusedConfig& cfg, int dt, int kx, int ky, size_t lds)
{
    if (cfg.grad) return fused_grad_blocks_per_cu(cfg, dt, kx, ky, lds);
    if (cfg.edge) return fused_edge_blocks_per_cu(cfg, dt, kx, ky, lds);
    int n = 0;
    hipError_t e = with_instance(cfg, dt, kx, ky, [&](auto ts, auto c, auto kxs, auto kys, auto r, auto il, auto ty) {
        using TS = decltype(ts);
        constexpr int TY = decltype(ty)::value, CPT = decltype(c)::value, TX = tile_x<TY>(), NT = TX * TY / CPT;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n,
            reinterpret_cast<const void*>(&fused_tile_kernel<TS, TX, TY, CPT, decltype(kxs)::value,
                                                             decltype(kys)::value, decltype(r)::value,
                                                             decltype(il)::value>),
            NT, lds);
    });
    return e == hipSuccess ? n : 0;
}

} // namespace ndl
