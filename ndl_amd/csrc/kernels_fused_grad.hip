// kernels_fused_grad.hip -- tap-gradient instances of the fused tile kernel
// (fused_tile.hpp, GRAD = true): x through the x/y taps is built in the tile
// and correlated with grad_out along the z role without leaving the chip.
// Used by ndl_sepconv_adjoint_taps for each axis that can take the z role
// (any axis except the unit-stride one), float32 images, FP32 accumulation
// of the passes and FP64 accumulation of the sums.
#include "fused_tile.hpp"

namespace ndl {
namespace {
using namespace fused_detail;

// f(kxs, kys, ring): compile-time x/y taps for the 9-tap filters, run-time
// taps for every odd ring period
template <class F> hipError_t with_grad_instance(const FusedConfig& cfg, int dt, int kx, int ky, F&& f)
{
    if (!cfg.grad || cfg.tx != 64 || cfg.ty != 16 || cfg.cpt != 2 || cfg.il != 1 || cfg.edge || dt != NDL_F32)
        return hipErrorInvalidValue;
    if (kx == 9 && ky == 9 && cfg.ring == 9) return f(IC<9>{}, IC<9>{}, IC<9>{});
    switch (cfg.ring) {
#define NDL_GR(R) case R: return f(IC<0>{}, IC<0>{}, IC<R>{});
        NDL_GR(3) NDL_GR(5) NDL_GR(7) NDL_GR(9) NDL_GR(11) NDL_GR(13) NDL_GR(15) NDL_GR(17) NDL_GR(19) NDL_GR(21)
        NDL_GR(23) NDL_GR(25)
#undef NDL_GR
    }
    return hipErrorInvalidValue;
}

} // namespace

bool fused_grad_supported(const FusedConfig& cfg, int dt, int kx, int ky)
{
    return with_grad_instance(cfg, dt, kx, ky, [](auto, auto, auto) { return hipSuccess; }) == hipSuccess;
}

hipError_t launch_fused_grad(const FusedArgs& a, const FusedConfig& cfg, int dt, size_t lds, hipStream_t s)
{
    if (!a.gpart) return hipErrorInvalidValue;
    return with_grad_instance(cfg, dt, a.kx, a.ky, [&](auto kxs, auto kys, auto r) {
        hipLaunchKernelGGL((fused_tile_kernel<float, 64, 16, 2, decltype(kxs)::value, decltype(kys)::value,
                                              decltype(r)::value, 1, false, true>),
                           dim3(a.ntiles), dim3(512), lds, s, a);
        return hipGetLastError();
    });
}

int fused_grad_blocks_per_cu(const FusedConfig& cfg, int dt, int kx, int ky, size_t lds)
{
    int n = 0;
    hipError_t e = with_grad_instance(cfg, dt, kx, ky, [&](auto kxs, auto kys, auto r) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n,
            reinterpret_cast<const void*>(&fused_tile_kernel<float, 64, 16, 2, decltype(kxs)::value,
                                                             decltype(kys)::value, decltype(r)::value, 1, false,
                                                             true>),
            512, lds);
    });
    return e == hipSuccess ? n : 0;
}

} // namespace ndl
