// kernels_slab.hip -- instances of the whole-row fused kernel (slab_tile.hpp)
// and its launch: float32 volumes whose rows fit one tile (x extent <= 256).
#include "slab_tile.hpp"

namespace ndl {
namespace {
using namespace slab_detail;

constexpr int kSlabD = 4; // slices of input loads in flight per Y wave

// compile-time tap counts (x, y, z) with an instance: Gaussian sigma 0.5,
// 1, 1.5, 2 (3, 7, 11, 13 taps per axis), 9-tap kernels, and the
// sigma (1.5, 1.5, 3) volume filter (11/11/19)
template <class F> hipError_t with_slab(int kx, int ky, int kz, F&& f)
{
#define NDL_SLAB_INST(KX, KY, KZ) \
    if (kx == KX && ky == KY && kz == KZ) return f(IC<KX>{}, IC<KY>{}, IC<KZ>{});
    NDL_SLAB_INST(11, 11, 19)
    NDL_SLAB_INST(11, 11, 11)
    NDL_SLAB_INST(13, 13, 13)
    NDL_SLAB_INST(9, 9, 9)
    NDL_SLAB_INST(7, 7, 7)
    NDL_SLAB_INST(3, 3, 3)
#undef NDL_SLAB_INST
    return hipErrorInvalidValue;
}

} // namespace

bool slab_supported(int kx, int ky, int kz)
{
    return with_slab(kx, ky, kz, [](auto, auto, auto) { return hipSuccess; }) == hipSuccess;
}

int slab_blocks_per_cu(int kx, int ky, int kz)
{
    int n = 0;
    hipError_t e = with_slab(kx, ky, kz, [&](auto x, auto y, auto z) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n,
            reinterpret_cast<const void*>(
                &slab_kernel<decltype(x)::value, decltype(y)::value, decltype(z)::value, kSlabD>),
            kSlabThreads, 0);
    });
    return e == hipSuccess ? n : 0;
}

hipError_t launch_slab(const FusedArgs& a, hipStream_t s)
{
    if (a.ntiles == 0) return hipSuccess;
    return with_slab(a.kx, a.ky, a.kz, [&](auto x, auto y, auto z) {
        hipLaunchKernelGGL((slab_kernel<decltype(x)::value, decltype(y)::value, decltype(z)::value, kSlabD>),
                           dim3(a.ntiles), dim3(kSlabThreads), 0, s, a);
        return hipGetLastError();
    });
}

} // namespace ndl
