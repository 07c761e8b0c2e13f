// kernels_downsample.hip -- downsample() (processing/convolution.h:241-305)
// for blocks of at most 16 positions (factor 2 on up to four axes, 3 or 4 on
// up to two).
//
// One output per thread, as the generic kernel, but the block's positions
// (2, 3, 4, 8, 9 or 16, a template parameter) are unrolled: their offsets (relative to the block's first element) and
// their steps along each decimated axis come from the host, so every load of
// the block is issued before the first is summed.  A position past a ragged
// far edge (its step reaches the extent) is skipped, as the reference's
// clipped block never visits it.  The sum is FP64 in block order (dimension
// 0 fastest), divided by the block size and stored with std::round for
// integral types: the reference's arithmetic.
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <class T, int P> __global__ __launch_bounds__(256) void downsample_fast_kernel(const DownArgs a)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t ioff = 0, ooff = 0;
    int lim[4] = {1, 1, 1, 1}; // the block's extent along each decimated axis
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        const uint32_t q = a.fdiv[d].div(rem);
        const int64_t c = (int64_t)(rem - q * (uint32_t)a.oext[d]);
        rem = q;
        ooff += c * a.ost[d];
        const int64_t lo = d == a.channel_axis ? c : c * a.factor;
        ioff += lo * a.ist[d];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < a.nq && d == a.qdim[k]) lim[k] = (int)min((int64_t)a.factor, a.ext[d] - lo);
    }
    const int64_t count = (int64_t)lim[0] * lim[1] * lim[2] * lim[3];
    const T* in = (const T*)a.in + ioff;
    // a full block (the common case) needs no per-position check
    bool full = true;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (k < a.nq && lim[k] < a.factor) full = false;
    double v[P];
    bool ok[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        ok[p] = true;
        if (!full) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k < a.nq && a.pdig[p][k] >= lim[k]) ok[p] = false;
        }
        v[p] = ok[p] ? load_f64(in + a.poff[p]) : 0.0;
    }
    double sum = 0.0;
#pragma unroll
    for (int p = 0; p < P; p++)
        if (ok[p]) sum = sum + v[p];
    store_from_f64((T*)a.out + ooff, sum / (double)count, NDL_ROUND_NEAREST);
}

} // namespace

hipError_t launch_downsample_fast(const DownArgs& a, int dt, hipStream_t s)
{
    if (a.P < 1 || a.P > kDownPos || a.nq > 4) return hipErrorInvalidValue;
    if (a.total == 0) return hipSuccess;
    const uint32_t blocks = (a.total + 255u) / 256u;
    auto go = [&](auto t) {
        using T = decltype(t);
        switch (a.P) {
        case 2: hipLaunchKernelGGL((downsample_fast_kernel<T, 2>), dim3(blocks), dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL((downsample_fast_kernel<T, 3>), dim3(blocks), dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((downsample_fast_kernel<T, 4>), dim3(blocks), dim3(256), 0, s, a); break;
        case 8: hipLaunchKernelGGL((downsample_fast_kernel<T, 8>), dim3(blocks), dim3(256), 0, s, a); break;
        case 9: hipLaunchKernelGGL((downsample_fast_kernel<T, 9>), dim3(blocks), dim3(256), 0, s, a); break;
        case 16: hipLaunchKernelGGL((downsample_fast_kernel<T, 16>), dim3(blocks), dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    };
    switch (dt) {
    case NDL_U8: return go(uint8_t{});
    case NDL_I8: return go(int8_t{});
    case NDL_U16: return go(uint16_t{});
    case NDL_I16: return go(int16_t{});
    case NDL_I32: return go(int32_t{});
    case NDL_F32: return go(float{});
    case NDL_F64: return go(double{});
    case NDL_F16: return go(f16_t{});
    case NDL_BF16: return go(bf16_t{});
    }
    return hipErrorInvalidValue;
}

} // namespace ndl
