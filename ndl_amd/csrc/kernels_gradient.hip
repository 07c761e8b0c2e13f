// kernels_gradient.hip -- gradient() (processing/convolution.h:359-388): the
// central difference along every axis into a leading component axis, all
// components in one launch instead of one convolve() per axis.
//
// One thread per source element (dimension 0 fastest, so a wave reads
// neighbouring elements along each axis together).  Component a is exactly
// the reference's convolve() with the kernel {-0.5, 0, 0.5} along a
// (convolution.h:52-71, the zero tap skipped, detail.h:164-174): total = 0,
// total = total + double(x[B(i-1)]) * -0.5, total = total + double(x[B(i+1)])
// * 0.5, every product and sum rounded separately (this file is built with
// -ffp-contract=off), then the truncating cast into the output type; B is
// the border map of the call (single fold, mathHelpers.h:23-25).
#include "ndl_internal.hpp"

namespace ndl {
namespace {

template <class TIn, class TOut> __global__ __launch_bounds__(256) void gradient_kernel(const GradArgs a)
{
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.total) return;
    int64_t c[NDL_MAX_DIM];
    int64_t ioff = 0, ooff = 0;
    uint32_t rem = idx;
    for (int d = 0; d < a.nd; d++) {
        const uint32_t q = a.fdiv[d].div(rem);
        c[d] = (int64_t)(rem - q * (uint32_t)a.ext[d]);
        rem = q;
        ioff += c[d] * a.ist[d];
        ooff += c[d] * a.ost[d];
    }
    const TIn* in = (const TIn*)a.in + ioff;
    TOut* out = (TOut*)a.out + ooff;
    for (int d = 0; d < a.nd; d++) {
        const int64_t m = a.ext[d];
        int64_t lo = c[d] - 1, hi = c[d] + 1;
        if (lo < 0) lo = border_map(a.border, m, lo);
        if (hi >= m) hi = border_map(a.border, m, hi);
        double total = 0.0;
        total = total + load_f64(in + (lo - c[d]) * a.ist[d]) * -0.5;
        total = total + load_f64(in + (hi - c[d]) * a.ist[d]) * 0.5;
        store_from_f64(out + d * a.cst, total, NDL_ROUND_TRUNC);
    }
}

template <class F> hipError_t with_grad_dtype(int dt, F&& f)
{
    switch (dt) {
    case NDL_U8: return f(uint8_t{});
    case NDL_I8: return f(int8_t{});
    case NDL_U16: return f(uint16_t{});
    case NDL_I16: return f(int16_t{});
    case NDL_I32: return f(int32_t{});
    case NDL_F32: return f(float{});
    case NDL_F64: return f(double{});
    case NDL_F16: return f(f16_t{});
    case NDL_BF16: return f(bf16_t{});
    }
    return hipErrorInvalidValue;
}

} // namespace

hipError_t launch_gradient(const GradArgs& a, int in_dt, int out_dt, hipStream_t s)
{
    if (a.total == 0) return hipSuccess;
    const uint32_t blocks = (a.total + 255u) / 256u;
    return with_grad_dtype(in_dt, [&](auto tin) {
        return with_grad_dtype(out_dt, [&](auto tout) {
            hipLaunchKernelGGL((gradient_kernel<decltype(tin), decltype(tout)>), dim3(blocks), dim3(256), 0, s, a);
            return hipGetLastError();
        });
    });
}

} // namespace ndl
