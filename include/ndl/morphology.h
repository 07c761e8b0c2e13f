// ndl/morphology.h -- drop-in erode()/dilate() and opening/closing on one
// MI355X.
//
// Same templates, defaults and semantics as the reference's
// processing/morphology.h:57-157: for every output the first included tap
// of the structuring element (kernel coordinate order, dimension 0 fastest;
// nonzero = included) and then every strictly smaller (erode) / larger
// (dilate) value, border resolved per tap, static_cast<DstT> of the winner.
// Bit-identical to the reference for every element type, NaN and signed
// zeros included.  An element whose included set is a product of per-axis
// sets (make_box_kernel, and any other product) runs as one min/max pass per
// axis (ndl_morph_separable); any other (make_cross_kernel, ...) runs the
// one-pass N-D kernel (ndl_morph_nd).  Host images are staged as in
// convolution.h; DeviceImage arguments stay in HBM.
#pragma once

#include <array>
#include <cassert>
#include <type_traits>
#include <vector>

#include "convolution.h"

namespace ndl {

namespace detail {

template <class SrcImageT, class DstImageT, class KernelT>
void morph(const SrcImageT& src, DstImageT& dst, const KernelT& kernel, BorderMode border, int32_t op)
{
    using SrcT = typename SrcImageT::value_type;
    using K = typename KernelT::value_type;
    static_assert(std::is_arithmetic_v<SrcT>, "ndl::erode()/dilate() require an arithmetic value_type");
    constexpr int DIM = dim_of<SrcImageT>;
    assert(dst.extent() == src.extent());
    auto kext = kernel.extent();
    std::vector<double> kd;
    kd.reserve(Image<K, DIM>::size(kext));
    for (const auto& c : kernel.coordinates()) kd.push_back(kernel.at(c) != K(0) ? 1.0 : 0.0);
    // per-axis inclusion: axis d includes c when any included tap has
    // coordinate c on d; the element is separable when it is their product
    std::array<std::vector<double>, DIM> axis;
    for (int d = 0; d < DIM; d++) axis[d].assign(kext[d], 0.0);
    std::size_t lin = 0;
    for (const auto& c : kernel.coordinates()) {
        if (kd[lin++] != 0.0)
            for (int d = 0; d < DIM; d++) axis[d][c[d]] = 1.0;
    }
    bool separable = true;
    lin = 0;
    for (const auto& c : kernel.coordinates()) {
        double p = 1.0;
        for (int d = 0; d < DIM; d++) p *= axis[d][c[d]];
        if (p != kd[lin++]) separable = false;
    }
    DeviceSource<SrcImageT> s(src);
    DeviceDest<DstImageT> o(dst);
    if (separable) {
        std::vector<int32_t> axes;
        std::vector<ndl_axis_kernel> k;
        for (int d = 0; d < DIM; d++)
            if (kext[d] > 1 || (d == DIM - 1 && axes.empty())) {
                axes.push_back(d);
                k.push_back(ndl_axis_kernel{axis[d].data(), (int32_t)kext[d], (int32_t)(kext[d] / 2), border_code(border)});
            }
        check(ndl_morph_separable(&s.t, &o.t, (int32_t)axes.size(), axes.data(), k.data(), op, nullptr));
    } else {
        std::array<int64_t, NDL_MAX_DIM> ke{};
        for (int d = 0; d < DIM; d++) ke[d] = kext[d];
        check(ndl_morph_nd(&s.t, &o.t, kd.data(), ke.data(), border_code(border), op, nullptr));
    }
    o.finish();
}

} // namespace detail

/// Morphological erosion: the minimum over the structuring element (morphology.h:57-79).
template <class SrcImageT, class DstImageT, class KernelT>
void erode(const SrcImageT& src, DstImageT& dst, const KernelT& kernel, BorderMode border = BorderMode::Clamp)
{
    detail::morph(src, dst, kernel, border, NDL_ERODE);
}

/// Morphological dilation: the maximum over the structuring element (morphology.h:80-101).
template <class SrcImageT, class DstImageT, class KernelT>
void dilate(const SrcImageT& src, DstImageT& dst, const KernelT& kernel, BorderMode border = BorderMode::Clamp)
{
    detail::morph(src, dst, kernel, border, NDL_DILATE);
}

/// erode() then dilate() through a caller-provided intermediate (morphology.h:133-148).
template <class SrcImageT, class DstImageT, class KernelT>
void morphological_open(const SrcImageT& src, DstImageT& dst, DstImageT& intermediate, const KernelT& kernel,
                        BorderMode border = BorderMode::Clamp)
{
    erode(src, intermediate, kernel, border);
    dilate(intermediate, dst, kernel, border);
}

/// dilate() then erode() through a caller-provided intermediate (morphology.h:149-157).
template <class SrcImageT, class DstImageT, class KernelT>
void morphological_close(const SrcImageT& src, DstImageT& dst, DstImageT& intermediate, const KernelT& kernel,
                         BorderMode border = BorderMode::Clamp)
{
    dilate(src, intermediate, kernel, border);
    erode(intermediate, dst, kernel, border);
}

} // namespace ndl
