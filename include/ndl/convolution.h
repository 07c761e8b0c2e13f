// ndl/convolution.h -- drop-in convolve()/gaussian_blur() on one MI355X.
//
// Same templates, defaults and semantics as the reference's
// processing/convolution.h:52-71 (convolve) and :107-186 (gaussian_blur),
// computed by libndl_gpu.so through the C ABI in ndl_gpu.h:
//  * convolve(src, dst, kernel, border): correlation, kernel centre K/2,
//    zero taps skipped, FP64 accumulation, static_cast into DstT.  A kernel
//    with extent > 1 on a single axis runs as one separable pass; any other
//    kernel runs the dense N-D kernel.  Bit-identical to the reference.
//  * gaussian_blur(src, dst, sigma, border): taps of convolution.h:116-126,
//    DIM passes in axis order through FP64 intermediates, std::round into an
//    integral DstT when DIM >= 2.  Bit-identical to the reference.
// Additions (no reference counterpart): per-axis sigma/taps/borders,
// FP32 accumulation (Accumulation::FP32, float/f16/bf16 data), the adjoints
// convolve_adjoint()/convolve_adjoint_taps(), and DeviceImage overloads that
// stay in HBM.
//
// Host images are staged: the source's memory span is copied to the device
// as is (any strides, incl. negative), the result comes back densely and is
// written through dst's own strides, so elements outside dst's view are
// never touched.  Errors from the library throw std::runtime_error; the
// reference's preconditions (equal extents, sigma > 0) are also asserted.
#pragma once

#include <array>
#include <cassert>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "image.h"

namespace ndl {

/// Accumulation contract of a call (see NDL_ACCUM_* in ndl_gpu.h).
enum class Accumulation { Reference = NDL_ACCUM_REF, FP32 = NDL_ACCUM_F32 };

namespace detail {

template <class ImageT> struct ImageDim;
template <class T, int DIM> struct ImageDim<Image<T, DIM>> { static constexpr int value = DIM; };
template <class T, int DIM> struct ImageDim<OwnedImage<T, DIM>> { static constexpr int value = DIM; };
template <class T, int DIM> struct ImageDim<DeviceImage<T, DIM>> { static constexpr int value = DIM; };
template <class ImageT> constexpr int dim_of = std::tuple_size<std::decay_t<decltype(std::declval<const ImageT&>().extent())>>::value;

template <class ImageT> struct is_device : std::false_type {};
template <class T, int DIM> struct is_device<DeviceImage<T, DIM>> : std::true_type {};

inline int32_t border_code(BorderMode b)
{
    return b == BorderMode::Wrap ? NDL_WRAP : (b == BorderMode::Reflect ? NDL_REFLECT : NDL_CLAMP);
}

template <class T, int DIM> ndl_tensor host_tensor(const std::array<int, DIM>& e, const std::array<int, DIM>& st)
{
    ndl_tensor t{};
    t.dtype = dtype_of<T>();
    t.ndim = DIM;
    for (int i = 0; i < DIM; i++) {
        t.extent[i] = e[i];
        t.stride[i] = st[i];
    }
    return t;
}

// A source on the device: DeviceImage as is, host views staged by span.
template <class ImageT> struct DeviceSource {
    using T = typename ImageT::value_type;
    static constexpr int DIM = dim_of<ImageT>;
    ndl_tensor t{};
    std::vector<DeviceBuffer> keep;
    explicit DeviceSource(const ImageT& img)
    {
        if constexpr (is_device<ImageT>::value) {
            t = img.tensor();
        } else {
            static_assert(has_stride_access_v<ImageT, DIM>,
                          "the GPU path needs an image with linear strides (stride() and a T& at()): Image, "
                          "OwnedImage, their views, DeviceImage, or a type of your own with that interface; convert "
                          "bit-packed or ring-buffer images first");
            static_assert(gpu_type_v<T>, "GPU convolution supports u8/i8/u16/i16/i32/float/double images");
            t = host_tensor<T, DIM>(img.extent(), img.stride());
            std::ptrdiff_t lo = 0, hi = 0;
            bool empty = false;
            for (int d = 0; d < DIM; d++) {
                if (img.extent()[d] == 0) empty = true;
                std::ptrdiff_t r = (std::ptrdiff_t)(img.extent()[d] - 1) * img.stride()[d];
                if (r < 0) lo += r; else hi += r;
            }
            if (empty) return;
            // the element at index 0 of every dimension, through the type's own at()
            const T* origin = &const_cast<ImageT&>(img).at(std::array<int, DIM>{});
            std::size_t bytes = (std::size_t)(hi - lo + 1) * sizeof(T);
            keep.emplace_back(bytes);
            check(ndl_memcpy_htod(keep.back().p, origin + lo, bytes, nullptr));
            t.data = (T*)keep.back().p - lo;
        }
    }
};

// A destination on the device; finish() brings a host view up to date.
template <class ImageT> struct DeviceDest {
    using T = typename ImageT::value_type;
    static constexpr int DIM = dim_of<ImageT>;
    ndl_tensor t{};
    std::vector<DeviceBuffer> keep;
    ImageT* host = nullptr;
    explicit DeviceDest(ImageT& img)
    {
        if constexpr (is_device<ImageT>::value) {
            t = img.tensor();
        } else {
            static_assert(has_stride_access_v<ImageT, DIM>,
                          "the GPU path needs a destination with linear strides (stride() and a T& at()): Image, "
                          "OwnedImage, their views, DeviceImage, or a type of your own with that interface");
            static_assert(gpu_type_v<T>, "GPU convolution supports u8/i8/u16/i16/i32/float/double images");
            host = &img;
            std::array<int, DIM> e = img.extent();
            Image<T, DIM> dense(nullptr, e);
            t = dense.tensor(dtype_of<T>());
            std::size_t n = Image<T, DIM>::size(e);
            keep.emplace_back(std::max<std::size_t>(n, 1) * sizeof(T));
            t.data = keep.back().p;
        }
    }
    void finish()
    {
        if constexpr (!is_device<ImageT>::value) {
            check(ndl_stream_synchronize(nullptr));
            std::size_t n = Image<T, DIM>::size(host->extent());
            if (n == 0) return;
            if constexpr (std::is_base_of_v<Image<T, DIM>, ImageT>) {
                if (host->stride() == default_stride<DIM>(host->extent())) {
                    check(ndl_memcpy_dtoh(host->data(), t.data, n * sizeof(T), nullptr));
                    check(ndl_stream_synchronize(nullptr));
                    return;
                }
            }
            std::vector<T> tmp(n);
            check(ndl_memcpy_dtoh(tmp.data(), t.data, n * sizeof(T), nullptr));
            check(ndl_stream_synchronize(nullptr));
            // through the destination's own at(): elements outside a view stay untouched
            Image<T, DIM> dense(tmp.data(), host->extent());
            std::array<int, DIM> c{};
            for (std::size_t i = 0; i < n; i++) {
                host->at(c) = dense.at(c);
                for (int d = 0; d < DIM; d++) {
                    if (++c[d] < host->extent()[d]) break;
                    c[d] = 0;
                }
            }
        } else {
            check(ndl_stream_synchronize(nullptr));
        }
    }
};

template <int DIM> struct AxisPlan {
    std::vector<std::vector<double>> taps;
    std::vector<ndl_axis_kernel> k;
    std::vector<int32_t> axes;
    void add(int axis, std::vector<double> t, int center, BorderMode b)
    {
        taps.push_back(std::move(t));
        axes.push_back(axis);
        k.push_back(ndl_axis_kernel{nullptr, (int32_t)taps.back().size(), (int32_t)center, border_code(b)});
    }
    void bind()
    {
        for (std::size_t i = 0; i < k.size(); i++) k[i].taps = taps[i].data();
    }
};

template <class Src, class Dst, int DIM>
void run_sepconv(const Src& src, Dst& dst, AxisPlan<DIM>& plan, Accumulation acc, int32_t rounding)
{
    assert(src.extent() == dst.extent());
    if (plan.axes.empty()) {
        // nothing to convolve: a cast copy through a 1-tap identity pass
        plan.add(0, {1.0}, 0, BorderMode::Clamp);
    }
    plan.bind();
    DeviceSource<Src> s(src);
    DeviceDest<Dst> d(dst);
    check(ndl_sepconv_forward(&s.t, &d.t, (int32_t)plan.axes.size(), plan.axes.data(), plan.k.data(),
                              (int32_t)acc, rounding, nullptr));
    d.finish();
}

inline std::vector<double> gaussian_taps(double sigma)
{
    std::vector<double> t(NDL_MAX_TAPS * 8);
    int n = ndl_gaussian_taps(sigma, t.data(), (int32_t)t.size());
    check(n);
    t.resize(n);
    return t;
}

} // namespace detail

/// Correlates `kernel` against `src` into `dst` (processing/convolution.h:52-71).
template <class SrcImageT, class DstImageT, class KernelT>
void convolve(const SrcImageT& src, DstImageT& dst, const KernelT& kernel, BorderMode border = BorderMode::Clamp)
{
    using SrcT = typename SrcImageT::value_type;
    using DstT = typename DstImageT::value_type;
    using K = typename KernelT::value_type;
    static_assert(std::is_arithmetic_v<SrcT>, "ndl::convolve() requires an arithmetic value_type");
    static_assert(std::is_arithmetic_v<DstT>, "ndl::convolve() requires an arithmetic value_type");
    constexpr int DIM = detail::dim_of<SrcImageT>;
    assert(dst.extent() == src.extent());
    auto kext = kernel.extent();
    // dense copy of the kernel, dimension 0 fastest
    std::vector<double> kd;
    kd.reserve(Image<K, DIM>::size(kext));
    for (const auto& c : kernel.coordinates()) kd.push_back(static_cast<double>(kernel.at(c)));
    int wide = -1, nwide = 0;
    for (int d = 0; d < DIM; d++)
        if (kext[d] > 1) { wide = d; nwide++; }
    if (nwide <= 1) {
        detail::AxisPlan<DIM> plan;
        int ax = wide < 0 ? 0 : wide;
        plan.add(ax, kd, (int)(kd.size() / 2), border);
        detail::run_sepconv(src, dst, plan, Accumulation::Reference, NDL_ROUND_TRUNC);
        return;
    }
    std::array<int64_t, NDL_MAX_DIM> ke{};
    for (int d = 0; d < DIM; d++) ke[d] = kext[d];
    detail::DeviceSource<SrcImageT> s(src);
    detail::DeviceDest<DstImageT> o(dst);
    detail::check(ndl_convolve_nd(&s.t, &o.t, kd.data(), ke.data(), detail::border_code(border), nullptr));
    o.finish();
}

/// Normalised Gaussian blur, one sigma and one border on every axis
/// (processing/convolution.h:107-186).
template <class SrcImageT, class DstImageT>
void gaussian_blur(const SrcImageT& src, DstImageT& dst, double sigma, BorderMode border = BorderMode::Clamp)
{
    using SrcT = typename SrcImageT::value_type;
    using DstT = typename DstImageT::value_type;
    static_assert(std::is_arithmetic_v<SrcT>, "ndl::gaussian_blur() requires an arithmetic value_type");
    constexpr int DIM = detail::dim_of<SrcImageT>;
    assert(sigma > 0);
    auto taps = detail::gaussian_taps(sigma);
    detail::AxisPlan<DIM> plan;
    for (int d = 0; d < DIM; d++) plan.add(d, taps, (int)taps.size() / 2, border);
    int32_t rounding = (DIM >= 2 && std::is_integral_v<DstT>) ? NDL_ROUND_NEAREST : NDL_ROUND_TRUNC;
    detail::run_sepconv(src, dst, plan, Accumulation::Reference, rounding);
}

/// Per-axis Gaussian: sigma[d] and border[d] on axis d (sigma <= 0 leaves the
/// axis untouched), passes in axis order.  Accumulation::Reference reproduces
/// the reference's gaussian_blur arithmetic chain exactly.
template <class SrcImageT, class DstImageT, int DIM>
void gaussian_blur(const SrcImageT& src, DstImageT& dst, const std::array<double, DIM>& sigma,
                   const std::array<BorderMode, DIM>& border, Accumulation acc = Accumulation::Reference)
{
    using DstT = typename DstImageT::value_type;
    static_assert(DIM == detail::dim_of<SrcImageT>, "one sigma per dimension");
    detail::AxisPlan<DIM> plan;
    for (int d = 0; d < DIM; d++) {
        if (!(sigma[d] > 0)) continue;
        auto taps = detail::gaussian_taps(sigma[d]);
        plan.add(d, taps, (int)taps.size() / 2, border[d]);
    }
    int32_t rounding = (plan.axes.size() >= 2 && std::is_integral_v<DstT>) ? NDL_ROUND_NEAREST : NDL_ROUND_TRUNC;
    detail::run_sepconv(src, dst, plan, acc, rounding);
}

/// Separable correlation with per-axis taps (an empty vector leaves the axis
/// untouched), centre count/2, passes in axis order through FP64 (Reference)
/// or FP32 intermediates; the final cast truncates like convolve().
template <class SrcImageT, class DstImageT, int DIM>
void convolve_separable(const SrcImageT& src, DstImageT& dst, const std::array<std::vector<double>, DIM>& taps,
                        const std::array<BorderMode, DIM>& border, Accumulation acc = Accumulation::Reference)
{
    static_assert(DIM == detail::dim_of<SrcImageT>, "one tap list per dimension");
    detail::AxisPlan<DIM> plan;
    for (int d = 0; d < DIM; d++)
        if (!taps[d].empty()) plan.add(d, taps[d], (int)taps[d].size() / 2, border[d]);
    detail::run_sepconv(src, dst, plan, acc, NDL_ROUND_TRUNC);
}

/// gaussian_blur() of one image at several sigmas: levels[l] =
/// gaussian_blur(src, sigma[l], border), bit-identical to separate calls
/// (the scale stack detect_keypoints builds, feature_detection.h:101-109);
/// the first axis pass of every level reads src once.
template <class SrcImageT, class LevelImageT>
void gaussian_blur_scales(const SrcImageT& src, std::vector<LevelImageT>& levels, const std::vector<double>& sigma,
                          BorderMode border = BorderMode::Clamp)
{
    using DstT = typename LevelImageT::value_type;
    constexpr int DIM = detail::dim_of<SrcImageT>;
    assert(levels.size() == sigma.size());
    for (double sg : sigma) assert(sg > 0);
    (void)sizeof(DstT);
    if (levels.empty()) return;
    detail::DeviceSource<SrcImageT> s(src);
    std::vector<std::unique_ptr<detail::DeviceDest<LevelImageT>>> o;
    std::vector<ndl_tensor> t;
    for (auto& lv : levels) {
        assert(lv.extent() == src.extent());
        o.emplace_back(new detail::DeviceDest<LevelImageT>(lv));
        t.push_back(o.back()->t);
    }
    const int32_t rounding = std::is_integral_v<DstT> && DIM >= 2 ? NDL_ROUND_NEAREST : NDL_ROUND_TRUNC;
    detail::check(ndl_gaussian_scales(&s.t, t.data(), (int32_t)t.size(), sigma.data(), detail::border_code(border),
                                      NDL_ACCUM_REF, rounding, nullptr));
    for (auto& d : o) d->finish();
}

/// Difference-of-Gaussians layers dogs[l] = levels[l+1] - levels[l]
/// (feature_detection.h:111-120), in FP64, stored as DogImageT's type.
template <class LevelImageT, class DogImageT>
void difference_of_gaussians(const std::vector<LevelImageT>& levels, std::vector<DogImageT>& dogs)
{
    assert(levels.size() >= 2 && dogs.size() + 1 == levels.size());
    std::vector<std::unique_ptr<detail::DeviceSource<LevelImageT>>> in;
    std::vector<std::unique_ptr<detail::DeviceDest<DogImageT>>> out;
    std::vector<ndl_tensor> lt, dt;
    for (const auto& lv : levels) {
        in.emplace_back(new detail::DeviceSource<LevelImageT>(lv));
        lt.push_back(in.back()->t);
    }
    for (auto& d : dogs) {
        out.emplace_back(new detail::DeviceDest<DogImageT>(d));
        dt.push_back(out.back()->t);
    }
    detail::check(ndl_dog(lt.data(), dt.data(), (int32_t)lt.size(), nullptr));
    for (auto& d : out) d->finish();
}

/// Block-mean shrink by `factor` along every axis but `channelAxis`
/// (convolution.h:241-305); returns a new OwnedImage.  `border` is accepted
/// for interface stability and, as in the reference, unused.
template <class ImageT>
auto downsample(const ImageT& src, int factor, int channelAxis = 0, BorderMode border = BorderMode::Clamp)
{
    (void)border;
    using T = typename ImageT::value_type;
    static_assert(std::is_arithmetic_v<T>, "ndl::downsample() requires an arithmetic value_type");
    constexpr int DIM = detail::dim_of<ImageT>;
    assert(factor > 1);
    auto e = src.extent();
    std::array<int, DIM> oe;
    for (int i = 0; i < DIM; i++) oe[i] = i == channelAxis ? e[i] : (e[i] - 1) / factor + 1;
    OwnedImage<T, DIM> result(oe);
    detail::DeviceSource<ImageT> s(src);
    detail::DeviceDest<OwnedImage<T, DIM>> o(result);
    detail::check(ndl_downsample(&s.t, &o.t, factor, channelAxis, nullptr));
    o.finish();
    return result;
}

/// downsample(gaussian_blur(src, sigma, border), factor, channelAxis) with the
/// blurred image in src's type (convolution.h:107-186 then :241-305): one
/// step of a Gaussian pyramid, bit-identical to the two calls.  The last
/// blur pass, its rounding and the block mean run in one kernel, so the
/// full-resolution blurred image is never stored.  Returns a new OwnedImage.
template <class ImageT>
auto gaussian_downsample(const ImageT& src, double sigma, int factor, int channelAxis = 0,
                         BorderMode border = BorderMode::Clamp)
{
    using T = typename ImageT::value_type;
    static_assert(std::is_arithmetic_v<T>, "ndl::gaussian_downsample() requires an arithmetic value_type");
    constexpr int DIM = detail::dim_of<ImageT>;
    static_assert(DIM >= 2, "ndl::gaussian_downsample() needs DIM >= 2");
    assert(sigma > 0 && factor > 1);
    auto e = src.extent();
    std::array<int, DIM> oe;
    for (int i = 0; i < DIM; i++) oe[i] = i == channelAxis ? e[i] : (e[i] - 1) / factor + 1;
    OwnedImage<T, DIM> result(oe);
    std::array<double, DIM> sg;
    sg.fill(sigma);
    const int32_t rounding = std::is_integral_v<T> ? NDL_ROUND_NEAREST : NDL_ROUND_TRUNC;
    detail::DeviceSource<ImageT> s(src);
    detail::DeviceDest<OwnedImage<T, DIM>> o(result);
    detail::check(ndl_gaussian_downsample(&s.t, &o.t, sg.data(), detail::border_code(border), factor, channelAxis,
                                          NDL_ACCUM_REF, rounding, nullptr));
    o.finish();
    return result;
}

/// Spatial gradient (processing/convolution.h:359-388): dst.slice(0, axis) is
/// the central difference {-0.5, 0, +0.5} along `axis` (the zero tap is
/// skipped, as in convolve()), for every axis; dst has a leading component
/// axis of size DIM.  All components in one launch.
template <class SrcImageT, class DstImageT>
void gradient(const SrcImageT& src, DstImageT& dst, BorderMode border = BorderMode::Reflect)
{
    using SrcT = typename SrcImageT::value_type;
    static_assert(std::is_arithmetic_v<SrcT>, "ndl::gradient() requires an arithmetic value_type");
    constexpr int DIM = detail::dim_of<SrcImageT>;
    constexpr int DstDIM = detail::dim_of<DstImageT>;
    static_assert(DstDIM == DIM + 1, "ndl::gradient() needs a destination with a leading component axis");
    assert(dst.extent()[0] == DIM);
    for (int d = 0; d < DIM; d++) assert(dst.extent()[d + 1] == src.extent()[d]);
    detail::DeviceSource<SrcImageT> s(src);
    detail::DeviceDest<DstImageT> o(dst);
    detail::check(ndl_gradient(&s.t, &o.t, detail::border_code(border), nullptr));
    o.finish();
}

/// Input gradient of L = <w, convolve_separable(x)>: grad_in = A^T grad_out,
/// the exact transpose including border folding (no reference counterpart).
template <class GradOutT, class GradInT, int DIM>
void convolve_adjoint(const GradOutT& grad_out, GradInT& grad_in, const std::array<std::vector<double>, DIM>& taps,
                      const std::array<BorderMode, DIM>& border, Accumulation acc = Accumulation::Reference)
{
    static_assert(std::is_floating_point_v<typename GradOutT::value_type> &&
                  std::is_floating_point_v<typename GradInT::value_type>, "adjoint fields are floating point");
    assert(grad_out.extent() == grad_in.extent());
    detail::AxisPlan<DIM> plan;
    for (int d = 0; d < DIM; d++)
        if (!taps[d].empty()) plan.add(d, taps[d], (int)taps[d].size() / 2, border[d]);
    plan.bind();
    detail::DeviceSource<GradOutT> s(grad_out);
    detail::DeviceDest<GradInT> o(grad_in);
    detail::check(ndl_sepconv_adjoint_input(&s.t, &o.t, (int32_t)plan.axes.size(), plan.axes.data(), plan.k.data(),
                                            (int32_t)acc, nullptr));
    o.finish();
}

/// Tap gradients of L = <grad_out, convolve_separable(x)>: one vector per
/// axis (empty for untouched axes), deterministic FP64 reduction.
template <class XImageT, class GradOutT, int DIM>
std::array<std::vector<double>, DIM> convolve_adjoint_taps(const XImageT& x, const GradOutT& grad_out,
                                                           const std::array<std::vector<double>, DIM>& taps,
                                                           const std::array<BorderMode, DIM>& border,
                                                           Accumulation acc = Accumulation::Reference)
{
    assert(x.extent() == grad_out.extent());
    detail::AxisPlan<DIM> plan;
    for (int d = 0; d < DIM; d++)
        if (!taps[d].empty()) plan.add(d, taps[d], (int)taps[d].size() / 2, border[d]);
    plan.bind();
    std::size_t total = 0;
    for (auto& t : plan.taps) total += t.size();
    detail::DeviceSource<XImageT> sx(x);
    detail::DeviceSource<GradOutT> sg(grad_out);
    detail::DeviceBuffer g(std::max<std::size_t>(total, 1) * sizeof(double));
    detail::check(ndl_sepconv_adjoint_taps(&sx.t, &sg.t, (int32_t)plan.axes.size(), plan.axes.data(), plan.k.data(),
                                           (double*)g.p, (int32_t)acc, nullptr));
    std::vector<double> flat(total);
    detail::check(ndl_stream_synchronize(nullptr));
    if (total) detail::check(ndl_memcpy_dtoh(flat.data(), g.p, total * sizeof(double), nullptr));
    detail::check(ndl_stream_synchronize(nullptr));
    std::array<std::vector<double>, DIM> out;
    std::size_t off = 0;
    for (std::size_t i = 0; i < plan.axes.size(); i++) {
        out[plan.axes[i]].assign(flat.begin() + off, flat.begin() + off + plan.taps[i].size());
        off += plan.taps[i].size();
    }
    return out;
}

} // namespace ndl
