// ref_shim.cpp -- extern "C" entry points over the REAL reference templates.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference headers where they lie (-I <reference root>) into
// oracle/_ref/libndl_ref.so; it is built only when the reference sources are
// present.  Used to pin the oracle restatement (tests/test_oracle.py,
// test_oracle_morph.py, test_oracle_downsample.py, test_image_io.py)
// and, optionally, as the timed CPU baseline.  No reference source is copied:
// this file only instantiates ndl::convolve / ndl::gaussian_blur
// (processing/convolution.h:52-71,107-186) on dense Image<T,DIM> views
// (image/core.h:228-235) and composes per-axis passes the way a caller of the
// reference would (double intermediates, as convolution.h:151-185 does);
// ndl::erode / ndl::dilate (processing/morphology.h:57-101) and
// ndl::downsample (processing/convolution.h:241-305) likewise; NRRD::save /
// NRRD::load (imageIO/NRRD/nrrd.h:114-164, 233-366) write and read the files
// the image_io tests compare against.
#include <array>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "imageIO/NRRD/nrrd.h"
#include "processing/convolution.h"
#include "processing/morphology.h"

#include "../include/ndl_gpu.h"

namespace {

template <int DIM>
std::array<int, DIM> ext_of(const ndl_tensor* t)
{
    std::array<int, DIM> e{};
    for (int d = 0; d < DIM; d++) e[d] = (int)t->extent[d];
    return e;
}

ndl::BorderMode bmode(int b)
{
    return b == NDL_WRAP ? ndl::BorderMode::Wrap
                         : (b == NDL_REFLECT ? ndl::BorderMode::Reflect : ndl::BorderMode::Clamp);
}

template <class T, int DIM>
int convolve_t(const ndl_tensor* src, const ndl_tensor* dst, const double* kernel,
               const int64_t* kextent, int border)
{
    ndl::Image<T, DIM> s((T*)src->data, ext_of<DIM>(src));
    ndl::Image<T, DIM> d((T*)dst->data, ext_of<DIM>(dst));
    std::array<int, DIM> ke{};
    for (int i = 0; i < DIM; i++) ke[i] = (int)kextent[i];
    ndl::Image<double, DIM> k(const_cast<double*>(kernel), ke);
    ndl::convolve(s, d, k, bmode(border));
    return 0;
}

template <class T, int DIM>
int gaussian_t(const ndl_tensor* src, const ndl_tensor* dst, double sigma, int border)
{
    ndl::Image<T, DIM> s((T*)src->data, ext_of<DIM>(src));
    ndl::Image<T, DIM> d((T*)dst->data, ext_of<DIM>(dst));
    ndl::gaussian_blur(s, d, sigma, bmode(border));
    return 0;
}

// A caller's per-axis chain: convolve() with {1,..,count,..,1} kernels
// through Image<double> buffers, last pass into dst (rounding TRUNC) or into a
// double buffer followed by std::round + cast (rounding NEAREST).
template <class T, int DIM>
int sepconv_t(const ndl_tensor* src, const ndl_tensor* dst, int naxes, const int32_t* axes,
              const ndl_axis_kernel* k, int rounding)
{
    auto ext = ext_of<DIM>(src);
    ndl::Image<T, DIM> s((T*)src->data, ext);
    ndl::Image<T, DIM> d((T*)dst->data, ext);
    size_t n = ndl::Image<double, DIM>::size(ext);
    std::vector<double> a(n), b(n);
    ndl::Image<double, DIM> A(a.data(), ext), B(b.data(), ext);
    std::vector<std::vector<double>> kd(naxes);
    std::vector<ndl::Image<double, DIM>> kimg;
    kimg.reserve(naxes);
    for (int i = 0; i < naxes; i++) {
        if (k[i].center != k[i].count / 2) return NDL_ERR_ARG;
        std::array<int, DIM> ke;
        ke.fill(1);
        ke[axes[i]] = k[i].count;
        kd[i].assign(k[i].taps, k[i].taps + k[i].count);
        kimg.emplace_back(kd[i].data(), ke);
    }
    bool round_last = rounding == NDL_ROUND_NEAREST && std::is_integral<T>::value;
    if (naxes == 1 && !round_last) {
        ndl::convolve(s, d, kimg[0], bmode(k[0].border));
        return 0;
    }
    ndl::convolve(s, A, kimg[0], bmode(k[0].border));
    ndl::Image<double, DIM>* cur = &A;
    ndl::Image<double, DIM>* oth = &B;
    for (int i = 1; i < naxes; i++) {
        if (i == naxes - 1 && !round_last) {
            ndl::convolve(*cur, d, kimg[i], bmode(k[i].border));
            return 0;
        }
        ndl::convolve(*cur, *oth, kimg[i], bmode(k[i].border));
        std::swap(cur, oth);
    }
    for (const auto& c : cur->coordinates())
        d.at(c) = static_cast<T>(std::round(cur->at(c)));
    return 0;
}

template <class T>
int dispatch_dim_conv(const ndl_tensor* s, const ndl_tensor* d, const double* k, const int64_t* ke,
                      int b)
{
    switch (s->ndim) {
    case 1: return convolve_t<T, 1>(s, d, k, ke, b);
    case 2: return convolve_t<T, 2>(s, d, k, ke, b);
    case 3: return convolve_t<T, 3>(s, d, k, ke, b);
    case 4: return convolve_t<T, 4>(s, d, k, ke, b);
    }
    return NDL_ERR_SHAPE;
}

template <class T>
int dispatch_dim_gauss(const ndl_tensor* s, const ndl_tensor* d, double sigma, int b)
{
    switch (s->ndim) {
    case 1: return gaussian_t<T, 1>(s, d, sigma, b);
    case 2: return gaussian_t<T, 2>(s, d, sigma, b);
    case 3: return gaussian_t<T, 3>(s, d, sigma, b);
    case 4: return gaussian_t<T, 4>(s, d, sigma, b);
    }
    return NDL_ERR_SHAPE;
}

template <class T>
int dispatch_dim_sep(const ndl_tensor* s, const ndl_tensor* d, int na, const int32_t* ax,
                     const ndl_axis_kernel* k, int r)
{
    switch (s->ndim) {
    case 1: return sepconv_t<T, 1>(s, d, na, ax, k, r);
    case 2: return sepconv_t<T, 2>(s, d, na, ax, k, r);
    case 3: return sepconv_t<T, 3>(s, d, na, ax, k, r);
    case 4: return sepconv_t<T, 4>(s, d, na, ax, k, r);
    }
    return NDL_ERR_SHAPE;
}

template <class T, int DIM>
int morph_t(const ndl_tensor* src, const ndl_tensor* dst, const double* kernel, const int64_t* kextent,
            int border, int op)
{
    ndl::Image<T, DIM> s((T*)src->data, ext_of<DIM>(src));
    ndl::Image<T, DIM> d((T*)dst->data, ext_of<DIM>(dst));
    std::array<int, DIM> ke{};
    for (int i = 0; i < DIM; i++) ke[i] = (int)kextent[i];
    ndl::Image<double, DIM> k(const_cast<double*>(kernel), ke);
    if (op == 0) ndl::erode(s, d, k, bmode(border));
    else ndl::dilate(s, d, k, bmode(border));
    return 0;
}

template <class T>
int dispatch_dim_morph(const ndl_tensor* s, const ndl_tensor* d, const double* k, const int64_t* ke, int b,
                       int op)
{
    switch (s->ndim) {
    case 1: return morph_t<T, 1>(s, d, k, ke, b, op);
    case 2: return morph_t<T, 2>(s, d, k, ke, b, op);
    case 3: return morph_t<T, 3>(s, d, k, ke, b, op);
    case 4: return morph_t<T, 4>(s, d, k, ke, b, op);
    }
    return NDL_ERR_SHAPE;
}

template <class T, int DIM>
int downsample_t(const ndl_tensor* src, const ndl_tensor* dst, int factor, int channel_axis)
{
    ndl::Image<T, DIM> s((T*)src->data, ext_of<DIM>(src));
    auto r = ndl::downsample(s, factor, channel_axis);
    auto e = r.extent();
    for (int d = 0; d < DIM; d++)
        if (e[d] != dst->extent[d]) return NDL_ERR_SHAPE;
    ndl::Image<T, DIM> d((T*)dst->data, e);
    for (const auto& c : r.coordinates()) d.at(c) = r.at(c);
    return 0;
}

template <class T>
int dispatch_dim_down(const ndl_tensor* s, const ndl_tensor* d, int factor, int channel_axis)
{
    switch (s->ndim) {
    case 2: return downsample_t<T, 2>(s, d, factor, channel_axis);
    case 3: return downsample_t<T, 3>(s, d, factor, channel_axis);
    case 4: return downsample_t<T, 4>(s, d, factor, channel_axis);
    }
    return NDL_ERR_SHAPE;
}

bool dense_same(const ndl_tensor* s, const ndl_tensor* d)
{
    if (s->dtype != d->dtype || s->ndim != d->ndim) return false;
    int64_t st = 1;
    for (int i = 0; i < s->ndim; i++) {
        if (s->extent[i] != d->extent[i] || s->stride[i] != st || d->stride[i] != st) return false;
        st *= s->extent[i];
    }
    return true;
}

} // namespace

#define NDL_REF_DISPATCH(FN, ...)                                        \
    switch (src->dtype) {                                                \
    case NDL_U8: return FN<uint8_t>(__VA_ARGS__);                         \
    case NDL_I8: return FN<int8_t>(__VA_ARGS__);                          \
    case NDL_U16: return FN<uint16_t>(__VA_ARGS__);                       \
    case NDL_I16: return FN<int16_t>(__VA_ARGS__);                        \
    case NDL_I32: return FN<int32_t>(__VA_ARGS__);                        \
    case NDL_F32: return FN<float>(__VA_ARGS__);                          \
    case NDL_F64: return FN<double>(__VA_ARGS__);                         \
    }                                                                    \
    return NDL_ERR_DTYPE;

extern "C" {

// convolve() on dense same-type src/dst
__attribute__((visibility("default"))) int ref_convolve(const ndl_tensor* src, const ndl_tensor* dst,
                                                        const double* kernel, const int64_t* kextent,
                                                        int border)
{
    if (!dense_same(src, dst)) return NDL_ERR_ARG;
    NDL_REF_DISPATCH(dispatch_dim_conv, src, dst, kernel, kextent, border)
}

// gaussian_blur() on dense same-type src/dst
__attribute__((visibility("default"))) int ref_gaussian_blur(const ndl_tensor* src,
                                                             const ndl_tensor* dst, double sigma,
                                                             int border)
{
    if (!dense_same(src, dst)) return NDL_ERR_ARG;
    NDL_REF_DISPATCH(dispatch_dim_gauss, src, dst, sigma, border)
}

// per-axis convolve() chain on dense same-type src/dst
__attribute__((visibility("default"))) int ref_sepconv(const ndl_tensor* src, const ndl_tensor* dst,
                                                       int naxes, const int32_t* axes,
                                                       const ndl_axis_kernel* k, int rounding)
{
    if (!dense_same(src, dst) || naxes < 1) return NDL_ERR_ARG;
    NDL_REF_DISPATCH(dispatch_dim_sep, src, dst, naxes, axes, k, rounding)
}

// erode() (op 0) / dilate() (op 1) on dense same-type src/dst
__attribute__((visibility("default"))) int ref_morph(const ndl_tensor* src, const ndl_tensor* dst,
                                                     const double* kernel, const int64_t* kextent, int border,
                                                     int op)
{
    if (!dense_same(src, dst) || (op != 0 && op != 1)) return NDL_ERR_ARG;
    NDL_REF_DISPATCH(dispatch_dim_morph, src, dst, kernel, kextent, border, op)
}

// downsample() into a dense dst of the result's extent (src dense)
__attribute__((visibility("default"))) int ref_downsample(const ndl_tensor* src, const ndl_tensor* dst, int factor,
                                                          int channel_axis)
{
    if (src->dtype != dst->dtype || factor < 2) return NDL_ERR_ARG;
    NDL_REF_DISPATCH(dispatch_dim_down, src, dst, factor, channel_axis)
}

// NRRD::save<T> of n-dimensional dense data (0 on success)
#define NDL_REF_NRRD_SAVE(NAME, T)                                                                    \
    __attribute__((visibility("default"))) int NAME(const char* path, const T* data, int n, const int* size) \
    {                                                                                                  \
        return NRRD::save<T>(path, data, n, size) ? 0 : 1;                                            \
    }
NDL_REF_NRRD_SAVE(ref_nrrd_save_u8, unsigned char)
NDL_REF_NRRD_SAVE(ref_nrrd_save_i16, short)
NDL_REF_NRRD_SAVE(ref_nrrd_save_u16, unsigned short)
NDL_REF_NRRD_SAVE(ref_nrrd_save_i32, int)
NDL_REF_NRRD_SAVE(ref_nrrd_save_u32, unsigned int)
NDL_REF_NRRD_SAVE(ref_nrrd_save_f32, float)
NDL_REF_NRRD_SAVE(ref_nrrd_save_f64, double)
#undef NDL_REF_NRRD_SAVE

// NRRD::load<double>: the file's values converted to double, its sizes and
// dimension; 0 on success, 1 when the reference's loader fails, 2 when the
// data does not fit `cap`
__attribute__((visibility("default"))) int ref_nrrd_load_f64(const char* path, double* out, long cap, int* sizes,
                                                             int* ndim)
{
    double* data = nullptr;
    std::vector<int> size;
    if (!NRRD::load<double>(path, &data, &size)) return 1;
    long n = 1;
    for (int v : size) n *= v;
    int rc = 0;
    if (n > cap || size.size() > 10) rc = 2;
    else {
        for (long i = 0; i < n; i++) out[i] = data[i];
        for (size_t d = 0; d < size.size(); d++) sizes[d] = size[d];
        *ndim = (int)size.size();
    }
    delete[] data;
    return rc;
}

// mathHelpers.h _kernelSize
__attribute__((visibility("default"))) int ref_kernel_size(double sigma)
{
    return ndl::_kernelSize(sigma);
}

} // extern "C"
