// ndl/image.h -- header-only strided image views for the GPU drop-in.
//
// The subset of the reference's Image<T,DIM> contract that the convolution
// path touches (image/core.h:50-972), rewritten here: a non-owning view of
// {data, extent, stride} with int extents and signed element strides,
// dimension 0 fastest, default strides {1, e0, e0*e1, ...}
// (image/core.h:858-864); at() is the stride dot product
// (image/core.h:557-558); view()/slice()/swap_axes()/mirror() compose
// zero-copy views with the reference's semantics (image/core.h:647-757);
// OwnedImage owns a std::vector (image/owned.h:84-151).  BorderMode has the
// reference's enumerators in the reference's order (image/border_mode.h:14-19).
// DeviceImage is this project's addition: the same view over device memory.
#pragma once

#include <algorithm>
#include <array>
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <iterator>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../ndl_gpu.h"

namespace ndl {

/// How an out-of-range neighbour is resolved (image/border_mode.h:14-19).
enum class BorderMode { Clamp, Wrap, Reflect };

namespace detail {
// _ceil (mathHelpers.h:17): truncation-based ceil of a non-negative value
inline int ceil_int(double num)
{
    int t = static_cast<int>(num);
    return (static_cast<double>(t) == num) ? t : t + ((num > 0) ? 1 : 0);
}
template <int DIM> std::array<int, DIM> default_stride(const std::array<int, DIM>& extent)
{
    std::array<int, DIM> s{};
    s[0] = 1;
    for (int i = 1; i < DIM; i++) s[i] = extent[i - 1] * s[i - 1];
    return s;
}
// every coordinate of `extent`, dimension 0 fastest (detail.h:127-143 order)
template <int DIM> std::vector<std::array<int, DIM>> coordinates_of(const std::array<int, DIM>& extent)
{
    std::vector<std::array<int, DIM>> out;
    std::size_t n = 1;
    for (int e : extent) n *= (std::size_t)std::max(e, 0);
    out.reserve(n);
    std::array<int, DIM> c{};
    for (std::size_t i = 0; i < n; i++) {
        out.push_back(c);
        for (int d = 0; d < DIM; d++) {
            if (++c[d] < extent[d]) break;
            c[d] = 0;
        }
    }
    return out;
}
} // namespace detail

template <class T, int DIM> class Image {
    static_assert(DIM >= 1 && DIM <= NDL_MAX_DIM, "Image<T,DIM>: DIM must be 1..8");

public:
    typedef T value_type;
    static constexpr int dimensions = DIM;

    /// View over `buffer` with the default (dense, dimension-0-fastest) strides.
    Image(T* buffer, std::array<int, DIM> extent)
        : extent_(extent), stride_(detail::default_stride<DIM>(extent)), data_(buffer) {}
    /// View with explicit signed element strides; `data` is element (0,...,0).
    Image(T* data, std::array<int, DIM> extent, std::array<int, DIM> stride)
        : extent_(extent), stride_(stride), data_(data) {}

    static std::size_t size(std::array<int, DIM> extent)
    {
        return std::accumulate(extent.begin(), extent.end(), std::size_t(1), std::multiplies<std::size_t>());
    }
    std::size_t size() const { return size(extent_); }
    const std::array<int, DIM>& extent() const { return extent_; }
    const std::array<int, DIM>& stride() const { return stride_; }
    T* data() const { return data_; }

    T& at(const std::array<int, DIM>& index)
    {
        assert(valid(index));
        return data_[offset(index)];
    }
    const T& at(const std::array<int, DIM>& index) const
    {
        assert(valid(index));
        return data_[offset(index)];
    }
    template <class... Args> T& operator()(Args... idx)
    {
        static_assert(sizeof...(Args) == DIM, "Image::operator() needs exactly DIM indices");
        return at({static_cast<int>(idx)...});
    }
    template <class... Args> const T& operator()(Args... idx) const
    {
        static_assert(sizeof...(Args) == DIM, "Image::operator() needs exactly DIM indices");
        return at({static_cast<int>(idx)...});
    }

    std::vector<std::array<int, DIM>> coordinates() const { return detail::coordinates_of<DIM>(extent_); }

    /// Zero-copy region/mirror/decimation view: per dimension inclusive
    /// [start, end] (negative counts from the end) and a step; a negative step
    /// walks backward anchored at `end` (image/core.h:647-703).
    Image view(std::initializer_list<int> start, std::initializer_list<int> end,
               std::initializer_list<int> step = {}) const
    {
        std::array<int, DIM> ext, str;
        T* p = data_;
        auto si = start.begin(), ei = end.begin(), ti = step.begin();
        for (int i = 0; i < DIM; i++) {
            int s = si != start.end() ? *si++ : 0;
            int e = ei != end.end() ? *ei++ : -1;
            int st = ti != step.end() ? *ti++ : 1;
            if (st == 0) throw std::out_of_range("Image::view: step must not be 0 (dimension " + std::to_string(i) + ")");
            if (s < 0) s += extent_[i];
            if (e < 0) e += extent_[i];
            if (s < 0 || s >= extent_[i] || e < 0 || e >= extent_[i] || e < s)
                throw std::out_of_range("Image::view: invalid range on dimension " + std::to_string(i));
            int mag = st < 0 ? -st : st;
            p += (std::ptrdiff_t)(st < 0 ? e : s) * stride_[i];
            ext[i] = detail::ceil_int((float)(1 + e - s) / mag);
            str[i] = stride_[i] * st;
        }
        return Image(p, ext, str);
    }
    /// Fixes one dimension: a (DIM-1)-dimensional view (image/core.h:709-712).
    Image<T, DIM - 1> slice(int dim, int index) const
    {
        static_assert(DIM > 1, "slice() needs DIM > 1");
        std::array<int, DIM - 1> ext, str;
        for (int i = 0, t = 0; i < DIM; i++)
            if (i != dim) { ext[t] = extent_[i]; str[t] = stride_[i]; t++; }
        return Image<T, DIM - 1>(data_ + (std::ptrdiff_t)index * stride_[dim], ext, str);
    }
    /// Two dimensions exchanged (image/core.h:717-731).
    Image swap_axes(int d1, int d2) const
    {
        std::array<int, DIM> ext = extent_, str = stride_;
        std::swap(ext[d1], ext[d2]);
        std::swap(str[d1], str[d2]);
        return Image(data_, ext, str);
    }
    /// One dimension reversed (image/core.h:735-757).
    Image mirror(int dim) const
    {
        std::array<int, DIM> str = stride_;
        str[dim] = -str[dim];
        return Image(data_ + (std::ptrdiff_t)(extent_[dim] - 1) * stride_[dim], extent_, str);
    }

    /// Every element set to `value` (make_box_kernel's `kernel = T(1)`).
    Image& operator=(const T& value)
    {
        for (const auto& c : coordinates()) at(c) = value;
        return *this;
    }
    template <class U> Image& assign(const Image<U, DIM>& rhs)
    {
        assert(rhs.extent() == extent_);
        for (const auto& c : coordinates()) at(c) = static_cast<T>(rhs.at(c));
        return *this;
    }
    Image(const Image&) = default;

    /// Element-wise equality of two same-extent views (values, not layout).
    template <class U> bool operator==(const Image<U, DIM>& rhs) const
    {
        if (rhs.extent() != extent_) return false;
        for (const auto& c : coordinates())
            if (!(at(c) == rhs.at(c))) return false;
        return true;
    }

    /// Forward iteration in coordinate order (dimension 0 fastest).
    template <bool IsConst> class basic_iterator {
    public:
        using iterator_category = std::forward_iterator_tag;
        using value_type = T;
        using difference_type = std::ptrdiff_t;
        using pointer = std::conditional_t<IsConst, const T*, T*>;
        using reference = std::conditional_t<IsConst, const T&, T&>;
        basic_iterator(const Image* img, bool end) : img_(img), idx_{}, done_(end || img->size() == 0) {}
        reference operator*() const { return img_->data_[img_->offset(idx_)]; }
        basic_iterator& operator++()
        {
            for (int d = 0; d < DIM; d++) {
                if (++idx_[d] < img_->extent_[d]) return *this;
                idx_[d] = 0;
            }
            done_ = true;
            return *this;
        }
        bool operator==(const basic_iterator& o) const { return done_ == o.done_ && (done_ || idx_ == o.idx_); }
        bool operator!=(const basic_iterator& o) const { return !(*this == o); }

    private:
        const Image* img_;
        std::array<int, DIM> idx_;
        bool done_;
    };
    using iterator = basic_iterator<false>;
    using const_iterator = basic_iterator<true>;
    iterator begin() { return iterator(this, false); }
    iterator end() { return iterator(this, true); }
    const_iterator begin() const { return const_iterator(this, false); }
    const_iterator end() const { return const_iterator(this, true); }

    /// Descriptor of this view for the C ABI (any memory space).
    ndl_tensor tensor(int32_t dtype) const
    {
        ndl_tensor t{};
        t.data = (void*)data_;
        t.dtype = dtype;
        t.ndim = DIM;
        for (int i = 0; i < DIM; i++) {
            t.extent[i] = extent_[i];
            t.stride[i] = stride_[i];
        }
        return t;
    }

protected:
    std::ptrdiff_t offset(const std::array<int, DIM>& idx) const
    {
        std::ptrdiff_t o = 0;
        for (int i = 0; i < DIM; i++) o += (std::ptrdiff_t)idx[i] * stride_[i];
        return o;
    }
    bool valid(const std::array<int, DIM>& idx) const
    {
        for (int i = 0; i < DIM; i++)
            if (idx[i] < 0 || idx[i] >= extent_[i]) return false;
        return true;
    }
    std::array<int, DIM> extent_;
    std::array<int, DIM> stride_;
    T* data_;
};

namespace detail {
template <class T> struct VecStore {
    explicit VecStore(std::size_t n) : storage_(n) {}
    std::vector<T> storage_;
};
} // namespace detail

/// Move-only Image that owns a zero-initialised std::vector (image/owned.h:84-151).
template <class T, int DIM> class OwnedImage : private detail::VecStore<T>, public Image<T, DIM> {
    static_assert(!std::is_same_v<T, bool>, "OwnedImage<bool,DIM>: std::vector<bool> has no data()");
    using Store = detail::VecStore<T>;

public:
    explicit OwnedImage(std::array<int, DIM> extent)
        : Store(Image<T, DIM>::size(extent)), Image<T, DIM>(Store::storage_.data(), extent) {}
    OwnedImage(std::array<int, DIM> extent, std::initializer_list<T> values)
        : Store(Image<T, DIM>::size(extent)), Image<T, DIM>(Store::storage_.data(), extent)
    {
        assert(values.size() == Store::storage_.size());
        std::copy(values.begin(), values.end(), Store::storage_.begin());
    }
    template <class U>
    explicit OwnedImage(const Image<U, DIM>& source)
        : Store(source.size()), Image<T, DIM>(Store::storage_.data(), source.extent())
    {
        this->assign(source);
    }
    OwnedImage(const OwnedImage&) = delete;
    OwnedImage& operator=(const OwnedImage&) = delete;
    OwnedImage(OwnedImage&& o) noexcept : Store(std::move(o)), Image<T, DIM>(Store::storage_.data(), o.extent()) {}
    using Image<T, DIM>::operator=;
    template <class U> static OwnedImage like(const Image<U, DIM>& source) { return OwnedImage(source.extent()); }
};

/// Fills `kernel` with ones: the box kernel (image/kernels.h:22-23).
template <class T, int DIM> void make_box_kernel(Image<T, DIM>& kernel) { kernel = T(1); }

/// Plus-sign structuring element (image/kernels.h:44-58): the centre and,
/// per axis, the taps that differ from the centre along that axis only.
template <class T, int DIM> void make_cross_kernel(Image<T, DIM>& kernel)
{
    kernel = T(0);
    std::array<int, DIM> centre;
    for (int i = 0; i < DIM; i++) centre[i] = kernel.extent()[i] / 2;
    kernel.at(centre) = T(1);
    for (int axis = 0; axis < DIM; axis++) {
        std::array<int, DIM> c = centre;
        for (int i = 0; i < kernel.extent()[axis]; i++) {
            c[axis] = i;
            kernel.at(c) = T(1);
        }
    }
}

/// Runs fn(srcChannel, dstChannel) on every slice along channelAxis
/// (image/kernels.h:78-87).
template <class T, int DIM, class Fn>
void per_channel(const Image<T, DIM>& src, Image<T, DIM>& dst, int channelAxis, Fn&& fn)
{
    for (int c = 0; c < src.extent()[channelAxis]; c++) {
        Image<T, DIM - 1> s = src.slice(channelAxis, c);
        Image<T, DIM - 1> d = dst.slice(channelAxis, c);
        fn(s, d);
    }
}

// ---------------------------------------------------------------- device side
namespace detail {
inline void check(int rc)
{
    if (rc < 0) throw std::runtime_error(std::string("ndl GPU error ") + std::to_string(rc) + ": " + ndl_last_error());
}
template <class T> constexpr int32_t dtype_of()
{
    if constexpr (std::is_same_v<T, uint8_t>) return NDL_U8;
    else if constexpr (std::is_same_v<T, int8_t>) return NDL_I8;
    else if constexpr (std::is_same_v<T, uint16_t>) return NDL_U16;
    else if constexpr (std::is_same_v<T, int16_t>) return NDL_I16;
    else if constexpr (std::is_same_v<T, int32_t>) return NDL_I32;
    else if constexpr (std::is_same_v<T, float>) return NDL_F32;
    else if constexpr (std::is_same_v<T, double>) return NDL_F64;
    else return -1;
}
template <class T> constexpr bool gpu_type_v = dtype_of<T>() >= 0;

// device buffer owned by RAII
struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(std::size_t bytes) { check(ndl_malloc(&p, bytes)); }
    ~DeviceBuffer() { if (p) ndl_free(p); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p(o.p) { o.p = nullptr; }
};
} // namespace detail

/// A dense image in device memory (x fastest).  Owning; move-only.
// Image-type contracts of the reference's free functions (image/detail.h:
// 48-57 "minimal interface", 90-98 "stride access"), stated here with a
// detection helper of this library: a minimal-interface image has extent()
// (convertible to std::array<int,DIM>), at(coord) and coordinates(); one with
// stride access also has stride() and an at() returning a real T& into
// linear memory.  The GPU path accepts every type with stride access (and
// DeviceImage); bit-packed or ring-buffer images must be converted first.
namespace detail {
template <class Void, template <class...> class Op, class... A> struct detected : std::false_type {};
template <template <class...> class Op, class... A>
struct detected<std::void_t<Op<A...>>, Op, A...> : std::true_type {};
template <class I> using extent_expr = decltype(std::declval<const I&>().extent());
template <class I> using stride_expr = decltype(std::declval<const I&>().stride());
template <class I> using coordinates_expr = decltype(std::declval<const I&>().coordinates());
template <class I, class C> using at_expr = decltype(std::declval<I&>().at(std::declval<C>()));
template <class I, int DIM, bool = detected<void, extent_expr, I>::value> struct extent_is : std::false_type {};
template <class I, int DIM>
struct extent_is<I, DIM, true> : std::is_convertible<extent_expr<I>, std::array<int, DIM>> {};
template <class I, int DIM, bool = detected<void, stride_expr, I>::value> struct stride_is : std::false_type {};
template <class I, int DIM>
struct stride_is<I, DIM, true> : std::is_convertible<stride_expr<I>, std::array<int, DIM>> {};
template <class I, int DIM, bool = detected<void, at_expr, I, std::array<int, DIM>>::value>
struct at_is_ref : std::false_type {};
template <class I, int DIM>
struct at_is_ref<I, DIM, true> : std::is_lvalue_reference<at_expr<I, std::array<int, DIM>>> {};

template <class ImageT, int DIM>
inline constexpr bool satisfies_minimal_interface_v =
    extent_is<ImageT, DIM>::value && detected<void, at_expr, const ImageT, std::array<int, DIM>>::value &&
    detected<void, coordinates_expr, ImageT>::value;
template <class ImageT, int DIM>
inline constexpr bool has_stride_access_v = stride_is<ImageT, DIM>::value && at_is_ref<ImageT, DIM>::value;
} // namespace detail

template <class T, int DIM> class DeviceImage {
    static_assert(detail::gpu_type_v<T>, "DeviceImage<T,DIM>: T must be u8/i8/u16/i16/i32/float/double");

public:
    typedef T value_type;
    explicit DeviceImage(std::array<int, DIM> extent)
        : extent_(extent), stride_(detail::default_stride<DIM>(extent)),
          buf_(std::max<std::size_t>(Image<T, DIM>::size(extent), 1) * sizeof(T)) {}
    template <class Host> explicit DeviceImage(const Host& host) : DeviceImage(host.extent()) { upload(host); }
    const std::array<int, DIM>& extent() const { return extent_; }
    const std::array<int, DIM>& stride() const { return stride_; }
    std::size_t size() const { return Image<T, DIM>::size(extent_); }
    T* data() const { return (T*)buf_.p; }
    ndl_tensor tensor() const { return Image<T, DIM>(data(), extent_, stride_).tensor(detail::dtype_of<T>()); }
    /// Copies a same-extent host image in (any strides).
    void upload(const Image<T, DIM>& host, void* stream = nullptr)
    {
        assert(host.extent() == extent_);
        if (host.stride() == stride_) {
            detail::check(ndl_memcpy_htod(buf_.p, host.data(), size() * sizeof(T), stream));
        } else {
            OwnedImage<T, DIM> dense(extent_);
            dense.assign(host);
            detail::check(ndl_memcpy_htod(buf_.p, dense.data(), size() * sizeof(T), stream));
        }
        detail::check(ndl_stream_synchronize(stream));
    }
    /// Copies into a same-extent host image (any strides).
    void download(Image<T, DIM>& host, void* stream = nullptr) const
    {
        assert(host.extent() == extent_);
        detail::check(ndl_stream_synchronize(stream));
        if (host.stride() == stride_) {
            detail::check(ndl_memcpy_dtoh(host.data(), buf_.p, size() * sizeof(T), stream));
            detail::check(ndl_stream_synchronize(stream));
        } else {
            OwnedImage<T, DIM> dense(extent_);
            detail::check(ndl_memcpy_dtoh(dense.data(), buf_.p, size() * sizeof(T), stream));
            detail::check(ndl_stream_synchronize(stream));
            host.assign(dense);
        }
    }
    OwnedImage<T, DIM> to_host() const
    {
        OwnedImage<T, DIM> h(extent_);
        download(h);
        return h;
    }

private:
    std::array<int, DIM> extent_;
    std::array<int, DIM> stride_;
    detail::DeviceBuffer buf_;
};

} // namespace ndl
