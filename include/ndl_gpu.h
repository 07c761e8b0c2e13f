/*
 * ndl_gpu.h -- C ABI of libndl_gpu.so: separable N-D convolution on one
 * MI355X (gfx950), the device side of ndl's convolve()/gaussian_blur().
 *
 * The reference exposes this path as header-only C++17 templates with no
 * FFI of its own:
 *   convolve(src, dst, kernel, border)        processing/convolution.h:52-71
 *   gaussian_blur(src, dst, sigma, border)    processing/convolution.h:107-186
 * over strided Image<T,DIM> views (image/core.h:50-972) with the border maps
 * of image/border_mode.h:14-19 and mathHelpers.h:23-25.  This header is the
 * plain-C boundary those templates call in the drop-in C++ layer
 * (include/ndl/convolution.h); every entry point below names the reference
 * interface it replaces.
 *
 * Conventions (all entry points):
 *  - Tensors are non-owning descriptors of DEVICE memory: `data` points at the
 *    element with index (0,...,0); strides are signed element counts, one per
 *    dimension, dimension 0 first (the reference's x-fastest order,
 *    image/core.h:858-864).  Extents and strides are 64-bit.
 *  - The caller owns every buffer.  The library never frees caller memory.
 *  - Calls are stream-ordered and asynchronous on `stream` (a hipStream_t;
 *    NULL = the default stream).  Host-side argument checks complete before
 *    the call returns.
 *  - Return value: NDL_OK (0) or a negative NDL_ERR_* code; the message of the
 *    last failure on the calling thread is returned by ndl_last_error().
 *    Nothing throws across this boundary.
 */
#ifndef NDL_GPU_H
#define NDL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NDL_MAX_DIM 8
#define NDL_MAX_TAPS 1024 /* per-axis tap limit of the separable entry points */

/* element types (Image<T,DIM>::value_type of the reference) */
enum ndl_dtype {
    NDL_U8 = 0,
    NDL_I8 = 1,
    NDL_U16 = 2,
    NDL_I16 = 3,
    NDL_I32 = 4,
    NDL_F32 = 5,
    NDL_F64 = 6,
    NDL_F16 = 7,  /* storage-only; not a reference type */
    NDL_BF16 = 8, /* storage-only; not a reference type */
    NDL_DTYPE_COUNT = 9
};

/* BorderMode, image/border_mode.h:14-19 (same order: Clamp, Wrap, Reflect) */
enum ndl_border { NDL_CLAMP = 0, NDL_WRAP = 1, NDL_REFLECT = 2 };

/* accumulation contract */
enum ndl_accum {
    /* The reference's own arithmetic: every product double(src)*double(tap)
     * and every sum rounded separately in FP64, taps in ascending order,
     * FP64 intermediates between axis passes (convolution.h:64-70,151-169).
     * Results are bit-identical to the reference for every element type. */
    NDL_ACCUM_REF = 0,
    /* FP32 fused multiply-add, FP32 intermediates.  Float types only. */
    NDL_ACCUM_F32 = 1
};

/* final narrowing to an integral destination type */
enum ndl_rounding {
    NDL_ROUND_TRUNC = 0,   /* static_cast<DstT>(v): convolve(), convolution.h:69 */
    NDL_ROUND_NEAREST = 1  /* std::round then cast: gaussian_blur() DIM>=2, :184 */
};

/* morphology operation (processing/morphology.h:57-101) */
enum ndl_morph_op {
    NDL_ERODE = 0,  /* erode(): minimum over the structuring element */
    NDL_DILATE = 1  /* dilate(): maximum over the structuring element */
};

enum ndl_status {
    NDL_OK = 0,
    NDL_ERR_ARG = -1,      /* null pointer, bad count, unsupported combination */
    NDL_ERR_DTYPE = -2,    /* unknown dtype, or a dtype the accumulation mode does not allow */
    NDL_ERR_SHAPE = -3,    /* extents differ, ndim out of range, empty axis list */
    NDL_ERR_RADIUS = -4,   /* wrap/reflect tap reach exceeds the axis extent (single-fold maps) */
    NDL_ERR_ALIAS = -5,    /* src and dst memory overlap */
    NDL_ERR_HIP = -6,      /* a HIP runtime call failed */
    NDL_ERR_NOMEM = -7     /* device allocation failed */
};

typedef struct ndl_tensor {
    void* data;                  /* device pointer to element (0,...,0) */
    int32_t dtype;               /* enum ndl_dtype */
    int32_t ndim;                /* 1..NDL_MAX_DIM */
    int64_t extent[NDL_MAX_DIM]; /* element count per dimension */
    int64_t stride[NDL_MAX_DIM]; /* signed, in elements */
} ndl_tensor;

/* One axis of a separable kernel: the {1,..,count,..,1} axis kernel that
 * gaussian_blur() builds (convolution.h:132-139).  Output i along the axis is
 * sum_t taps[t] * src[B(i + t - center)], t ascending; taps that are exactly
 * zero are skipped as detail::kernelIncludedTaps() does (detail.h:164-174).
 * center defaults to count/2 (detail::kernelCenter, detail.h:155-163). */
typedef struct ndl_axis_kernel {
    const double* taps; /* host pointer, `count` values */
    int32_t count;      /* 1..NDL_MAX_TAPS */
    int32_t center;     /* 0..count-1 */
    int32_t border;     /* enum ndl_border */
} ndl_axis_kernel;

/* ---- separable forward -------------------------------------------------
 * Replaces the per-axis pass chain of gaussian_blur() (convolution.h:151-185)
 * and, with naxes == 1, a convolve() call whose kernel has extent > 1 on one
 * axis only (convolution.h:52-71).  Axes are applied in the order given
 * (gaussian_blur applies 0..DIM-1); axes not listed are left untouched, which
 * is what per_channel() slicing gives (image/kernels.h:78-87).  src and dst
 * must have equal extents and must not overlap. */
int ndl_sepconv_forward(const ndl_tensor* src, const ndl_tensor* dst, int32_t naxes,
                        const int32_t* axes, const ndl_axis_kernel* k, int32_t accum,
                        int32_t rounding, void* stream);

/* ---- dense N-D kernel ---------------------------------------------------
 * convolve(src, dst, kernel, border) for an arbitrary (non-separable) N-D
 * kernel, convolution.h:52-71: taps in kernel coordinate order (dimension 0
 * fastest), zero taps skipped, FP64 accumulation, static_cast output.
 * kernel: host pointer to prod(kextent) doubles, dense, dimension 0 fastest.
 * Integer images with an integer rank-1 kernel (a box, outer products of
 * binomials), whose FP64 sums are exact, run as one pass per axis: the same
 * result. */
int ndl_convolve_nd(const ndl_tensor* src, const ndl_tensor* dst, const double* kernel,
                    const int64_t* kextent, int32_t border, void* stream);

/* ---- morphology -----------------------------------------------------------
 * erode(src, dst, kernel, border) / dilate(...), morphology.h:57-101: for each
 * output, the first included tap (kernel coordinate order, dimension 0
 * fastest; included = nonzero, detail.h:164-174) and then every strictly
 * smaller (erode) / larger (dilate) value, border resolved per tap
 * (detail.h:178-194); dst gets static_cast<DstT> of the winner (SrcT{} when
 * no tap is included).  Results are bit-identical to the reference.
 *
 * ndl_morph_separable: structuring elements whose included set is a product
 * of per-axis sets (make_box_kernel, image/kernels.h:22-23): k[i].taps is the
 * inclusion mask of axis axes[i] (nonzero = included), each with its own
 * border and centre; axes not listed have extent 1.  One min/max pass per
 * axis.  Same checks as ndl_sepconv_forward; an axis may be listed once. */
int ndl_morph_separable(const ndl_tensor* src, const ndl_tensor* dst, int32_t naxes,
                        const int32_t* axes, const ndl_axis_kernel* k, int32_t op, void* stream);

/* ndl_morph_nd: any N-D mask (e.g. make_cross_kernel, image/kernels.h:44-58),
 * one pass; mask as ndl_convolve_nd's kernel. */
int ndl_morph_nd(const ndl_tensor* src, const ndl_tensor* dst, const double* mask,
                 const int64_t* kextent, int32_t border, int32_t op, void* stream);

/* ---- scale stacks ------------------------------------------------------------
 * Several gaussian_blur() levels of one image: levels[l] = gaussian_blur(src,
 * sigmas[l], border) over every axis (convolution.h:107-186), the scale stack
 * of detect_keypoints (feature_detection.h:101-109).  accum / rounding as in
 * ndl_sepconv_forward (the C++ layer passes NDL_ACCUM_REF and, for integral
 * images with ndim >= 2, NDL_ROUND_NEAREST: bit-identical to the reference).
 * In NDL_ACCUM_REF the first axis pass of every level reads src once.
 * nlevels <= 16; each level like a dst of ndl_sepconv_forward; levels must not
 * overlap src or each other. */
int ndl_gaussian_scales(const ndl_tensor* src, const ndl_tensor* levels, int32_t nlevels,
                        const double* sigmas, int32_t border, int32_t accum, int32_t rounding,
                        void* stream);

/* Difference-of-Gaussians layers (feature_detection.h:111-120):
 * dogs[l] = levels[l+1] - levels[l] for l < nlevels-1, computed in FP64 and
 * stored as dogs' element type (exact for float64 levels, as the reference's
 * OwnedImage<double> layers).  Equal extents, no overlap with the levels. */
int ndl_dog(const ndl_tensor* levels, const ndl_tensor* dogs, int32_t nlevels, void* stream);

/* ---- gradient --------------------------------------------------------------
 * gradient(src, dst, border), convolution.h:359-388: dst has one more
 * dimension than src, a leading component axis of src->ndim entries, and
 * dst[a, p] = convolve() of src with {-0.5, 0, 0.5} along axis a at p (FP64,
 * the zero tap skipped, truncating cast): bit-identical to the reference's
 * one convolve() per axis, all components in one launch. */
int ndl_gradient(const ndl_tensor* src, const ndl_tensor* dst, int32_t border, void* stream);

/* ---- downsample ------------------------------------------------------------
 * downsample(src, factor, channelAxis), convolution.h:241-305: every axis but
 * channel_axis shrinks to (extent-1)/factor + 1; each output is the FP64 mean
 * of its factor-sized source block clipped at the far edge (block order
 * dimension 0 fastest, divided by the block size), std::round then cast for
 * integral types.  dst must have that extent and src's element type; ndim >= 2.
 * Bit-identical to the reference. */
int ndl_downsample(const ndl_tensor* src, const ndl_tensor* dst, int32_t factor,
                   int32_t channel_axis, void* stream);

/* ---- blur then downsample --------------------------------------------------
 * downsample(gaussian_blur(src, sigma, border), factor, channelAxis) with the
 * blurred image held in dst's element type: the reduce step of a Gaussian
 * pyramid, composed from convolution.h:107-186 and :241-305.  sigma has ndim
 * entries; dimension d is blurred with gaussian_blur()'s taps for sigma[d]
 * (passes in dimension order) unless sigma[d] <= 0; at least one is > 0.  The
 * reference composition is every sigma[d] equal, NDL_ACCUM_REF and, for an
 * integral dst, NDL_ROUND_NEAREST (gaussian_blur's final rounding); then the
 * result is bit-identical to the two calls.  The last blurred pass, that
 * rounding and the block mean run in one kernel: the full-resolution blurred
 * image is never stored.  accum NDL_ACCUM_F32 needs f32/f16/bf16 src and dst.
 * dst extents as for ndl_downsample; src and dst must not overlap. */
int ndl_gaussian_downsample(const ndl_tensor* src, const ndl_tensor* dst, const double* sigma,
                            int32_t border, int32_t factor, int32_t channel_axis, int32_t accum,
                            int32_t rounding, void* stream);

/* ---- adjoint (no counterpart in the reference) -------------------------
 * Gradient of L = <w, sepconv_forward(x)> with respect to x: grad_in =
 * A^T grad_out, the exact transpose of the forward operator including its
 * clamp/wrap/reflect index folding at the borders.  grad_out/grad_in are
 * float (f32/f64/f16/bf16); accum NDL_ACCUM_REF computes in FP64. */
int ndl_sepconv_adjoint_input(const ndl_tensor* grad_out, const ndl_tensor* grad_in,
                              int32_t naxes, const int32_t* axes, const ndl_axis_kernel* k,
                              int32_t accum, void* stream);

/* Gradient of L with respect to every tap: grad_taps (DEVICE pointer to
 * sum(k[a].count) doubles, axis order as given) receives
 * dL/dk_a[t] = sum_p grad_out[p] * U_a[B_a(p + t - c_a)], U_a being x passed
 * through every axis except a.  The reduction is deterministic. */
int ndl_sepconv_adjoint_taps(const ndl_tensor* x, const ndl_tensor* grad_out, int32_t naxes,
                             const int32_t* axes, const ndl_axis_kernel* k, double* grad_taps,
                             int32_t accum, void* stream);

/* ---- allocation-free input adjoint (no counterpart in the reference) -----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The result
 * of ndl_sepconv_adjoint_input(..., NDL_ACCUM_F32, ...) when that adjoint
 * runs as one fused pass: grad_in of grad_out's type (f32/f16/bf16), at most
 * three convolved axes, no zero tap, a unit-stride dimension in both.  The
 * edge weight rows of clamp / asymmetric-reflect axes go to a library-owned
 * device array of NDL_NOALLOC_SCRATCH_BYTES on each device (held until the
 * process ends); nothing is allocated or freed.  Returns NDL_ERR_ARG when the
 * call would need the per-axis chain (use ndl_sepconv_adjoint_input) and
 * NDL_ERR_NOMEM when the rows exceed the array.  Every call on one device
 * shares that array: calls that overlap in time on one device (other
 * streams, other host threads) race. */
#define NDL_NOALLOC_SCRATCH_BYTES (256 * 1024)
int ndl_sepconv_adjoint_input_noalloc(const ndl_tensor* grad_out, const ndl_tensor* grad_in,
                                      int32_t naxes, const int32_t* axes, const ndl_axis_kernel* k,
                                      void* stream);

/* ---- Gaussian taps --------------------------------------------------------
 * The taps gaussian_blur() builds (convolution.h:116-126): radius
 * _kernelSize(sigma) (mathHelpers.h:17-18), w_i = exp(-d^2/(2 sigma^2)),
 * normalised by their ascending sum.  Writes 2*radius+1 values into taps
 * (capacity cap) and returns that count, or a negative NDL_ERR_*. */
int ndl_gaussian_taps(double sigma, double* taps, int32_t cap);

/* ---- batched plans (no counterpart in the reference) --------------------
 * Many independent separable forwards of uneven extents, dimensions, taps
 * and borders.  The plan copies the descriptors and taps, groups the items
 * by storage type, tile shape and ring period, and owns one device job list
 * per group; each group runs as a persistent work-queue kernel, the groups
 * side by side on up to four streams forked from and joined back onto the
 * caller's stream (or, if that measured faster, one launch per item).
 * Items the fused kernel cannot take run through the per-axis chain.
 * The first ndl_plan_execute times every candidate schedule on the
 * caller's stream and waits for those timings before it returns (the only
 * call that blocks the host); it keeps the fastest and captures it into a
 * HIP graph, and later calls only launch that graph, asynchronously. */
typedef struct ndl_batch_item {
    ndl_tensor src;
    ndl_tensor dst;
    int32_t naxes;
    int32_t axes[NDL_MAX_DIM];
    ndl_axis_kernel k[NDL_MAX_DIM];
} ndl_batch_item;
typedef struct ndl_plan ndl_plan;
int ndl_plan_create(const ndl_batch_item* items, int32_t n, int32_t accum, ndl_plan** plan);
int ndl_plan_execute(ndl_plan* plan, void* stream);
int ndl_plan_destroy(ndl_plan* plan);

/* ---- device memory and stream helpers used by the C++ layer -------------- */
int ndl_malloc(void** ptr, size_t bytes);
int ndl_free(void* ptr);
int ndl_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int ndl_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int ndl_memset(void* dst, int value, size_t bytes, void* stream);
int ndl_stream_synchronize(void* stream);
/* Page-locked host memory (hipHostMalloc) for staging files into device
 * images (include/ndl/image_io.h); no counterpart in the reference. */
int ndl_host_alloc(void** ptr, size_t bytes);
int ndl_host_free(void* ptr);
int ndl_device_count(int* count);

const char* ndl_last_error(void); /* thread-local; never NULL */
int ndl_version(void);            /* major*10000 + minor*100 + patch */

#ifdef __cplusplus
}
#endif
#endif /* NDL_GPU_H */
