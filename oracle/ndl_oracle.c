/*
 * ndl_oracle.c -- CPU restatement of ndl's separable convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load this library, and only as the checker
 * or as the timed CPU baseline -- never as the product path.
 *
 * Build: oracle/Makefile (gcc -O3 -std=c11 -ffp-contract=off), so every
 * product and sum is rounded separately in FP64 exactly as the reference's
 * default build does (no FMA on baseline x86-64, CMakeLists.txt:86).
 *
 * What each function follows (reference file:line):
 *   oracle_kernel_size      mathHelpers.h:17-18   (_ceil, _kernelSize)
 *   oracle_gaussian_taps    processing/convolution.h:116-126
 *   border_map              mathHelpers.h:23-25 via image/detail.h:178-194
 *   oracle_convolve         processing/convolution.h:52-71, detail.h:155-194,
 *                           image/core.h:557-558 (at = stride dot product),
 *                           detail.h:127-143 (coordinate order, dim 0 fastest)
 *   oracle_sepconv          processing/convolution.h:141-185 (per-axis passes
 *                           through double buffers, final round/cast)
 *   cvtt_i32, store_elem    convolution.h:69 static_cast<DstT>(double): x86-64
 *                           cvttsd2si to int32 then the low bits (UB in C++
 *                           for out-of-range values; this is what the
 *                           reference's compiled code does)
 *   oracle_sepconv_adjoint_input / _taps: no reference counterpart; the exact
 *                           transpose / tap derivative of oracle_sepconv in FP64.
 *   oracle_downsample       processing/convolution.h:241-305 (block mean:
 *                           every axis but channelAxis shrinks to
 *                           (e-1)/factor+1; each output sums its clipped
 *                           source block in FP64, block coordinates
 *                           dimension 0 fastest, divides by the block size,
 *                           std::round for integral types, static_cast)
 *   oracle_morph            processing/morphology.h:57-101 (erode / dilate:
 *                           first included tap in kernel-coordinate order,
 *                           then strictly smaller / larger values replace it;
 *                           static_cast<DstT> of the winner), same tap list
 *                           and border resolution as oracle_convolve.
 *
 * How each is pinned (tests/test_oracle.py): the reference unit-test and
 * tutorial golden values (image_convolution_tests.cpp:13-67,
 * composite_tests.cpp:105-133, type_coverage_tests.cpp:141-149,
 * convolution_tutorial.md:61-66,83-85,138-149), brute force and
 * scipy.ndimage.correlate1d on small inputs, the FFT circular-correlation
 * identity (fft_tests.cpp:164-223), dense-Jacobian and dot-product identities
 * for the adjoints, and -- when the reference sources are present -- the
 * compiled reference itself (oracle/_ref).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ndl_gpu.h"

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- threads
 * The passes below compute every output element independently (gathers) or
 * one line at a time in the serial order (the transposed scatter), so
 * splitting their index range over threads changes no result bit.  Default
 * one thread: the timed CPU baseline is the reference's serial loop. */
static int g_threads = 1;

EXPORT void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : (n > 256 ? 256 : n); }
EXPORT int oracle_get_threads(void) { return g_threads; }

typedef void (*range_fn)(void* ctx, int64_t lo, int64_t hi);
struct range_job { range_fn fn; void* ctx; int64_t lo, hi; };

static void* range_thread(void* p)
{
    struct range_job* j = (struct range_job*)p;
    j->fn(j->ctx, j->lo, j->hi);
    return NULL;
}

/* fn(ctx, lo, hi) over [0, n) in g_threads contiguous ranges */
static void run_ranges(int64_t n, range_fn fn, void* ctx)
{
    int nt = g_threads;
    if (nt > 1 && n < 4096) nt = 1;
    if (nt <= 1) {
        fn(ctx, 0, n);
        return;
    }
    struct range_job jobs[256];
    pthread_t tid[256];
    int started[256];
    for (int i = 0; i < nt; i++) {
        jobs[i].fn = fn;
        jobs[i].ctx = ctx;
        jobs[i].lo = n * i / nt;
        jobs[i].hi = n * (i + 1) / nt;
        started[i] = i > 0 && pthread_create(&tid[i], NULL, range_thread, &jobs[i]) == 0;
    }
    for (int i = 0; i < nt; i++)
        if (!started[i]) fn(ctx, jobs[i].lo, jobs[i].hi);
    for (int i = 1; i < nt; i++)
        if (started[i]) pthread_join(tid[i], NULL);
}

/* coordinates (dimension 0 fastest) of linear index lin */
static void unravel(int64_t lin, int nd, const int64_t* extent, int64_t* coord)
{
    for (int d = 0; d < nd; d++) {
        coord[d] = extent[d] ? lin % extent[d] : 0;
        lin = extent[d] ? lin / extent[d] : 0;
    }
}

/* ---------------------------------------------------------------- scalars */

static double half_to_double(uint16_t h)
{
    uint32_t sign = (h >> 15) & 1u, exp = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    double v;
    if (exp == 0)
        v = ldexp((double)man, -24);
    else if (exp == 31)
        v = man ? NAN : INFINITY;
    else
        v = ldexp((double)(man | 0x400u), (int)exp - 25);
    return sign ? -v : v;
}

static double bf16_to_double(uint16_t b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return (double)f;
}

/* round-to-nearest-even double -> binary16 */
static uint16_t double_to_half(double v)
{
    uint16_t sign = signbit(v) ? 0x8000u : 0;
    double a = fabs(v);
    if (isnan(v)) return 0x7e00u;
    if (a >= 65520.0) return sign | 0x7c00u; /* rounds to inf */
    if (a < ldexp(1.0, -14)) {               /* subnormal: units of 2^-24 */
        double q = nearbyint(a * ldexp(1.0, 24));
        return sign | (uint16_t)q;           /* q==1024 carries into the exponent */
    }
    int e;
    frexp(a, &e); /* a = m*2^e, m in [0.5,1) */
    int E = e - 1; /* a in [2^E, 2^(E+1)) */
    double q = nearbyint(ldexp(a, 10 - E)); /* 1024..2048 */
    if (q >= 2048.0) { q /= 2; E += 1; }
    if (E > 15) return sign | 0x7c00u;
    return sign | (uint16_t)((E + 15) << 10) | (uint16_t)((uint32_t)q & 0x3ffu);
}

/* round-to-nearest-even double -> bfloat16 (via the exactly rounded float) */
static uint16_t double_to_bf16(double v)
{
    float f = (float)v; /* RNE to float first: a rare double rounding is possible,
                           harmless where the oracle stores bf16 (error statistics) */
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

static double load_elem(const void* base, int dt, int64_t off)
{
    switch (dt) {
    case NDL_U8: return (double)((const uint8_t*)base)[off];
    case NDL_I8: return (double)((const int8_t*)base)[off];
    case NDL_U16: return (double)((const uint16_t*)base)[off];
    case NDL_I16: return (double)((const int16_t*)base)[off];
    case NDL_I32: return (double)((const int32_t*)base)[off];
    case NDL_F32: return (double)((const float*)base)[off];
    case NDL_F64: return ((const double*)base)[off];
    case NDL_F16: return half_to_double(((const uint16_t*)base)[off]);
    case NDL_BF16: return bf16_to_double(((const uint16_t*)base)[off]);
    }
    return NAN;
}

/* static_cast<integral DstT>(double) as compiled for x86-64: cvttsd2si into a
 * 32-bit register (NaN and out-of-range give INT32_MIN), then the low bits. */
static int32_t cvtt_i32(double v)
{
    if (v > -2147483649.0 && v < 2147483648.0) return (int32_t)v;
    return INT32_MIN;
}

static int is_integral(int dt) { return dt <= NDL_I32; }

static void store_elem(void* base, int dt, int64_t off, double v, int rounding)
{
    if (is_integral(dt) && rounding == NDL_ROUND_NEAREST) v = round(v);
    switch (dt) {
    case NDL_U8: ((uint8_t*)base)[off] = (uint8_t)(uint32_t)cvtt_i32(v); break;
    case NDL_I8: ((int8_t*)base)[off] = (int8_t)(uint8_t)(uint32_t)cvtt_i32(v); break;
    case NDL_U16: ((uint16_t*)base)[off] = (uint16_t)(uint32_t)cvtt_i32(v); break;
    case NDL_I16: ((int16_t*)base)[off] = (int16_t)(uint16_t)(uint32_t)cvtt_i32(v); break;
    case NDL_I32: ((int32_t*)base)[off] = cvtt_i32(v); break;
    case NDL_F32: ((float*)base)[off] = (float)v; break;
    case NDL_F64: ((double*)base)[off] = v; break;
    case NDL_F16: ((uint16_t*)base)[off] = double_to_half(v); break;
    case NDL_BF16: ((uint16_t*)base)[off] = double_to_bf16(v); break;
    }
}

/* _clamp/_wrap/_reflect, mathHelpers.h:23-25 (single fold each) */
static int64_t border_map(int border, int64_t m, int64_t x)
{
    switch (border) {
    case NDL_WRAP: return x < 0 ? x + m : (x >= m ? x - m : x);
    case NDL_REFLECT: return x < 0 ? -x - 1 : (x >= m ? 2 * m - x - 1 : x);
    default: return x < 0 ? 0 : (x >= m ? m - 1 : x);
    }
}

/* The single-fold maps are only defined for -m <= x < 2m; the reference reads
 * out of bounds beyond that, the oracle refuses. */
static int reach_ok(int border, int64_t m, int64_t lo_reach, int64_t hi_reach)
{
    if (border == NDL_CLAMP) return 1;
    return lo_reach <= m && hi_reach <= m;
}

/* ------------------------------------------------------------- exported */

EXPORT int oracle_kernel_size(double sigma)
{
    /* _kernelSize: (int)_ceil(3.0f * sigma); 3.0f*sigma is a double product */
    double num = 3.0 * sigma;
    int t = (int)num;
    if ((double)t == num) return t;
    return t + (num > 0 ? 1 : 0);
}

EXPORT int oracle_gaussian_taps(double sigma, double* taps, int cap)
{
    if (!(sigma > 0)) return NDL_ERR_ARG;
    int radius = oracle_kernel_size(sigma);
    int width = 2 * radius + 1;
    if (width > cap) return NDL_ERR_ARG;
    double wsum = 0;
    for (int i = 0; i < width; i++) {
        double d = i - radius;
        taps[i] = exp(-(d * d) / (2 * sigma * sigma));
        wsum += taps[i];
    }
    for (int i = 0; i < width; i++) taps[i] /= wsum;
    return width;
}

static int64_t numel(const ndl_tensor* t)
{
    int64_t n = 1;
    for (int d = 0; d < t->ndim; d++) n *= t->extent[d];
    return n;
}

/* convolve(src, dst, kernel, border), convolution.h:52-71 */
EXPORT int oracle_convolve(const ndl_tensor* src, const ndl_tensor* dst, const double* kernel,
                           const int64_t* kextent, int border)
{
    int nd = src->ndim;
    if (nd < 1 || nd > NDL_MAX_DIM || dst->ndim != nd) return NDL_ERR_SHAPE;
    for (int d = 0; d < nd; d++)
        if (src->extent[d] != dst->extent[d] || kextent[d] < 1) return NDL_ERR_SHAPE;
    int64_t center[NDL_MAX_DIM], kn = 1;
    for (int d = 0; d < nd; d++) {
        center[d] = kextent[d] / 2; /* detail::kernelCenter */
        if (!reach_ok(border, src->extent[d], center[d], kextent[d] - 1 - center[d]))
            return NDL_ERR_RADIUS;
        kn *= kextent[d];
    }
    /* detail::kernelIncludedTaps: kernel coordinates dim 0 fastest, nonzero only */
    int64_t* tapc = (int64_t*)malloc(sizeof(int64_t) * (size_t)(kn * nd + 1));
    double* tapv = (double*)malloc(sizeof(double) * (size_t)(kn + 1));
    int64_t ntaps = 0;
    for (int64_t lin = 0; lin < kn; lin++) {
        if (kernel[lin] == 0.0) continue;
        int64_t r = lin;
        for (int d = 0; d < nd; d++) {
            tapc[ntaps * nd + d] = r % kextent[d] - center[d];
            r /= kextent[d];
        }
        tapv[ntaps++] = kernel[lin];
    }
    int64_t n = numel(src), coord[NDL_MAX_DIM] = {0};
    for (int64_t lin = 0; lin < n; lin++) {
        double total = 0;
        for (int64_t t = 0; t < ntaps; t++) {
            int64_t off = 0;
            for (int d = 0; d < nd; d++) {
                int64_t x = border_map(border, src->extent[d], coord[d] + tapc[t * nd + d]);
                off += x * src->stride[d];
            }
            total += load_elem(src->data, src->dtype, off) * tapv[t];
        }
        int64_t doff = 0;
        for (int d = 0; d < nd; d++) doff += coord[d] * dst->stride[d];
        store_elem(dst->data, dst->dtype, doff, total, NDL_ROUND_TRUNC);
        for (int d = 0; d < nd; d++) { /* dim 0 fastest */
            if (++coord[d] < src->extent[d]) break;
            coord[d] = 0;
        }
    }
    free(tapc);
    free(tapv);
    return NDL_OK;
}

/* one axis pass: out = sum_t k[t] * in[B(i + t - c)], FP64, ascending nonzero taps */
struct pass_ctx {
    const void* in; int in_dt; const int64_t* in_stride;
    void* out; int out_dt; const int64_t* out_stride;
    int nd; const int64_t* extent; int axis; const ndl_axis_kernel* k; int rounding;
};

static void axis_pass_range(void* p, int64_t lo, int64_t hi)
{
    const struct pass_ctx* c = (const struct pass_ctx*)p;
    const int nd = c->nd, axis = c->axis;
    const ndl_axis_kernel* k = c->k;
    int64_t m = c->extent[axis];
    int64_t coord[NDL_MAX_DIM] = {0};
    unravel(lo, nd, c->extent, coord);
    for (int64_t lin = lo; lin < hi; lin++) {
        int64_t base = 0, doff = 0;
        for (int d = 0; d < nd; d++) {
            if (d != axis) base += coord[d] * c->in_stride[d];
            doff += coord[d] * c->out_stride[d];
        }
        double total = 0;
        for (int t = 0; t < k->count; t++) {
            if (k->taps[t] == 0.0) continue;
            int64_t x = border_map(k->border, m, coord[axis] + t - k->center);
            total += load_elem(c->in, c->in_dt, base + x * c->in_stride[axis]) * k->taps[t];
        }
        store_elem(c->out, c->out_dt, doff, total, c->rounding);
        for (int d = 0; d < nd; d++) {
            if (++coord[d] < c->extent[d]) break;
            coord[d] = 0;
        }
    }
}

static void axis_pass(const void* in, int in_dt, const int64_t* in_stride, void* out, int out_dt,
                      const int64_t* out_stride, int nd, const int64_t* extent, int axis,
                      const ndl_axis_kernel* k, int rounding)
{
    struct pass_ctx c = {in, in_dt, in_stride, out, out_dt, out_stride, nd, extent, axis, k, rounding};
    int64_t n = 1;
    for (int d = 0; d < nd; d++) n *= extent[d];
    run_ranges(n, axis_pass_range, &c);
}

static int check_axes(const ndl_tensor* t, int naxes, const int32_t* axes, const ndl_axis_kernel* k)
{
    if (naxes < 0 || naxes > NDL_MAX_DIM) return NDL_ERR_ARG;
    for (int a = 0; a < naxes; a++) {
        if (axes[a] < 0 || axes[a] >= t->ndim) return NDL_ERR_SHAPE;
        if (k[a].count < 1 || k[a].center < 0 || k[a].center >= k[a].count) return NDL_ERR_ARG;
        if (k[a].border < 0 || k[a].border > 2) return NDL_ERR_ARG;
        if (!reach_ok(k[a].border, t->extent[axes[a]], k[a].center, k[a].count - 1 - k[a].center))
            return NDL_ERR_RADIUS;
    }
    return NDL_OK;
}

static void dense_stride(int nd, const int64_t* extent, int64_t* stride)
{
    stride[0] = 1;
    for (int d = 1; d < nd; d++) stride[d] = stride[d - 1] * extent[d - 1];
}

/* gaussian_blur's pass chain, convolution.h:141-185, with per-axis taps and
 * borders: pass 0 reads src, every pass writes a double buffer, the last pass
 * casts into dst (a single pass casts straight into dst, as DIM==1 does). */
EXPORT int oracle_sepconv(const ndl_tensor* src, const ndl_tensor* dst, int naxes,
                          const int32_t* axes, const ndl_axis_kernel* k, int rounding)
{
    int nd = src->ndim;
    if (nd < 1 || nd > NDL_MAX_DIM || dst->ndim != nd) return NDL_ERR_SHAPE;
    for (int d = 0; d < nd; d++)
        if (src->extent[d] != dst->extent[d]) return NDL_ERR_SHAPE;
    int rc = check_axes(src, naxes, axes, k);
    if (rc) return rc;
    if (naxes < 1) return NDL_ERR_ARG;
    int64_t n = numel(src), ds[NDL_MAX_DIM];
    dense_stride(nd, src->extent, ds);
    double* buf[2] = {NULL, NULL};
    if (naxes > 1) {
        buf[0] = (double*)malloc(sizeof(double) * (size_t)n);
        buf[1] = (double*)malloc(sizeof(double) * (size_t)n);
        if (!buf[0] || !buf[1]) { free(buf[0]); free(buf[1]); return NDL_ERR_NOMEM; }
    }
    const void* cur = src->data;
    int cur_dt = src->dtype;
    const int64_t* cur_stride = src->stride;
    for (int a = 0; a < naxes; a++) {
        int last = a == naxes - 1;
        void* out = last ? dst->data : (void*)buf[a & 1];
        int out_dt = last ? dst->dtype : NDL_F64;
        const int64_t* out_stride = last ? dst->stride : ds;
        axis_pass(cur, cur_dt, cur_stride, out, out_dt, out_stride, nd, src->extent, axes[a], &k[a],
                  last ? rounding : NDL_ROUND_TRUNC);
        cur = out;
        cur_dt = NDL_F64;
        cur_stride = ds;
    }
    free(buf[0]);
    free(buf[1]);
    return NDL_OK;
}

/* transpose of one FP64 axis pass: out[B(i+t-c)] += k[t] * in[i] (dense
 * buffers).  Every scatter stays inside its line along the axis, and a line
 * is processed in ascending position order, so threads split the lines. */
struct tpass_ctx {
    const double* in; double* out; int nd; const int64_t* extent; const int64_t* stride;
    int axis; const ndl_axis_kernel* k;
};

static void axis_pass_transpose_lines(void* p, int64_t lo, int64_t hi)
{
    const struct tpass_ctx* c = (const struct tpass_ctx*)p;
    const int nd = c->nd, axis = c->axis;
    const ndl_axis_kernel* k = c->k;
    int64_t m = c->extent[axis], ext[NDL_MAX_DIM], coord[NDL_MAX_DIM] = {0};
    int od = 0;
    for (int d = 0; d < nd; d++)
        if (d != axis) ext[od++] = c->extent[d];
    unravel(lo, od, ext, coord);
    for (int64_t line = lo; line < hi; line++) {
        int64_t base = 0;
        for (int d = 0, j = 0; d < nd; d++)
            if (d != axis) base += coord[j++] * c->stride[d];
        for (int64_t i = 0; i < m; i++) {
            double g = c->in[base + i * c->stride[axis]];
            for (int t = 0; t < k->count; t++) {
                if (k->taps[t] == 0.0) continue;
                int64_t x = border_map(k->border, m, i + t - k->center);
                c->out[base + x * c->stride[axis]] += k->taps[t] * g;
            }
        }
        for (int d = 0; d < od; d++) {
            if (++coord[d] < ext[d]) break;
            coord[d] = 0;
        }
    }
}

static void axis_pass_transpose(const double* in, double* out, int nd, const int64_t* extent,
                                const int64_t* stride, int axis, const ndl_axis_kernel* k)
{
    int64_t n = 1;
    for (int d = 0; d < nd; d++) n *= extent[d];
    memset(out, 0, sizeof(double) * (size_t)n);
    struct tpass_ctx c = {in, out, nd, extent, stride, axis, k};
    run_ranges(extent[axis] ? n / extent[axis] : 0, axis_pass_transpose_lines, &c);
}

static int gather_dense(const ndl_tensor* t, double* out)
{
    int64_t n = numel(t), coord[NDL_MAX_DIM] = {0};
    for (int64_t lin = 0; lin < n; lin++) {
        int64_t off = 0;
        for (int d = 0; d < t->ndim; d++) off += coord[d] * t->stride[d];
        out[lin] = load_elem(t->data, t->dtype, off);
        for (int d = 0; d < t->ndim; d++) {
            if (++coord[d] < t->extent[d]) break;
            coord[d] = 0;
        }
    }
    return NDL_OK;
}

static void scatter_dense(const double* in, const ndl_tensor* t)
{
    int64_t n = numel(t), coord[NDL_MAX_DIM] = {0};
    for (int64_t lin = 0; lin < n; lin++) {
        int64_t off = 0;
        for (int d = 0; d < t->ndim; d++) off += coord[d] * t->stride[d];
        store_elem(t->data, t->dtype, off, in[lin], NDL_ROUND_TRUNC);
        for (int d = 0; d < t->ndim; d++) {
            if (++coord[d] < t->extent[d]) break;
            coord[d] = 0;
        }
    }
}

/* grad_in = A^T grad_out, A = A_{n-1} ... A_0 (FP64 throughout) */
EXPORT int oracle_sepconv_adjoint_input(const ndl_tensor* g, const ndl_tensor* gx, int naxes,
                                        const int32_t* axes, const ndl_axis_kernel* k)
{
    int nd = g->ndim;
    if (nd < 1 || nd > NDL_MAX_DIM || gx->ndim != nd) return NDL_ERR_SHAPE;
    for (int d = 0; d < nd; d++)
        if (g->extent[d] != gx->extent[d]) return NDL_ERR_SHAPE;
    int rc = check_axes(g, naxes, axes, k);
    if (rc) return rc;
    int64_t n = numel(g), ds[NDL_MAX_DIM];
    dense_stride(nd, g->extent, ds);
    double* a = (double*)malloc(sizeof(double) * (size_t)n);
    double* b = (double*)malloc(sizeof(double) * (size_t)n);
    if (!a || !b) { free(a); free(b); return NDL_ERR_NOMEM; }
    gather_dense(g, a);
    for (int i = naxes - 1; i >= 0; i--) {
        axis_pass_transpose(a, b, nd, g->extent, ds, axes[i], &k[i]);
        double* t = a; a = b; b = t;
    }
    scatter_dense(a, gx);
    free(a);
    free(b);
    return NDL_OK;
}

/* one tap's sum, in the serial position order (dimension 0 fastest) */
struct tap_ctx {
    const double* gd; const double* u; int nd; const int64_t* extent; const int64_t* ds; int ax;
    const ndl_axis_kernel* k; double* out;
};

static void tap_sum_one(void* p, int64_t t0, int64_t t1)
{
    const struct tap_ctx* c = (const struct tap_ctx*)p;
    int64_t n = 1;
    for (int d = 0; d < c->nd; d++) n *= c->extent[d];
    const int ax = c->ax;
    const int64_t m = c->extent[ax];
    for (int64_t t = t0; t < t1; t++) {
        double acc = 0;
        int64_t coord[NDL_MAX_DIM] = {0};
        for (int64_t lin = 0; lin < n; lin++) {
            int64_t base = lin - coord[ax] * c->ds[ax];
            int64_t xi = border_map(c->k->border, m, coord[ax] + t - c->k->center);
            acc += c->gd[lin] * c->u[base + xi * c->ds[ax]];
            for (int d = 0; d < c->nd; d++) {
                if (++coord[d] < c->extent[d]) break;
                coord[d] = 0;
            }
        }
        c->out[t] = acc;
    }
}

/* fn(ctx, i, i+1) for every i < n, one item per thread (up to g_threads) */
static void run_ranges_each(int64_t n, range_fn fn, void* ctx)
{
    int saved = g_threads;
    int64_t done = 0;
    while (done < n) {
        int64_t chunk = n - done < saved ? n - done : saved;
        struct range_job jobs[256];
        pthread_t tid[256];
        int started[256];
        for (int64_t i = 0; i < chunk; i++) {
            jobs[i].fn = fn;
            jobs[i].ctx = ctx;
            jobs[i].lo = done + i;
            jobs[i].hi = done + i + 1;
            started[i] = i > 0 && pthread_create(&tid[i], NULL, range_thread, &jobs[i]) == 0;
        }
        for (int64_t i = 0; i < chunk; i++)
            if (!started[i]) fn(ctx, jobs[i].lo, jobs[i].hi);
        for (int64_t i = 1; i < chunk; i++)
            if (started[i]) pthread_join(tid[i], NULL);
        done += chunk;
    }
}

/* dL/dk_a[t] = sum_p g[p] * U_a[B_a(p + t - c_a)], U_a = x through every other
 * axis (FP64, axis order as given), p in coordinate order. */
EXPORT int oracle_sepconv_adjoint_taps(const ndl_tensor* x, const ndl_tensor* g, int naxes,
                                       const int32_t* axes, const ndl_axis_kernel* k, double* grad)
{
    int nd = x->ndim;
    if (nd < 1 || nd > NDL_MAX_DIM || g->ndim != nd) return NDL_ERR_SHAPE;
    for (int d = 0; d < nd; d++)
        if (g->extent[d] != x->extent[d]) return NDL_ERR_SHAPE;
    int rc = check_axes(x, naxes, axes, k);
    if (rc) return rc;
    int64_t n = numel(x), ds[NDL_MAX_DIM];
    dense_stride(nd, x->extent, ds);
    double* u = (double*)malloc(sizeof(double) * (size_t)n);
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    double* gd = (double*)malloc(sizeof(double) * (size_t)n);
    if (!u || !v || !gd) { free(u); free(v); free(gd); return NDL_ERR_NOMEM; }
    gather_dense(g, gd);
    int off = 0;
    for (int a = 0; a < naxes; a++) {
        gather_dense(x, u);
        for (int b = 0; b < naxes; b++) {
            if (b == a) continue;
            axis_pass(u, NDL_F64, ds, v, NDL_F64, ds, nd, x->extent, axes[b], &k[b], NDL_ROUND_TRUNC);
            double* t = u; u = v; v = t;
        }
        int ax = axes[a];
        int64_t m = x->extent[ax];
        struct tap_ctx c = {gd, u, nd, x->extent, ds, ax, &k[a], grad + off};
        run_ranges_each(k[a].count, tap_sum_one, &c);
        off += k[a].count;
    }
    free(u);
    free(v);
    free(gd);
    return NDL_OK;
}

/* element codecs, exposed so tests can check the storage conversions */
/* erode (op 0) / dilate (op 1), morphology.h:57-101.  Values compare as
 * doubles: every storage type converts exactly and in order, NaN and signed
 * zeros included, so `v < best` is the reference's comparison. */
EXPORT int oracle_morph(const ndl_tensor* src, const ndl_tensor* dst, const double* kernel,
                        const int64_t* kextent, int border, int op)
{
    int nd = src->ndim;
    if (nd < 1 || nd > NDL_MAX_DIM || dst->ndim != nd) return NDL_ERR_SHAPE;
    if (op != 0 && op != 1) return NDL_ERR_ARG;
    int64_t center[NDL_MAX_DIM], kn = 1;
    for (int d = 0; d < nd; d++) {
        if (src->extent[d] != dst->extent[d] || kextent[d] < 1) return NDL_ERR_SHAPE;
        center[d] = kextent[d] / 2;
        if (!reach_ok(border, src->extent[d], center[d], kextent[d] - 1 - center[d])) return NDL_ERR_RADIUS;
        kn *= kextent[d];
    }
    int64_t* tapc = (int64_t*)malloc(sizeof(int64_t) * (size_t)(kn * nd + 1));
    int64_t ntaps = 0;
    for (int64_t lin = 0; lin < kn; lin++) {
        if (kernel[lin] == 0.0) continue;
        int64_t r = lin;
        for (int d = 0; d < nd; d++) {
            tapc[ntaps * nd + d] = r % kextent[d] - center[d];
            r /= kextent[d];
        }
        ntaps++;
    }
    int64_t n = numel(src), coord[NDL_MAX_DIM] = {0};
    for (int64_t lin = 0; lin < n; lin++) {
        double best = 0; /* SrcT{} when no tap is included */
        for (int64_t t = 0; t < ntaps; t++) {
            int64_t off = 0;
            for (int d = 0; d < nd; d++)
                off += border_map(border, src->extent[d], coord[d] + tapc[t * nd + d]) * src->stride[d];
            double v = load_elem(src->data, src->dtype, off);
            if (t == 0 || (op == 0 ? v < best : v > best)) best = v;
        }
        int64_t doff = 0;
        for (int d = 0; d < nd; d++) doff += coord[d] * dst->stride[d];
        store_elem(dst->data, dst->dtype, doff, best, NDL_ROUND_TRUNC);
        for (int d = 0; d < nd; d++) {
            if (++coord[d] < src->extent[d]) break;
            coord[d] = 0;
        }
    }
    free(tapc);
    return NDL_OK;
}

EXPORT int oracle_downsample(const ndl_tensor* src, const ndl_tensor* dst, int factor, int channel_axis)
{
    int nd = src->ndim;
    if (nd < 2 || nd > NDL_MAX_DIM || dst->ndim != nd || factor < 2 || channel_axis < 0 || channel_axis >= nd)
        return NDL_ERR_ARG;
    for (int d = 0; d < nd; d++) {
        int64_t want = d == channel_axis ? src->extent[d] : (src->extent[d] - 1) / factor + 1;
        if (src->extent[d] < 1 || dst->extent[d] != want) return NDL_ERR_SHAPE;
    }
    int64_t n = numel(dst), coord[NDL_MAX_DIM] = {0};
    for (int64_t lin = 0; lin < n; lin++) {
        int64_t lo[NDL_MAX_DIM], hi[NDL_MAX_DIM], cur[NDL_MAX_DIM], count = 1;
        for (int d = 0; d < nd; d++) {
            if (d == channel_axis) {
                lo[d] = coord[d];
                hi[d] = coord[d] + 1;
            } else {
                lo[d] = coord[d] * factor;
                hi[d] = lo[d] + factor < src->extent[d] ? lo[d] + factor : src->extent[d];
            }
            count *= hi[d] - lo[d];
            cur[d] = lo[d];
        }
        double sum = 0;
        for (int64_t k = 0; k < count; k++) {
            int64_t off = 0;
            for (int d = 0; d < nd; d++) off += cur[d] * src->stride[d];
            sum += load_elem(src->data, src->dtype, off);
            for (int d = 0; d < nd; d++) { /* block odometer, dimension 0 fastest */
                if (d == channel_axis) continue;
                if (++cur[d] < hi[d]) break;
                cur[d] = lo[d];
            }
        }
        double avg = sum / (double)count;
        int64_t doff = 0;
        for (int d = 0; d < nd; d++) doff += coord[d] * dst->stride[d];
        store_elem(dst->data, dst->dtype, doff, avg, NDL_ROUND_NEAREST);
        for (int d = 0; d < nd; d++) {
            if (++coord[d] < dst->extent[d]) break;
            coord[d] = 0;
        }
    }
    return NDL_OK;
}

EXPORT double oracle_half_to_double(uint16_t h) { return half_to_double(h); }
EXPORT uint16_t oracle_double_to_half(double v) { return double_to_half(v); }
EXPORT double oracle_bf16_to_double(uint16_t b) { return bf16_to_double(b); }
EXPORT uint16_t oracle_double_to_bf16(double v) { return double_to_bf16(v); }
EXPORT int32_t oracle_cvtt_i32(double v) { return cvtt_i32(v); }
