"""ndl_amd -- separable N-D convolution on one MI355X.

Python binding over the C ABI of libndl_gpu.so (include/ndl_gpu.h), used by
the parity tests and the benchmark.  The drop-in for C++ users of ndl is the
header-only layer in include/ndl/; this module mirrors the same operations
for numpy arrays and device buffers:

    gaussian_blur(img, sigma, border)          ndl::gaussian_blur, convolution.h:107-186
    convolve(img, kernel, border)              ndl::convolve, convolution.h:52-71
    sepconv(img, axes, taps, borders, ...)     per-axis pass chain (FP64 or FP32)
    adjoint_input / adjoint_taps               gradients of <w, sepconv(x)>

Arrays follow ndl_amd.abi: numpy shape is the reversed ndl extent, and every
`axes` argument uses ndl axis numbers (0 = x, the fastest dimension).

There is no CPU fallback: if libndl_gpu.so is missing or no GPU is present,
the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libndl_gpu.so")

_P = C.POINTER
_T = _P(abi.Tensor)
_K = _P(abi.AxisKernel)
_I32P = _P(C.c_int32)
_I64P = _P(C.c_int64)
_DP = _P(C.c_double)
_VP = C.c_void_p

# every entry point declared in include/ndl_gpu.h, with its ctypes signature
SIGNATURES = {
    "ndl_sepconv_forward": (C.c_int, [_T, _T, C.c_int32, _I32P, _K, C.c_int32, C.c_int32, _VP]),
    "ndl_convolve_nd": (C.c_int, [_T, _T, _DP, _I64P, C.c_int32, _VP]),
    "ndl_morph_separable": (C.c_int, [_T, _T, C.c_int32, _I32P, _K, C.c_int32, _VP]),
    "ndl_morph_nd": (C.c_int, [_T, _T, _DP, _I64P, C.c_int32, C.c_int32, _VP]),
    "ndl_gradient": (C.c_int, [_T, _T, C.c_int32, _VP]),
    "ndl_downsample": (C.c_int, [_T, _T, C.c_int32, C.c_int32, _VP]),
    "ndl_gaussian_downsample": (C.c_int, [_T, _T, _DP, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                          _VP]),
    "ndl_gaussian_scales": (C.c_int, [_T, _T, C.c_int32, _DP, C.c_int32, C.c_int32, C.c_int32, _VP]),
    "ndl_dog": (C.c_int, [_T, _T, C.c_int32, _VP]),
    "ndl_sepconv_adjoint_input": (C.c_int, [_T, _T, C.c_int32, _I32P, _K, C.c_int32, _VP]),
    "ndl_sepconv_adjoint_input_noalloc": (C.c_int, [_T, _T, C.c_int32, _I32P, _K, _VP]),
    "ndl_sepconv_adjoint_taps": (C.c_int, [_T, _T, C.c_int32, _I32P, _K, _VP, C.c_int32, _VP]),
    "ndl_gaussian_taps": (C.c_int, [C.c_double, _DP, C.c_int32]),
    "ndl_plan_create": (C.c_int, [_P(abi.BatchItem), C.c_int32, C.c_int32, _P(_VP)]),
    "ndl_plan_execute": (C.c_int, [_VP, _VP]),
    "ndl_plan_destroy": (C.c_int, [_VP]),
    "ndl_malloc": (C.c_int, [_P(_VP), C.c_size_t]),
    "ndl_free": (C.c_int, [_VP]),
    "ndl_memcpy_htod": (C.c_int, [_VP, _VP, C.c_size_t, _VP]),
    "ndl_memcpy_dtoh": (C.c_int, [_VP, _VP, C.c_size_t, _VP]),
    "ndl_memset": (C.c_int, [_VP, C.c_int, C.c_size_t, _VP]),
    "ndl_stream_synchronize": (C.c_int, [_VP]),
    "ndl_host_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "ndl_host_free": (C.c_int, [_VP]),
    "ndl_device_count": (C.c_int, [_P(C.c_int)]),
    "ndl_last_error": (C.c_char_p, []),
    "ndl_version": (C.c_int, []),
}

_lib = None


def lib():
    """The loaded libndl_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ndl_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class NdlError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"ndl status {rc}: {msg}")
        self.rc = rc


def check(rc):
    if rc < 0:
        raise NdlError(rc, lib().ndl_last_error().decode())
    return rc


# ------------------------------------------------------------- device memory
class DeviceArray:
    """A device buffer with numpy-style shape/strides (in elements)."""

    def __init__(self, shape, dtype, bf16=False, strides=None, _base=None, _offset=0):
        self.shape = tuple(int(s) for s in shape)
        self.np_dtype = np.dtype(dtype)
        self.bf16 = bf16
        try:
            self.ndl_dtype = abi.BF16 if bf16 else abi.dtype_of(np.empty(0, self.np_dtype))
        except TypeError:
            self.ndl_dtype = -1  # raw buffer (e.g. uint32 scratch); not usable as a tensor
        if strides is None:
            st, acc = [], 1
            for s in reversed(self.shape):
                st.append(acc)
                acc *= s
            strides = tuple(reversed(st))
        self.strides = tuple(int(s) for s in strides)
        if _base is None:
            self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.np_dtype.itemsize
            p = _VP()
            check(lib().ndl_malloc(C.byref(p), max(self.nbytes, 1)))
            self._alloc = p.value
            self._owner = None
            self.ptr = p.value
        else:
            self._alloc = None
            self._owner = _base
            self.ptr = _base.ptr + _offset * self.np_dtype.itemsize

    @classmethod
    def from_numpy(cls, arr, bf16=False, stream=None):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype, bf16=bf16)
        check(lib().ndl_memcpy_htod(d.ptr, arr.ctypes.data, arr.nbytes, stream))
        check(lib().ndl_stream_synchronize(stream))
        return d

    def to_numpy(self, stream=None):
        assert self._owner is None, "copy out of the owning array"
        out = np.empty(self.shape, dtype=self.np_dtype)
        check(lib().ndl_stream_synchronize(stream))
        check(lib().ndl_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, stream))
        check(lib().ndl_stream_synchronize(stream))
        return out

    def view(self, index):
        """numpy-style basic-slicing view (ints, slices with any step)."""
        if not isinstance(index, tuple):
            index = (index,)
        index = index + (slice(None),) * (len(self.shape) - len(index))
        off, shape, strides = 0, [], []
        for ix, n, st in zip(index, self.shape, self.strides):
            if isinstance(ix, slice):
                start, stop, step = ix.indices(n)
                cnt = len(range(start, stop, step))
                off += start * st
                shape.append(cnt)
                strides.append(st * step)
            else:
                i = ix + n if ix < 0 else ix
                off += i * st
        base = self._owner or self
        return DeviceArray(shape, self.np_dtype, bf16=self.bf16, strides=strides,
                           _base=base, _offset=(self.ptr - base.ptr) // self.np_dtype.itemsize + off)

    def tensor(self):
        return abi.make_tensor(self.ptr, self.ndl_dtype, tuple(reversed(self.shape)),
                               tuple(reversed(self.strides)))

    def free(self):
        if self._alloc is not None:
            lib().ndl_free(self._alloc)
            self._alloc = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream=None):
    check(lib().ndl_stream_synchronize(stream))


def device_count():
    n = C.c_int(0)
    rc = lib().ndl_device_count(C.byref(n))
    return n.value if rc == 0 else 0


# ------------------------------------------------------------ device calls
def _kern(taps, borders, centers):
    if isinstance(borders, (str, int)):
        borders = [borders] * len(taps)
    return abi.Kernels(taps, borders, centers)


def sepconv_dev(src, dst, axes, taps, borders, accum=abi.ACCUM_F32, rounding=abi.ROUND_TRUNC,
                centers=None, stream=None):
    ks = _kern(taps, borders, centers)
    ts, td = src.tensor(), dst.tensor()
    return check(lib().ndl_sepconv_forward(C.byref(ts), C.byref(td), len(axes), abi.Kernels.axes(axes),
                                           ks.arr, accum, rounding, stream))


def adjoint_input_dev(g, gx, axes, taps, borders, accum=abi.ACCUM_F32, centers=None, stream=None):
    ks = _kern(taps, borders, centers)
    tg, tx = g.tensor(), gx.tensor()
    return check(lib().ndl_sepconv_adjoint_input(C.byref(tg), C.byref(tx), len(axes),
                                                 abi.Kernels.axes(axes), ks.arr, accum, stream))


def adjoint_taps_dev(x, g, axes, taps, borders, out_dev, accum=abi.ACCUM_F32, centers=None, stream=None):
    ks = _kern(taps, borders, centers)
    tx, tg = x.tensor(), g.tensor()
    return check(lib().ndl_sepconv_adjoint_taps(C.byref(tx), C.byref(tg), len(axes), abi.Kernels.axes(axes),
                                                ks.arr, out_dev.ptr, accum, stream))


def convolve_nd_dev(src, dst, kernel, border="clamp", stream=None):
    kernel = np.ascontiguousarray(kernel, dtype=np.float64)
    kext = (C.c_int64 * abi.NDL_MAX_DIM)(*reversed(kernel.shape))
    b = abi.BORDERS[border] if isinstance(border, str) else int(border)
    ts, td = src.tensor(), dst.tensor()
    return check(lib().ndl_convolve_nd(C.byref(ts), C.byref(td), kernel.ctypes.data_as(_DP), kext, b, stream))


MORPH_OPS = {"erode": 0, "dilate": 1}


def morph_separable_dev(src, dst, axes, masks, borders, op, centers=None, stream=None):
    """erode/dilate with a structuring element that is the product of the
    per-axis inclusion masks `masks` (nonzero = included)."""
    ks = _kern(masks, borders, centers)
    ts, td = src.tensor(), dst.tensor()
    return check(lib().ndl_morph_separable(C.byref(ts), C.byref(td), len(axes), abi.Kernels.axes(axes), ks.arr,
                                           MORPH_OPS[op], stream))


def morph_nd_dev(src, dst, mask, border, op, stream=None):
    """erode/dilate with an N-D mask (numpy order, nonzero = included)."""
    mask = np.ascontiguousarray(mask, dtype=np.float64)
    kext = (C.c_int64 * abi.NDL_MAX_DIM)(*reversed(mask.shape))
    b = abi.BORDERS[border] if isinstance(border, str) else int(border)
    ts, td = src.tensor(), dst.tensor()
    return check(lib().ndl_morph_nd(C.byref(ts), C.byref(td), mask.ctypes.data_as(_DP), kext, b, MORPH_OPS[op],
                                    stream))


def gaussian_scales_dev(src, levels, sigmas, border="clamp", accum=abi.ACCUM_REF, rounding=abi.ROUND_TRUNC,
                        stream=None):
    """levels[l] = gaussian_blur(src, sigmas[l], border) for DeviceArrays."""
    arr = (abi.Tensor * len(levels))(*[lv.tensor() for lv in levels])
    sg = np.ascontiguousarray(sigmas, dtype=np.float64)
    b = abi.BORDERS[border] if isinstance(border, str) else int(border)
    ts = src.tensor()
    return check(lib().ndl_gaussian_scales(C.byref(ts), arr, len(levels), sg.ctypes.data_as(_DP), b, accum,
                                           rounding, stream))


def dog_dev(levels, dogs, stream=None):
    """dogs[l] = levels[l+1] - levels[l]."""
    la = (abi.Tensor * len(levels))(*[lv.tensor() for lv in levels])
    da = (abi.Tensor * max(1, len(dogs)))(*[d.tensor() for d in dogs])
    return check(lib().ndl_dog(la, da, len(levels), stream))


def gradient_dev(src, dst, border="reflect", stream=None):
    """ndl::gradient for DeviceArrays: dst has a leading component axis (the
    last numpy axis, of size src's ndim), component a = d/dx_a (ndl axis a)."""
    b = abi.BORDERS[border] if isinstance(border, str) else int(border)
    ts, td = src.tensor(), dst.tensor()
    return check(lib().ndl_gradient(C.byref(ts), C.byref(td), b, stream))


def downsample_dev(src, dst, factor, channel_axis=0, stream=None):
    ts, td = src.tensor(), dst.tensor()
    return check(lib().ndl_downsample(C.byref(ts), C.byref(td), int(factor), int(channel_axis), stream))


def gaussian_downsample_dev(src, dst, sigma, factor, channel_axis=0, border="clamp", accum=abi.ACCUM_REF,
                            rounding=abi.ROUND_TRUNC, stream=None):
    """downsample(gaussian_blur(src, sigma, border), factor, channel_axis) for
    DeviceArrays; sigma is a scalar (every axis) or one value per ndl axis
    (<= 0: not blurred)."""
    nd = len(src.shape)
    sg = np.full(nd, float(sigma)) if np.isscalar(sigma) else np.ascontiguousarray(sigma, dtype=np.float64)
    if sg.shape != (nd,):
        raise ValueError(f"sigma needs {nd} entries")
    b = abi.BORDERS[border] if isinstance(border, str) else int(border)
    ts, td = src.tensor(), dst.tensor()
    return check(lib().ndl_gaussian_downsample(C.byref(ts), C.byref(td), sg.ctypes.data_as(_DP), b, int(factor),
                                               int(channel_axis), accum, rounding, stream))


def separable_mask(mask):
    """Per-axis inclusion masks (numpy order) when the nonzero set of `mask`
    is their product, else None."""
    inc = np.asarray(mask) != 0
    axes = [inc.any(axis=tuple(j for j in range(inc.ndim) if j != i)) for i in range(inc.ndim)]
    prod = axes[0]
    for a in axes[1:]:
        prod = np.multiply.outer(prod, a)
    return axes if np.array_equal(prod, inc) else None


def gaussian_taps(sigma):
    buf = np.zeros(4 * abi.NDL_MAX_TAPS, dtype=np.float64)
    n = check(lib().ndl_gaussian_taps(float(sigma), buf.ctypes.data_as(_DP), len(buf)))
    return buf[:n].copy()


# ------------------------------------------------------- host conveniences
def sepconv(img, axes, taps, borders, accum=abi.ACCUM_F32, rounding=abi.ROUND_TRUNC, centers=None,
            out_dtype=None, bf16=False):
    """Upload, run the per-axis chain on the GPU, download."""
    s = DeviceArray.from_numpy(img, bf16=bf16)
    d = DeviceArray(img.shape, out_dtype or img.dtype, bf16=bf16)
    sepconv_dev(s, d, axes, taps, borders, accum, rounding, centers)
    return d.to_numpy()


def gaussian_blur(img, sigma, border="clamp", accum=None):
    """ndl::gaussian_blur semantics: one sigma, one border, every axis in order."""
    if accum is None:
        accum = abi.ACCUM_REF if img.dtype in (np.float64,) or np.issubdtype(img.dtype, np.integer) else abi.ACCUM_F32
    taps = gaussian_taps(sigma)
    nd = img.ndim
    rounding = abi.ROUND_NEAREST if (nd >= 2 and np.issubdtype(img.dtype, np.integer)) else abi.ROUND_TRUNC
    return sepconv(img, list(range(nd)), [taps] * nd, [border] * nd, accum, rounding)


def convolve(img, kernel, border="clamp", out_dtype=None):
    """ndl::convolve with an N-D kernel (numpy order), FP64 reference arithmetic."""
    s = DeviceArray.from_numpy(img)
    d = DeviceArray(img.shape, out_dtype or img.dtype)
    convolve_nd_dev(s, d, kernel, border)
    return d.to_numpy()


def morph(img, mask, op, border="clamp", out_dtype=None, force_nd=False):
    """ndl::erode / ndl::dilate (morphology.h:57-101) with an N-D mask (numpy
    order): one min/max pass per axis when the mask is separable (a box),
    otherwise the one-pass N-D kernel."""
    s = DeviceArray.from_numpy(img)
    d = DeviceArray(img.shape, out_dtype or img.dtype)
    mask = np.asarray(mask, dtype=np.float64)
    sep = None if force_nd else separable_mask(mask)
    if sep is not None:
        n = img.ndim
        axes = [n - 1 - i for i in range(n) if mask.shape[i] > 1]
        masks = [sep[i].astype(np.float64) for i in range(n) if mask.shape[i] > 1]
        if not axes:  # a 1x..x1 element
            axes, masks = [0], [sep[n - 1].astype(np.float64)]
        morph_separable_dev(s, d, axes, masks, [border] * len(axes), op)
    else:
        morph_nd_dev(s, d, mask, border, op)
    return d.to_numpy()


def downsample(img, factor, channel_axis=0):
    """ndl::downsample (convolution.h:241-305); channel_axis is an ndl axis
    (0 = numpy's last)."""
    n = img.ndim
    shape = tuple(e if (n - 1 - i) == channel_axis else (e - 1) // factor + 1 for i, e in enumerate(img.shape))
    s = DeviceArray.from_numpy(img)
    d = DeviceArray(shape, img.dtype)
    downsample_dev(s, d, factor, channel_axis)
    return d.to_numpy()


def gaussian_downsample(img, sigma, factor, channel_axis=0, border="clamp"):
    """downsample(gaussian_blur(img, sigma, border), factor, channel_axis) in
    one call, bit-identical to the two (the blurred image in img's type)."""
    n = img.ndim
    shape = tuple(e if (n - 1 - i) == channel_axis else (e - 1) // factor + 1 for i, e in enumerate(img.shape))
    rounding = abi.ROUND_NEAREST if (n >= 2 and np.issubdtype(img.dtype, np.integer)) else abi.ROUND_TRUNC
    s = DeviceArray.from_numpy(img)
    d = DeviceArray(shape, img.dtype)
    gaussian_downsample_dev(s, d, sigma, factor, channel_axis, border, abi.ACCUM_REF, rounding)
    return d.to_numpy()


# ------------------------------------------------------------- HIP events
class _Hip:
    """Minimal HIP runtime calls for timing (events, streams) via the HIP
    runtime libndl_gpu.so itself loaded."""

    def __init__(self):
        lib()
        self.h = C.CDLL("libamdhip64.so.7")
        for n, a in [("hipEventCreate", [_P(_VP)]), ("hipEventCreateWithFlags", [_P(_VP), C.c_uint]),
                     ("hipEventRecord", [_VP, _VP]),
                     ("hipEventSynchronize", [_VP]), ("hipEventDestroy", [_VP]),
                     ("hipEventElapsedTime", [_P(C.c_float), _VP, _VP]),
                     ("hipStreamCreate", [_P(_VP)]), ("hipStreamDestroy", [_VP]),
                     ("hipDeviceSynchronize", []), ("hipSetDevice", [C.c_int])]:
            f = getattr(self.h, n)
            f.restype = C.c_int
            f.argtypes = a

    def ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    # hipEventDisableSystemFence: the event does not wait for a system-scope
    # release of the preceding work, only for its completion on the device
    EVENT_NO_SYSTEM_FENCE = 0x20000000

    def event(self, flags=0):
        e = _VP()
        self.ok(self.h.hipEventCreateWithFlags(C.byref(e), flags), "hipEventCreateWithFlags")
        return e.value

    def record(self, ev, stream=None):
        self.ok(self.h.hipEventRecord(ev, stream), "hipEventRecord")

    def elapsed_ms(self, a, b):
        self.ok(self.h.hipEventSynchronize(b), "hipEventSynchronize")
        ms = C.c_float()
        self.ok(self.h.hipEventElapsedTime(C.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value

    def stream(self):
        s = _VP()
        self.ok(self.h.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s.value

    def device_synchronize(self):
        self.ok(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")


_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = _Hip()
    return _hip
