"""numpy front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, as the checker or the timed CPU baseline; the product
path (ndl_amd, include/ndl) never does.

liboracle.so is the plain-C restatement (oracle/ndl_oracle.c).  When the
reference sources were present at build time, _ref/libndl_ref.so wraps the
reference's own templates (oracle/ref_shim.cpp) and is exposed as `ref`.
Arrays follow ndl_amd.abi's convention: numpy shape is reversed ndl extent,
ndl axis d is numpy axis ndim-1-d.
"""
import ctypes as C
import os

import numpy as np

from ndl_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libndl_ref.so")

_P = C.POINTER
_T = _P(abi.Tensor)
_K = _P(abi.AxisKernel)
_I32P = _P(C.c_int32)
_I64P = _P(C.c_int64)
_DP = _P(C.c_double)


def build():
    """Compile liboracle.so (and _ref when the reference is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (run make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    abi.bind(lib, "oracle_convolve", C.c_int, _T, _T, _DP, _I64P, C.c_int)
    abi.bind(lib, "oracle_morph", C.c_int, _T, _T, _DP, _I64P, C.c_int, C.c_int)
    abi.bind(lib, "oracle_downsample", C.c_int, _T, _T, C.c_int, C.c_int)
    abi.bind(lib, "oracle_sepconv", C.c_int, _T, _T, C.c_int, _I32P, _K, C.c_int)
    abi.bind(lib, "oracle_sepconv_adjoint_input", C.c_int, _T, _T, C.c_int, _I32P, _K)
    abi.bind(lib, "oracle_sepconv_adjoint_taps", C.c_int, _T, _T, C.c_int, _I32P, _K, _DP)
    abi.bind(lib, "oracle_gaussian_taps", C.c_int, C.c_double, _DP, C.c_int)
    abi.bind(lib, "oracle_kernel_size", C.c_int, C.c_double)
    abi.bind(lib, "oracle_half_to_double", C.c_double, C.c_uint16)
    abi.bind(lib, "oracle_double_to_half", C.c_uint16, C.c_double)
    abi.bind(lib, "oracle_bf16_to_double", C.c_double, C.c_uint16)
    abi.bind(lib, "oracle_double_to_bf16", C.c_uint16, C.c_double)
    abi.bind(lib, "oracle_cvtt_i32", C.c_int32, C.c_double)
    abi.bind(lib, "oracle_set_threads", None, C.c_int)
    abi.bind(lib, "oracle_get_threads", C.c_int)
    return lib


def _load_ref():
    if not os.path.exists(REF_PATH):
        return None
    lib = C.CDLL(REF_PATH)
    abi.bind(lib, "ref_convolve", C.c_int, _T, _T, _DP, _I64P, C.c_int)
    abi.bind(lib, "ref_gaussian_blur", C.c_int, _T, _T, C.c_double, C.c_int)
    abi.bind(lib, "ref_sepconv", C.c_int, _T, _T, C.c_int, _I32P, _K, C.c_int)
    abi.bind(lib, "ref_kernel_size", C.c_int, C.c_double)
    if hasattr(lib, "ref_morph"):
        abi.bind(lib, "ref_morph", C.c_int, _T, _T, _DP, _I64P, C.c_int, C.c_int)
    if hasattr(lib, "ref_downsample"):
        abi.bind(lib, "ref_downsample", C.c_int, _T, _T, C.c_int, C.c_int)
    return lib


lib = _load()
ref_lib = _load_ref()


class threads:
    """with oracle.threads(n): run the oracle's passes on n host threads.  The
    passes split independent outputs (and transposed scatters line by line),
    so results are bit-identical to one thread."""

    def __init__(self, n=None):
        self.n = int(n or os.cpu_count() or 1)

    def __enter__(self):
        self.saved = lib.oracle_get_threads()
        lib.oracle_set_threads(self.n)
        return self

    def __exit__(self, *exc):
        lib.oracle_set_threads(self.saved)
        return False


class OracleError(RuntimeError):
    def __init__(self, rc, what):
        super().__init__(f"{what} failed with status {rc}")
        self.rc = rc


def _check(rc, what):
    if rc != 0:
        raise OracleError(rc, what)


def _border(b):
    return abi.BORDERS[b] if isinstance(b, str) else int(b)


def kernel_size(sigma):
    return lib.oracle_kernel_size(float(sigma))


def gaussian_taps(sigma):
    buf = np.zeros(abi.NDL_MAX_TAPS * 4, dtype=np.float64)
    n = lib.oracle_gaussian_taps(float(sigma), buf.ctypes.data_as(_DP), len(buf))
    _check(n if n < 0 else 0, "gaussian_taps")
    return buf[:n].copy()


def _out_like(src, dst_dtype):
    return np.zeros(src.shape, dtype=src.dtype if dst_dtype is None else dst_dtype)


def convolve(src, kernel, border="clamp", dst=None, src_bf16=False, dst_bf16=False, use_ref=False):
    """ndl::convolve(src, dst, kernel, border) with an N-D kernel (numpy order)."""
    kernel = np.ascontiguousarray(kernel, dtype=np.float64)
    assert kernel.ndim == src.ndim
    if dst is None:
        dst = _out_like(src, None)
    kext = (C.c_int64 * abi.NDL_MAX_DIM)(*reversed(kernel.shape))
    ts, td = abi.host_tensor(src, src_bf16), abi.host_tensor(dst, dst_bf16)
    kp = kernel.ctypes.data_as(_DP)
    if use_ref:
        _check(ref_lib.ref_convolve(C.byref(ts), C.byref(td), kp, kext, _border(border)), "ref_convolve")
    else:
        _check(lib.oracle_convolve(C.byref(ts), C.byref(td), kp, kext, _border(border)), "oracle_convolve")
    return dst


def morph(src, kernel, op, border="clamp", dst=None, use_ref=False):
    """ndl::erode (op "erode") / ndl::dilate (op "dilate") with an N-D
    structuring element (numpy order; nonzero = included),
    morphology.h:57-101."""
    kernel = np.ascontiguousarray(kernel, dtype=np.float64)
    assert kernel.ndim == src.ndim and op in ("erode", "dilate")
    if dst is None:
        dst = _out_like(src, None)
    kext = (C.c_int64 * abi.NDL_MAX_DIM)(*reversed(kernel.shape))
    ts, td = abi.host_tensor(src), abi.host_tensor(dst)
    kp = kernel.ctypes.data_as(_DP)
    o = 0 if op == "erode" else 1
    if use_ref:
        _check(ref_lib.ref_morph(C.byref(ts), C.byref(td), kp, kext, _border(border), o), "ref_morph")
    else:
        _check(lib.oracle_morph(C.byref(ts), C.byref(td), kp, kext, _border(border), o), "oracle_morph")
    return dst


def downsample_shape(shape, factor, channel_axis=0):
    """Result shape (numpy order) of ndl::downsample; channel_axis is an ndl
    axis (0 = numpy's last)."""
    n = len(shape)
    return tuple(e if (n - 1 - i) == channel_axis else (e - 1) // factor + 1 for i, e in enumerate(shape))


def downsample(src, factor, channel_axis=0, use_ref=False):
    """ndl::downsample(src, factor, channelAxis), convolution.h:241-305."""
    src = np.ascontiguousarray(src)
    dst = np.zeros(downsample_shape(src.shape, factor, channel_axis), dtype=src.dtype)
    ts, td = abi.host_tensor(src), abi.host_tensor(dst)
    if use_ref:
        _check(ref_lib.ref_downsample(C.byref(ts), C.byref(td), factor, channel_axis), "ref_downsample")
    else:
        _check(lib.oracle_downsample(C.byref(ts), C.byref(td), factor, channel_axis), "oracle_downsample")
    return dst


def sepconv(src, axes, taps, borders, rounding=abi.ROUND_TRUNC, dst=None, centers=None,
            src_bf16=False, dst_bf16=False, use_ref=False):
    """Per-axis pass chain (ndl axis numbers), FP64 intermediates, final cast."""
    if dst is None:
        dst = _out_like(src, None)
    ks = abi.Kernels(taps, borders, centers)
    ax = abi.Kernels.axes(axes)
    ts, td = abi.host_tensor(src, src_bf16), abi.host_tensor(dst, dst_bf16)
    if use_ref:
        _check(ref_lib.ref_sepconv(C.byref(ts), C.byref(td), len(axes), ax, ks.arr, rounding), "ref_sepconv")
    else:
        _check(lib.oracle_sepconv(C.byref(ts), C.byref(td), len(axes), ax, ks.arr, rounding), "oracle_sepconv")
    return dst


def gaussian_blur(src, sigma, border="clamp", dst=None, use_ref=False):
    """ndl::gaussian_blur: same sigma and border on every axis."""
    if dst is None:
        dst = _out_like(src, None)
    if use_ref:
        ts, td = abi.host_tensor(src), abi.host_tensor(dst)
        _check(ref_lib.ref_gaussian_blur(C.byref(ts), C.byref(td), float(sigma), _border(border)),
               "ref_gaussian_blur")
        return dst
    taps = gaussian_taps(sigma)
    nd = src.ndim
    integral = np.issubdtype(dst.dtype, np.integer)
    rounding = abi.ROUND_NEAREST if (nd >= 2 and integral) else abi.ROUND_TRUNC
    return sepconv(src, list(range(nd)), [taps] * nd, [border] * nd, rounding, dst=dst)


def adjoint_input(g, axes, taps, borders, centers=None, dst_dtype=np.float64):
    gx = np.zeros(g.shape, dtype=dst_dtype)
    ks = abi.Kernels(taps, borders, centers)
    ax = abi.Kernels.axes(axes)
    tg, tx = abi.host_tensor(g), abi.host_tensor(gx)
    _check(lib.oracle_sepconv_adjoint_input(C.byref(tg), C.byref(tx), len(axes), ax, ks.arr),
           "oracle_sepconv_adjoint_input")
    return gx


def adjoint_taps(x, g, axes, taps, borders, centers=None):
    ks = abi.Kernels(taps, borders, centers)
    ax = abi.Kernels.axes(axes)
    out = np.zeros(sum(len(t) for t in taps), dtype=np.float64)
    tx, tg = abi.host_tensor(x), abi.host_tensor(g)
    _check(lib.oracle_sepconv_adjoint_taps(C.byref(tx), C.byref(tg), len(axes), ax, ks.arr,
                                           out.ctypes.data_as(_DP)), "oracle_sepconv_adjoint_taps")
    return out


# ---- storage codecs (bit-level) ----------------------------------------

def bf16_to_f64(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def f64_to_bf16(v):
    return np.array([lib.oracle_double_to_bf16(float(x)) for x in np.ravel(v)], dtype=np.uint16).reshape(np.shape(v))
