"""ctypes mirror of include/ndl_gpu.h (the C ABI of libndl_gpu.so).

Array convention: ndl extents are x-fastest (dimension 0 has stride 1 in a
dense image, image/core.h:858-864).  A numpy array of shape
(e[n-1], ..., e[1], e[0]) in C order is the ndl image of extent
(e[0], ..., e[n-1]); ndl axis d is numpy axis n-1-d.  Strides, including
negative ones from numpy views, are carried over in elements.
"""
import ctypes as C

import numpy as np

NDL_MAX_DIM = 8
NDL_MAX_TAPS = 1024

U8, I8, U16, I16, I32, F32, F64, F16, BF16 = range(9)
CLAMP, WRAP, REFLECT = 0, 1, 2
ACCUM_REF, ACCUM_F32 = 0, 1
ROUND_TRUNC, ROUND_NEAREST = 0, 1

OK = 0
ERR_ARG, ERR_DTYPE, ERR_SHAPE, ERR_RADIUS, ERR_ALIAS, ERR_HIP, ERR_NOMEM = range(-1, -8, -1)

BORDERS = {"clamp": CLAMP, "wrap": WRAP, "reflect": REFLECT}

# numpy storage dtype for each ndl dtype (bf16 is carried as raw uint16 bits)
NP_OF = {
    U8: np.uint8, I8: np.int8, U16: np.uint16, I16: np.int16, I32: np.int32,
    F32: np.float32, F64: np.float64, F16: np.float16, BF16: np.uint16,
}
ITEMSIZE = {k: np.dtype(v).itemsize for k, v in NP_OF.items()}
NAME = {U8: "u8", I8: "i8", U16: "u16", I16: "i16", I32: "i32", F32: "f32",
        F64: "f64", F16: "f16", BF16: "bf16"}
FLOAT_TYPES = (F32, F64, F16, BF16)


def dtype_of(arr, bf16=False):
    """ndl dtype code of a numpy array (uint16 arrays are bf16 when bf16=True)."""
    if bf16:
        assert arr.dtype == np.uint16
        return BF16
    for k, v in NP_OF.items():
        if k != BF16 and arr.dtype == np.dtype(v):
            return k
    raise TypeError(f"no ndl dtype for {arr.dtype}")


class Tensor(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("dtype", C.c_int32),
        ("ndim", C.c_int32),
        ("extent", C.c_int64 * NDL_MAX_DIM),
        ("stride", C.c_int64 * NDL_MAX_DIM),
    ]


class AxisKernel(C.Structure):
    _fields_ = [
        ("taps", C.POINTER(C.c_double)),
        ("count", C.c_int32),
        ("center", C.c_int32),
        ("border", C.c_int32),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("src", Tensor),
        ("dst", Tensor),
        ("naxes", C.c_int32),
        ("axes", C.c_int32 * NDL_MAX_DIM),
        ("k", AxisKernel * NDL_MAX_DIM),
    ]


def make_tensor(ptr, dtype, extent, stride):
    """Descriptor for memory at `ptr` (element (0,..,0)); extent/stride x-fastest."""
    t = Tensor()
    t.data = C.c_void_p(int(ptr))
    t.dtype = int(dtype)
    t.ndim = len(extent)
    assert 1 <= t.ndim <= NDL_MAX_DIM
    for d, (e, s) in enumerate(zip(extent, stride)):
        t.extent[d] = int(e)
        t.stride[d] = int(s)
    return t


def host_tensor(arr, bf16=False):
    """Descriptor of a numpy array (host memory) in ndl axis order."""
    isz = arr.dtype.itemsize
    ext = tuple(reversed(arr.shape))
    st = tuple(s // isz for s in reversed(arr.strides))
    return make_tensor(arr.ctypes.data, dtype_of(arr, bf16), ext, st)


class Kernels:
    """Keeps per-axis tap arrays alive while an AxisKernel array is in use."""

    def __init__(self, taps_list, borders, centers=None):
        n = len(taps_list)
        self._bufs = [np.ascontiguousarray(t, dtype=np.float64) for t in taps_list]
        self.arr = (AxisKernel * max(n, 1))()
        for i, b in enumerate(self._bufs):
            self.arr[i].taps = b.ctypes.data_as(C.POINTER(C.c_double))
            self.arr[i].count = len(b)
            self.arr[i].center = len(b) // 2 if centers is None or centers[i] is None else int(centers[i])
            bb = borders[i]
            self.arr[i].border = BORDERS[bb] if isinstance(bb, str) else int(bb)
        self.n = n

    @staticmethod
    def axes(axes):
        a = (C.c_int32 * NDL_MAX_DIM)()
        for i, x in enumerate(axes):
            a[i] = int(x)
        return a


def bind(lib, name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f
