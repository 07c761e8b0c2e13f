"""libndl_gpu.so on the CPU: it loads, exports every entry point include/*.h
declares, and rejects bad arguments before touching the GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import ndl_amd
from ndl_amd import abi, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ndl_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", ndl_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ndl_\w+)", out))
    missing = syms - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    assert syms == set(ndl_amd.SIGNATURES), "python binding out of sync with the header"
    lib = ndl_amd.lib()
    for s in syms:
        assert hasattr(lib, s)


def test_version_and_errors_without_gpu():
    lib = ndl_amd.lib()
    assert lib.ndl_version() >= 100
    ks = abi.Kernels([np.ones(3)], ["clamp"])
    ax = abi.Kernels.axes([0])
    t = abi.make_tensor(0x1000, abi.F32, (8, 8), (1, 8))
    # null descriptors
    assert lib.ndl_sepconv_forward(None, C.byref(t), 1, ax, ks.arr, 0, 0, None) == abi.ERR_ARG
    # shape mismatch
    t2 = abi.make_tensor(0x100000, abi.F32, (8, 9), (1, 8))
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(t2), 1, ax, ks.arr, 0, 0, None) == abi.ERR_SHAPE
    assert b"extent" in lib.ndl_last_error()
    # aliasing
    t3 = abi.make_tensor(0x1000 + 16, abi.F32, (8, 8), (1, 8))
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(t3), 1, ax, ks.arr, 0, 0, None) == abi.ERR_ALIAS
    # reach beyond the extent under wrap/reflect (single-fold maps)
    far = abi.make_tensor(0x100000, abi.F32, (8, 8), (1, 8))
    kw = abi.Kernels([np.ones(19)], ["wrap"])
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(far), 1, ax, kw.arr, 0, 0, None) == abi.ERR_RADIUS
    # FP32 accumulation on an integer image
    tu = abi.make_tensor(0x1000, abi.U8, (8, 8), (1, 8))
    tu2 = abi.make_tensor(0x100000, abi.U8, (8, 8), (1, 8))
    assert lib.ndl_sepconv_forward(C.byref(tu), C.byref(tu2), 1, ax, ks.arr, abi.ACCUM_F32, 0, None) == abi.ERR_DTYPE
    # axis out of range, bad border, bad center, too many taps
    bad = abi.Kernels.axes([2])
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(far), 1, bad, ks.arr, 0, 0, None) == abi.ERR_SHAPE
    kb = abi.Kernels([np.ones(3)], [7])
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(far), 1, ax, kb.arr, 0, 0, None) == abi.ERR_ARG
    kc = abi.Kernels([np.ones(3)], ["clamp"], centers=[3])
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(far), 1, ax, kc.arr, 0, 0, None) == abi.ERR_ARG
    kt = abi.Kernels([np.ones(abi.NDL_MAX_TAPS + 1)], ["clamp"])
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(far), 1, ax, kt.arr, 0, 0, None) == abi.ERR_ARG
    # adjoint on integer fields
    assert lib.ndl_sepconv_adjoint_input(C.byref(tu), C.byref(tu2), 1, ax, ks.arr, 0, None) == abi.ERR_DTYPE
    # plan with an invalid item is refused and no plan is returned
    item = abi.BatchItem()
    item.src, item.dst, item.naxes = t, t3, 1
    item.axes[0] = 0
    item.k[0] = ks.arr[0]
    plan = C.c_void_p(1)
    assert lib.ndl_plan_create(C.byref(item), 1, abi.ACCUM_F32, C.byref(plan)) == abi.ERR_ALIAS
    assert plan.value is None


def test_empty_extent_is_a_no_op_without_gpu():
    lib = ndl_amd.lib()
    ks = abi.Kernels([np.ones(3)], ["clamp"])
    t = abi.make_tensor(0, abi.F32, (0, 8), (1, 1))
    t2 = abi.make_tensor(0, abi.F32, (0, 8), (1, 1))
    assert lib.ndl_sepconv_forward(C.byref(t), C.byref(t2), 1, abi.Kernels.axes([1]), ks.arr, 0, 0, None) == 0


def test_gaussian_taps_match_reference_construction():
    lib = ndl_amd.lib()
    for s in [0.5, 1.0, 1.5, 2.0, 3.0, 1 / 3]:
        buf = np.zeros(64)
        n = lib.ndl_gaussian_taps(s, buf.ctypes.data_as(C.POINTER(C.c_double)), 64)
        np.testing.assert_array_equal(buf[:n], workloads.gaussian_taps(s))
    import oracle
    for s in [0.5, 1.5, 3.0, 0.7777]:
        np.testing.assert_array_equal(workloads.gaussian_taps(s),
                                      oracle.gaussian_taps(s))
    buf = np.zeros(4)
    assert lib.ndl_gaussian_taps(2.0, buf.ctypes.data_as(C.POINTER(C.c_double)), 4) == abi.ERR_ARG
    assert lib.ndl_gaussian_taps(-1.0, buf.ctypes.data_as(C.POINTER(C.c_double)), 4) == abi.ERR_ARG


def test_workload_rng_is_counter_based():
    a = workloads.uniform(3, 0, 1000)
    b = workloads.uniform(3, 0, 400, start=600)
    np.testing.assert_array_equal(a[600:], b)
    assert 0 <= a.min() and a.max() < 1
    assert abs(a.mean() - 0.5) < 0.05
    n = workloads.normal(1, 0, 20000)
    assert abs(n.mean()) < 0.05 and abs(n.std() - 1) < 0.05
    items = workloads.config4()
    assert len(items) == 256
    assert sum(len(s) == 2 for s, _, _ in items) == 192
