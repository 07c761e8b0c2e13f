"""ndl_sepconv_adjoint_input_noalloc against ndl_sepconv_adjoint_input.

The allocation-free entry point keeps its edge weight rows in a
library-owned device array; it must give the allocating entry point's
result bit for bit (the same fused kernel), on two calls in a row with the
second one smaller (the array keeps the first call's rows), for clamp
(weight rows on every axis), asymmetric reflect, wrap and symmetric reflect
(no rows), and refuse what needs the per-axis chain.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import ndl_amd
from ndl_amd import abi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(entry, g, taps, borders, storage):
    bf = storage == "bf16"
    dg = ndl_amd.DeviceArray.from_numpy(g, bf16=bf)
    dx = ndl_amd.DeviceArray(g.shape, g.dtype, bf16=bf)
    ks = abi.Kernels(taps, borders)
    ax = abi.Kernels.axes(list(range(len(taps))))
    tg, tx = dg.tensor(), dx.tensor()
    lib = ndl_amd.lib()
    if entry == "noalloc":
        rc = lib.ndl_sepconv_adjoint_input_noalloc(C.byref(tg), C.byref(tx), len(taps), ax, ks.arr, None)
    else:
        rc = lib.ndl_sepconv_adjoint_input(C.byref(tg), C.byref(tx), len(taps), ax, ks.arr, abi.ACCUM_F32, None)
    ndl_amd.check(rc)
    ndl_amd.synchronize()
    return dx.to_numpy()


CASES = [
    # (borders, taps (ndl axis order), storage)
    (["clamp"] * 3, [11, 11, 19], "f32"),
    (["reflect", "clamp", "wrap"], [9, 7, 5], "f32"),
    (["reflect"] * 3, [11, 11, 19], "f32"),     # symmetric reflect: no weight rows
    (["reflect"] * 3, [9, 9, 9], "bf16"),       # half storage (symmetric: no rows)
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_adjoint_matches_allocating_entry(case):
    borders, ks, storage = CASES[case]
    rng = np.random.default_rng(700 + case)
    # asymmetric taps (except the symmetric case) so reflect needs weight rows
    taps = []
    for k in ks:
        t = rng.standard_normal(k) if case < 2 else np.array(ndl_amd.gaussian_taps((k - 1) / 6.0))
        taps.append(t)
    for shape in [(70, 48, 100), (23, 30, 41)]:  # second call smaller
        say("noalloc adjoint", borders, ks, storage, shape)
        g = rng.standard_normal(shape).astype(np.float32)
        if storage == "bf16":
            u = g.view(np.uint32).astype(np.uint64)
            g = ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)).astype(np.uint16)
        want = _run("alloc", g, taps, borders, storage)
        got = _run("noalloc", g, taps, borders, storage)
        np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))


def test_noalloc_adjoint_refuses_the_chain():
    # grad_in of another type than grad_out needs the per-axis chain
    lib = ndl_amd.lib()
    dg = ndl_amd.DeviceArray((10, 12, 14), np.float32)
    dx = ndl_amd.DeviceArray((10, 12, 14), np.float64)
    ks = abi.Kernels([np.ones(3)] * 3, ["clamp"] * 3)
    ax = abi.Kernels.axes([0, 1, 2])
    tg, tx = dg.tensor(), dx.tensor()
    assert lib.ndl_sepconv_adjoint_input_noalloc(C.byref(tg), C.byref(tx), 3, ax, ks.arr, None) == abi.ERR_DTYPE
    dx2 = ndl_amd.DeviceArray((10, 12, 14), np.float32)
    tx2 = dx2.tensor()
    kz = abi.Kernels([np.array([0.0, 1.0, 0.5])] * 3, ["clamp"] * 3)  # a zero tap: the chain skips it
    assert lib.ndl_sepconv_adjoint_input_noalloc(C.byref(tg), C.byref(tx2), 3, ax, kz.arr, None) == abi.ERR_ARG
    assert b"per-axis chain" in lib.ndl_last_error()
