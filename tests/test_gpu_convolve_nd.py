"""GPU parity of convolve() with a dense N-D kernel image
(processing/convolution.h:52-71; ndl_convolve_nd).

Integer images with an integer rank-1 kernel (boxes, outer products of
binomials and derivative stencils, zero entries included) take the
separable route: the reference's single FP64 sum is exact for them, so one
pass per axis gives the identical value.  Each case is bit-identical to the
oracle's N-D sum and to the one-output-per-thread N-D kernel
(NDL_ND_SEPARABLE=0); kernels that are not such products (a cross, float
weights) and float images stay on the N-D kernel.
"""
import sys

import numpy as np
import pytest

import ndl_amd
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def both(x, k, border, monkeypatch, out_dtype=None):
    monkeypatch.setenv("NDL_ND_SEPARABLE", "1")
    a = ndl_amd.convolve(x, k, border, out_dtype=out_dtype)
    monkeypatch.setenv("NDL_ND_SEPARABLE", "0")
    b = ndl_amd.convolve(x, k, border, out_dtype=out_dtype)
    monkeypatch.delenv("NDL_ND_SEPARABLE")
    return a, b


def outer(*vs):
    k = np.array(1.0)
    for v in vs:
        k = np.multiply.outer(k, np.asarray(v, dtype=np.float64))
    return k


CASES = [
    # (numpy image shape, dtype, kernel (numpy order), border)
    ((200, 300, 3), np.uint8, np.ones((7, 7, 1)), "wrap"),          # colour box blur (packed bytes)
    ((128, 256), np.uint8, np.ones((3, 9)), "clamp"),
    ((90, 70), np.int16, outer([1, 4, 6, 4, 1], [1, -2, 1]), "reflect"),
    ((20, 30, 40), np.uint16, outer([1, 2, 1], [1, 0, -1], [1, 2, 1]), "clamp"),
    ((64, 80), np.int32, outer([2, 0, 2], [1, 1, 1, 1, 1]), "wrap"),
    ((50, 60, 3), np.uint8, outer([1, 2, 1], [1, 2, 1], [1]), "reflect"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_separable_integer_kernels(case, monkeypatch):
    shape, dtype, k, border = CASES[case]
    rng = np.random.default_rng(100 + case)
    info = np.iinfo(dtype)
    x = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
    say("convolve nd", shape, np.dtype(dtype).name, k.shape, border)
    a, b = both(x, k, border, monkeypatch)
    want = oracle.convolve(x, k, border)
    np.testing.assert_array_equal(a, want)
    np.testing.assert_array_equal(b, want)


def test_wider_destination_type(monkeypatch):
    rng = np.random.default_rng(110)
    x = rng.integers(0, 256, size=(60, 90)).astype(np.uint8)
    k = np.ones((5, 5))
    a, b = both(x, k, "reflect", monkeypatch, out_dtype=np.int32)
    want = oracle.convolve(x, k, "reflect", dst=np.zeros(x.shape, np.int32))
    np.testing.assert_array_equal(a, want)
    np.testing.assert_array_equal(b, want)


@pytest.mark.parametrize("name", ["cross", "float weights", "float image"])
def test_other_kernels_stay_exact(name, monkeypatch):
    rng = np.random.default_rng(111)
    if name == "float image":
        x = rng.random((40, 50))
        k = np.ones((3, 3))
    else:
        x = rng.integers(0, 256, size=(40, 50)).astype(np.uint8)
        k = (np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=np.float64) if name == "cross"
             else outer([0.25, 0.5, 0.25], [0.25, 0.5, 0.25]))
    a, b = both(x, k, "clamp", monkeypatch)
    want = oracle.convolve(x, k, "clamp")
    np.testing.assert_array_equal(a, want)
    np.testing.assert_array_equal(b, want)
