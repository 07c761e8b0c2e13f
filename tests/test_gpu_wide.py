"""GPU parity of the wide-tap pass (kernels_wide.hip): one axis with a
run-time tap count beyond the one-axis kernels' 25 (sigma 5, 9, 12: 31, 55
and 73 taps), a thread per 8 outputs of a line.  Reference arithmetic is
bit-identical to the oracle and to the per-axis pass kernel (NDL_WIDE=0)
for float32, float64, uint8 and uint16 images, every border, the
unit-stride axis and the others, off-centre kernels, extents shorter than
the 8-output chunk and strided views; FP32 accumulation is held to 1e-6
relative on positive data.
"""
import sys

import numpy as np
import pytest

import ndl_amd
import oracle
from ndl_amd import abi

pytestmark = pytest.mark.gpu
REF, F32 = abi.ACCUM_REF, abi.ACCUM_F32
NEAR, TRUNC = abi.ROUND_NEAREST, abi.ROUND_TRUNC


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(x, axes, taps, borders, accum, rounding, monkeypatch, wide, centers=None, view=None):
    monkeypatch.setenv("NDL_WIDE", "1" if wide else "0")
    d = ndl_amd.DeviceArray.from_numpy(x)
    src = d.view(view) if view is not None else d
    out = ndl_amd.DeviceArray(src.shape, x.dtype)
    ndl_amd.sepconv_dev(src, out, axes, taps, borders, accum, rounding, centers=centers)
    return out.to_numpy()


def data(dtype, shape, rng):
    if np.issubdtype(dtype, np.integer):
        return rng.integers(0, np.iinfo(dtype).max, size=shape, endpoint=True).astype(dtype)
    return rng.random(shape).astype(dtype)


@pytest.mark.parametrize("sigma", [5.0, 9.0, 12.0])
@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8, np.uint16])
def test_wide_gaussian_bit_exact(sigma, border, dtype, monkeypatch):
    rng = np.random.default_rng(140)
    x = data(dtype, (40, 90, 100), rng)
    g = oracle.gaussian_taps(sigma)
    nd = 3
    rounding = NEAR if np.issubdtype(dtype, np.integer) else TRUNC
    say("wide", sigma, len(g), border, np.dtype(dtype).name)
    a = run(x, [0, 1, 2], [g] * nd, [border] * nd, REF, rounding, monkeypatch, True)
    b = run(x, [0, 1, 2], [g] * nd, [border] * nd, REF, rounding, monkeypatch, False)
    want = oracle.sepconv(x, [0, 1, 2], [g] * nd, [border] * nd, rounding)
    np.testing.assert_array_equal(a.view(np.uint8), want.view(np.uint8))
    np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_wide_one_axis_off_centre_and_short_extent(axis, monkeypatch):
    rng = np.random.default_rng(141)
    x = rng.standard_normal((6, 33, 5)).astype(np.float64)   # extents 5, 33, 6 (ndl order)
    t = rng.random(29) + 0.1
    for c in (0, 11, 28):
        a = run(x, [axis], [t], ["clamp"], REF, TRUNC, monkeypatch, True, centers=[c])
        b = run(x, [axis], [t], ["clamp"], REF, TRUNC, monkeypatch, False, centers=[c])
        want = oracle.sepconv(x, [axis], [t], ["clamp"], centers=[c])
        np.testing.assert_array_equal(a, want)
        np.testing.assert_array_equal(a, b)


def test_wide_strided_view(monkeypatch):
    rng = np.random.default_rng(142)
    x = rng.random((50, 80, 120)).astype(np.float32)
    view = (slice(None, None, -1), slice(3, 79, 2), slice(None, None, -1))
    g = oracle.gaussian_taps(6.0)
    a = run(x, [0, 1, 2], [g] * 3, ["reflect"] * 3, REF, TRUNC, monkeypatch, True, view=view)
    want = oracle.sepconv(np.ascontiguousarray(x[view]), [0, 1, 2], [g] * 3, ["reflect"] * 3)
    np.testing.assert_array_equal(a, want)


def test_wide_fp32_accumulation(monkeypatch):
    monkeypatch.setenv("NDL_DISABLE_FUSED", "1")
    rng = np.random.default_rng(143)
    x = rng.random((30, 70, 90)).astype(np.float32)
    g = oracle.gaussian_taps(5.0)
    a = run(x, [0, 1, 2], [g] * 3, ["reflect"] * 3, F32, TRUNC, monkeypatch, True)
    b = run(x, [0, 1, 2], [g] * 3, ["reflect"] * 3, F32, TRUNC, monkeypatch, False)
    want = oracle.sepconv(x.astype(np.float64), [0, 1, 2], [g] * 3, ["reflect"] * 3)
    for got in (a, b):
        assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-30)) <= 1e-6
