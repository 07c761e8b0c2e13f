"""GPU parity of the short-axis pass (kernels_short.hip): one pass along an
axis of at most 16 elements, a thread per line, the border fold resolved
into element indices on the host (processing/convolution.h:52-71 per
output).  Each case is bit-identical to the oracle and to the per-axis pass
kernel (NDL_SHORT_AXIS=0): axis extents 1..16, every border with taps
reaching up to the extent (clamp: beyond it), zero taps, off-centre kernels,
the axis first, in the middle and last, strided views, the reference chain
into and out of its FP64 intermediate, and FP32 accumulation; axes of at
most 4 elements also against the run-time-tap form (NDL_SHORT4=0).
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


def run(x, axes, taps, borders, accum, rounding, monkeypatch, short, centers=None, view=None):
    monkeypatch.setenv("NDL_SHORT_AXIS", "1" if short else "0")
    d = ndl_amd.DeviceArray.from_numpy(x)
    src = d.view(view) if view is not None else d
    out = ndl_amd.DeviceArray(src.shape, x.dtype)
    ndl_amd.sepconv_dev(src, out, axes, taps, borders, accum, rounding, centers=centers)
    return out.to_numpy()


def check(x, axes, taps, borders, monkeypatch, accum=REF, centers=None, view=None):
    integral = np.issubdtype(x.dtype, np.integer)
    rounding = NEAR if integral and len(axes) >= 2 else TRUNC
    a = run(x, axes, taps, borders, accum, rounding, monkeypatch, True, centers, view)
    b = run(x, axes, taps, borders, accum, rounding, monkeypatch, False, centers, view)
    np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))
    # axes of at most 4 elements: the compile-time-tap form against the
    # run-time one
    monkeypatch.setenv("NDL_SHORT4", "0")
    c = run(x, axes, taps, borders, accum, rounding, monkeypatch, True, centers, view)
    monkeypatch.delenv("NDL_SHORT4")
    np.testing.assert_array_equal(a.view(np.uint8), c.view(np.uint8))
    if accum == REF:
        sub = np.ascontiguousarray(x[view]) if view is not None else x
        want = oracle.sepconv(sub, axes, taps, borders, rounding, centers=centers)
        np.testing.assert_array_equal(a.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("m", [1, 2, 3, 4, 7, 16])
@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
def test_extents_and_borders(m, border, monkeypatch):
    rng = np.random.default_rng(80 + m)
    x = (rng.random((20, 30, m)) * 255).astype(np.uint8)
    r = m if border != "clamp" else m + 3
    taps = oracle.gaussian_taps(r / 3.0)[: 2 * r + 1] if r > 0 else np.array([1.0])
    say("short", m, border, len(taps))
    # the axis alone, and first in a chain (uint8 -> FP64 -> uint8)
    check(x, [0], [taps], [border], monkeypatch)
    check(x, [0, 1, 2], [taps, oracle.gaussian_taps(1.0), oracle.gaussian_taps(1.0)],
          [border, "clamp", "reflect"], monkeypatch)


def test_zero_and_off_centre_taps(monkeypatch):
    rng = np.random.default_rng(90)
    x = rng.standard_normal((12, 40, 5))
    t = np.array([0.5, 0.0, -1.0, 0.0, 2.0, 0.25])
    for c in (0, 2, 5):
        check(x, [0], [t], ["clamp"], monkeypatch, centers=[c])
    check(x, [0], [np.array([0.0, 1.5, 0.0])], ["wrap"], monkeypatch)


def test_axis_in_the_middle_and_last(monkeypatch):
    rng = np.random.default_rng(91)
    x = (rng.random((9, 4, 70)) * 1000).astype(np.int16)
    g = oracle.gaussian_taps(1.0)
    # dimension 1 has 4 elements (middle), dimension 2 has 9 (last)
    check(x, [0, 1, 2], [g, g, g], ["reflect", "clamp", "wrap"], monkeypatch)
    check(x, [1], [g], ["clamp"], monkeypatch)


def test_strided_view(monkeypatch):
    rng = np.random.default_rng(92)
    x = (rng.random((40, 50, 3)) * 255).astype(np.uint8)
    view = (slice(None, None, -1), slice(2, 48, 3), slice(None, None, -1))
    g = oracle.gaussian_taps(0.8)
    check(x, [0, 1, 2], [g, g, g], ["reflect"] * 3, monkeypatch, view=view)


@pytest.mark.parametrize("dtype", [np.float32, np.float16])
def test_fp32_accumulation(dtype, monkeypatch):
    # the per-axis chain (fused kernels off): the short axis takes this
    # kernel, bit-identical to the per-axis pass (same fmaf order)
    monkeypatch.setenv("NDL_DISABLE_FUSED", "1")
    rng = np.random.default_rng(93)
    x = rng.random((30, 40, 3)).astype(dtype)
    g = oracle.gaussian_taps(1.0)
    check(x, [0], [g], ["clamp"], monkeypatch, accum=F32)
    check(x, [0, 1], [g, g], ["clamp", "reflect"], monkeypatch, accum=F32)
