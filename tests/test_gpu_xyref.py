"""GPU parity of the two-pass reference kernel (kernels_xyref.hip): two
consecutive passes of the NDL_ACCUM_REF chain in one launch, the first
pass's FP64 results held in LDS (gaussian_blur's chain,
processing/convolution.h:151-185).

Every case is bit-identical to the oracle and to the chain run one axis per
launch (NDL_XY_REF=0): element types u8/u16/i16/f32/f64 (i8 and i32 take
the chain) with the final rounding of integral outputs, partial 64 x 32
tiles, x extents at the 16 minimum, tap counts 3..25 (25 on float64 exceeds
the LDS budget and takes the chain), off-centre kernels, kernels with zero
taps or unequal counts (the chain), every border, batch dimensions, axis
pairs other than (0, 1) and in reverse order, and strided (mirrored,
stepped) views.  Chains that end in an integral image run their last pass
from the FP64 intermediate (a colour image: the channel pass, then x and y
here; a volume: x and y here, then the z line march), rounded to nearest.
"""
import sys

import numpy as np
import pytest

import ndl_amd
import oracle
from ndl_amd import abi

pytestmark = pytest.mark.gpu

REF = abi.ACCUM_REF
NEAR, TRUNC = abi.ROUND_NEAREST, abi.ROUND_TRUNC


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def data(dtype, shape, rng):
    if np.issubdtype(dtype, np.floating):
        return (rng.standard_normal(shape) * 50.0).astype(dtype)
    info = np.iinfo(dtype)
    return rng.integers(max(info.min, -100000), min(info.max, 100000), size=shape, endpoint=True).astype(dtype)


def run(x, axes, taps, borders, rounding, centers=None, view=None, monkeypatch=None, pair=True):
    monkeypatch.setenv("NDL_XY_REF", "1" if pair else "0")
    d = ndl_amd.DeviceArray.from_numpy(x)
    src = d.view(view) if view is not None else d
    out = ndl_amd.DeviceArray(src.shape, x.dtype)
    ndl_amd.sepconv_dev(src, out, axes, taps, borders, REF, rounding, centers=centers)
    return out.to_numpy()


def check(x, axes, taps, borders, monkeypatch, centers=None, view=None):
    integral = np.issubdtype(x.dtype, np.integer)
    rounding = NEAR if integral and len(axes) >= 2 else TRUNC
    a = run(x, axes, taps, borders, rounding, centers, view, monkeypatch, True)
    b = run(x, axes, taps, borders, rounding, centers, view, monkeypatch, False)
    sub = np.ascontiguousarray(x[view]) if view is not None else x
    want = oracle.sepconv(sub, axes, taps, borders, rounding, centers=centers)
    assert a.dtype == want.dtype
    np.testing.assert_array_equal(a.view(np.uint8), want.view(np.uint8))
    np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


def gauss(k):
    return oracle.gaussian_taps((k // 2) / 3.0)


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.int32, np.float32, np.float64])
@pytest.mark.parametrize("shape", [(100, 130), (37, 70, 5)[::-1], (9, 33, 16), (3, 4, 40, 66), (50, 70, 3)])
def test_types_and_shapes(dtype, shape, monkeypatch):
    rng = np.random.default_rng(71)
    x = data(dtype, shape, rng)
    nd = len(shape)
    g = gauss(7)
    say("xy", np.dtype(dtype).name, shape)
    check(x, list(range(nd)), [g] * nd, ["reflect", "clamp", "wrap", "clamp"][:nd], monkeypatch)


@pytest.mark.parametrize("k", [3, 5, 13, 25])
@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
def test_tap_counts_and_borders(k, border, monkeypatch):
    rng = np.random.default_rng(72 + k)
    for dtype in (np.float32, np.float64, np.uint8):
        x = data(dtype, (45, 77), rng)
        g = gauss(k)
        say("xy taps", k, border, np.dtype(dtype).name)
        check(x, [0, 1], [g, g], [border, border], monkeypatch)


def test_off_centre_and_zero_taps(monkeypatch):
    rng = np.random.default_rng(73)
    x = data(np.float32, (40, 50, 70), rng)
    tx = np.array([0.5, 0.0, -1.25, 2.0, 0.0])
    ty = np.array([0.0, 1.0, 0.25, 0.0, 0.0, 3.0, -0.5])
    check(x, [0, 1, 2], [tx, ty, gauss(5)], ["reflect", "wrap", "clamp"], monkeypatch, centers=[0, 6, 2])
    check(x, [0, 1], [tx, ty], ["clamp", "reflect"], monkeypatch, centers=[4, 1])
    # equal counts, no zero tap, off-centre: this kernel
    ta = np.array([0.25, -0.5, 1.5, 0.125, 2.0, -1.0, 0.75])
    tb = np.array([1.0, 0.5, 0.25, -0.125, 3.0, 0.5, -2.0])
    check(x, [0, 1, 2], [ta, tb, gauss(3)], ["wrap", "reflect", "clamp"], monkeypatch, centers=[1, 5, 1])
    check(x, [1, 2], [tb, ta], ["clamp", "wrap"], monkeypatch, centers=[6, 0])


@pytest.mark.parametrize("axes", [[1, 2], [0, 2], [2, 0], [1, 0, 2]])
def test_axis_pairs(axes, monkeypatch):
    rng = np.random.default_rng(74)
    x = data(np.float64, (36, 41, 66), rng)
    taps = [gauss(7), gauss(7), gauss(5)][:len(axes)]
    check(x, axes, taps, ["reflect", "clamp", "wrap"][:len(axes)], monkeypatch)


@pytest.mark.parametrize("name,view", [
    ("mirrored x", (slice(None), slice(None), slice(None, None, -1))),
    ("stepped y, window x", (slice(None), slice(0, 60, 2), slice(5, 90))),
    ("mirrored z, stepped x", (slice(None, None, -1), slice(None), slice(0, 100, 3))),
])
def test_strided_views(name, view, monkeypatch):
    rng = np.random.default_rng(75)
    x = data(np.uint8, (12, 60, 100), rng)
    say("xy view", name)
    check(x, [0, 1, 2], [gauss(7)] * 3, ["clamp"] * 3, monkeypatch, view=view)


def test_gaussian_blur_2d_uint8_one_launch(monkeypatch):
    # the drop-in's default on an 8-bit image: both passes in this kernel,
    # rounded to nearest on the way into uint8 (convolution.h:177-185)
    rng = np.random.default_rng(76)
    x = rng.integers(0, 256, size=(300, 517)).astype(np.uint8)
    got = ndl_amd.gaussian_blur(x, 2.0, "reflect", accum=REF)
    np.testing.assert_array_equal(got, oracle.gaussian_blur(x, 2.0, "reflect"))
