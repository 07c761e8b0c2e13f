"""GPU parity of the three-pass reference kernel (cxy_ref_kernel,
kernels_xyref.hip): gaussian_blur()'s FP64 chain (processing/convolution.h:
151-185) over a short first axis (the channels of a colour image, at most
four elements), then x, then y, in one launch with both FP64 intermediates
in LDS.

Every case is bit-identical to the oracle and to the same chain with the
channel pass on its own kernel (NDL_CXY_REF=0: the short-axis kernel into
an FP64 image, then the two-pass kernel): uint8 / uint16 / int16 / float32
images with 1-4 channels, tap counts 3..9 (the reach of a channel pass
stays within the channel count), every border, partial tiles,
a volume whose chain continues after the three passes (FP64 output, then
the z line march), batch dimensions and a mirrored view.
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


def run(x, axes, taps, borders, rounding, fused, monkeypatch, view=None):
    monkeypatch.setenv("NDL_CXY_REF", "1" if fused else "0")
    d = ndl_amd.DeviceArray.from_numpy(x)
    src = d.view(view) if view is not None else d
    out = ndl_amd.DeviceArray(src.shape, x.dtype)
    ndl_amd.sepconv_dev(src, out, axes, taps, borders, REF, rounding)
    return out.to_numpy()


def check(x, sigma, borders, monkeypatch, view=None):
    nd = x.ndim
    axes = list(range(nd))
    taps = [ndl_amd.gaussian_taps(sigma)] * nd
    rounding = NEAR if np.issubdtype(x.dtype, np.integer) else TRUNC
    a = run(x, axes, taps, borders, rounding, True, monkeypatch, view)
    b = run(x, axes, taps, borders, rounding, False, monkeypatch, view)
    sub = np.ascontiguousarray(x[view]) if view is not None else x
    want = oracle.sepconv(sub, axes, taps, borders, rounding)
    np.testing.assert_array_equal(b, want)
    np.testing.assert_array_equal(a, want)


# (numpy shape: y, x, channels), dtype, sigma (taps), border; the taps
# reach at most the channel count past each end (the reference rejects more)
CASES = [
    ((96, 160, 3), np.uint8, 1.0, "clamp"),     # 7 taps
    ((70, 130, 3), np.uint8, 0.9, "reflect"),   # 7 taps, partial tiles
    ((64, 64, 4), np.uint8, 1.2, "wrap"),       # 9 taps, four channels
    ((40, 80, 1), np.uint8, 0.3, "reflect"),    # one channel, 3 taps
    ((50, 70, 2), np.uint16, 0.6, "clamp"),     # 5 taps
    ((45, 90, 3), np.int16, 0.8, "wrap"),
    ((60, 100, 3), np.float32, 1.0, "reflect"),
    ((33, 17, 3), np.float32, 0.4, "clamp"),    # x at the 16 minimum plus one
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_colour_images(case, monkeypatch):
    shape, dtype, sigma, border = CASES[case]
    rng = np.random.default_rng(300 + case)
    x = data(dtype, shape, rng)
    say("cxy", shape, np.dtype(dtype).name, sigma, border)
    check(x, sigma, [border] * 3, monkeypatch)


def test_mixed_borders(monkeypatch):
    rng = np.random.default_rng(310)
    x = data(np.uint8, (72, 120, 3), rng)
    check(x, 1.0, ["wrap", "clamp", "reflect"], monkeypatch)


def test_volume_continues_after_three_passes(monkeypatch):
    # {3, x, y, z}: the three passes write the FP64 intermediate, z follows
    rng = np.random.default_rng(311)
    x = data(np.uint8, (12, 40, 64, 3), rng)
    check(x, 0.8, ["reflect"] * 4, monkeypatch)


def test_batch_dimension(monkeypatch):
    # the chain over the first three axes only; the fourth is a batch
    rng = np.random.default_rng(312)
    x = data(np.float32, (5, 40, 48, 3), rng)
    taps = [ndl_amd.gaussian_taps(1.0)] * 3
    a = run(x, [0, 1, 2], taps, ["clamp"] * 3, TRUNC, True, monkeypatch)
    b = run(x, [0, 1, 2], taps, ["clamp"] * 3, TRUNC, False, monkeypatch)
    want = oracle.sepconv(x, [0, 1, 2], taps, ["clamp"] * 3, TRUNC)
    np.testing.assert_array_equal(b, want)
    np.testing.assert_array_equal(a, want)


def test_mirrored_view(monkeypatch):
    rng = np.random.default_rng(313)
    x = data(np.uint8, (64, 100, 3), rng)
    check(x, 1.0, ["reflect"] * 3, monkeypatch, view=(slice(None, None, -1), slice(None), slice(None)))
