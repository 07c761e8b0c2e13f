"""GPU parity of the fused blur-then-downsample call (ndl_gaussian_downsample)
against the oracle's composition downsample(gaussian_blur(x, sigma, border),
factor, channelAxis) (convolution.h:107-186 then :241-305), through the C ABI.

The bar is bit-exact for the reference arithmetic (every element type, the
blurred image held in the destination's type with gaussian_blur's rounding),
and 1e-6 relative for FP32 accumulation on positive data.  Cases cover colour
images with the channel axis first, 2-D and 3-D images, factors 2..4 with
ragged last blocks, every border, per-axis sigmas (an unblurred channel axis,
a single blurred axis read straight from the source), wide kernels (taps
beyond those carried in kernel arguments) and strided sources.  Each case
also matches the two separate GPU calls bit for bit; the march kernel
(lanes marching along the last blurred axis) matches the generic one, and
the two-pass kernel with block means (the last two blurred axes, factor 2 or
4) matches the march.
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


def same_bits(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def data(dtype, shape, rng):
    if np.issubdtype(dtype, np.floating):
        return (rng.random(shape) * 200.0 - 50.0).astype(dtype)
    info = np.iinfo(dtype)
    lo, hi = max(info.min, -30000), min(info.max, 60000)
    return rng.integers(lo, hi, size=shape, endpoint=True).astype(dtype)


def two_gpu_calls(x, sigma, factor, ca, border):
    """gaussian_blur then downsample as two separate GPU calls."""
    return ndl_amd.downsample(ndl_amd.gaussian_blur(x, sigma, border, accum=REF), factor, ca)


CASES = [
    # (numpy shape, sigma, factor, channel axis (ndl), border)
    ((30, 40, 3), 1.0, 2, 0, "clamp"),
    ((31, 45, 3), 1.0, 2, 0, "reflect"),
    ((29, 37, 3), 0.8, 3, 0, "wrap"),
    ((64, 80), 1.6, 2, 1, "reflect"),
    ((63, 81), 2.0, 4, 0, "clamp"),
    ((20, 33, 41), 1.2, 2, 0, "reflect"),
    ((19, 34, 40), 1.0, 3, 2, "wrap"),
    ((5, 9, 11, 6), 0.7, 2, 1, "clamp"),
]


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.int32, np.float32, np.float64])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-s{c[1]}-f{c[2]}-ca{c[3]}-{c[4]}")
def test_gaussian_downsample_bit_exact(case, dtype):
    shape, sigma, f, ca, border = case
    rng = np.random.default_rng(abs(hash((shape, f))) % 2**32)
    x = data(dtype, shape, rng)
    say("blur-downsample", np.dtype(dtype).name, shape, sigma, f, ca, border)
    want = oracle.downsample(oracle.gaussian_blur(x, sigma, border), f, ca)
    got = ndl_amd.gaussian_downsample(x, sigma, f, ca, border)
    assert same_bits(got, want)
    assert same_bits(got, two_gpu_calls(x, sigma, f, ca, border))


def test_reference_constant_colour_stays_constant():
    # image_convolution_tests.cpp:70-91 with a blur in front: a constant image
    # is a fixed point of both steps
    x = np.full((30, 40, 3), 77, dtype=np.uint8)
    out = ndl_amd.gaussian_downsample(x, 1.5, 2)
    assert out.shape == (15, 20, 3) and np.all(out == 77)


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_channel_axis_not_blurred(dtype):
    # sigma 0 on the channel axis: the spatial axes only, in axis order
    rng = np.random.default_rng(61)
    x = data(dtype, (33, 47, 3), rng)
    taps = oracle.gaussian_taps(1.3)
    rounding = NEAR if np.issubdtype(dtype, np.integer) else TRUNC
    blurred = oracle.sepconv(x, [1, 2], [taps, taps], ["reflect", "reflect"], rounding)
    want = oracle.downsample(blurred, 2, 0)
    s = ndl_amd.DeviceArray.from_numpy(x)
    d = ndl_amd.DeviceArray(want.shape, dtype)
    ndl_amd.gaussian_downsample_dev(s, d, [0.0, 1.3, 1.3], 2, 0, "reflect", REF, rounding)
    assert same_bits(d.to_numpy(), want)


@pytest.mark.parametrize("axis", [0, 1])
def test_single_blurred_axis_reads_the_source(axis):
    # one blurred axis: the fused pass reads the uint8 source itself, and the
    # blur store truncates (a single pass, as convolve())
    rng = np.random.default_rng(62)
    x = data(np.uint8, (50, 70), rng)
    taps = oracle.gaussian_taps(1.1)
    blurred = oracle.sepconv(x, [axis], [taps], ["clamp"], TRUNC)
    want = oracle.downsample(blurred, 3, 1 - axis)
    sg = [0.0, 0.0]
    sg[axis] = 1.1
    s = ndl_amd.DeviceArray.from_numpy(x)
    d = ndl_amd.DeviceArray(want.shape, np.uint8)
    ndl_amd.gaussian_downsample_dev(s, d, sg, 3, 1 - axis, "clamp", REF, TRUNC)
    assert same_bits(d.to_numpy(), want)


def test_wide_sigma_taps_on_the_device():
    # sigma 12 gives 73 taps: the fused pass reads them from device memory
    rng = np.random.default_rng(63)
    x = rng.random((90, 130)) * 255.0
    want = oracle.downsample(oracle.gaussian_blur(x, 12.0, "reflect"), 2, 0)
    got = ndl_amd.gaussian_downsample(x, 12.0, 2, 0, "reflect")
    assert same_bits(got, want)


def test_strided_source_view():
    rng = np.random.default_rng(64)
    full = data(np.uint8, (40, 60, 3), rng)
    view = (slice(None, None, -1), slice(4, 56, 2), slice(None))
    dense = np.ascontiguousarray(full[view])
    want = oracle.downsample(oracle.gaussian_blur(dense, 1.0, "clamp"), 2, 0)
    s = ndl_amd.DeviceArray.from_numpy(full).view(view)
    d = ndl_amd.DeviceArray(want.shape, np.uint8)
    ndl_amd.gaussian_downsample_dev(s, d, 1.0, 2, 0, "clamp", REF, NEAR)
    assert same_bits(d.to_numpy(), want)


def test_fp32_accumulation():
    rng = np.random.default_rng(65)
    x = rng.random((40, 50, 60)).astype(np.float32)
    want = oracle.downsample(oracle.gaussian_blur(x.astype(np.float64), 1.5, "reflect"), 2, 0)
    s = ndl_amd.DeviceArray.from_numpy(x)
    d = ndl_amd.DeviceArray(want.shape, np.float32)
    ndl_amd.gaussian_downsample_dev(s, d, 1.5, 2, 0, "reflect", F32, TRUNC)
    got = d.to_numpy().astype(np.float64)
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-30)) <= 1e-6


def test_argument_errors():
    x = ndl_amd.DeviceArray((10, 12), np.uint8)
    good = ndl_amd.DeviceArray((5, 12), np.uint8)
    with pytest.raises(ndl_amd.NdlError):
        ndl_amd.gaussian_downsample_dev(x, ndl_amd.DeviceArray((6, 12), np.uint8), 1.0, 2, 0)  # wrong extent
    with pytest.raises(ndl_amd.NdlError):
        ndl_amd.gaussian_downsample_dev(x, good, 1.0, 1, 0)  # factor < 2
    with pytest.raises(ndl_amd.NdlError):
        ndl_amd.gaussian_downsample_dev(x, good, [0.0, 0.0], 2, 0)  # nothing blurred
    with pytest.raises(ndl_amd.NdlError):
        ndl_amd.gaussian_downsample_dev(x, good, 1.0, 2, 0, accum=F32)  # FP32 on uint8
    with pytest.raises(ndl_amd.NdlError):
        ndl_amd.gaussian_downsample_dev(x, good, 6.0, 2, 0, "reflect")  # taps reach past 12


@pytest.mark.parametrize("case", CASES[:4] + CASES[6:], ids=lambda c: f"{c[0]}-f{c[2]}-ca{c[3]}")
def test_generic_kernel_matches_the_march(case, monkeypatch):
    # the cases above run the march kernel (lanes marching along the last
    # blurred axis); the generic one-output-per-thread kernel must agree
    shape, sigma, f, ca, border = case
    rng = np.random.default_rng(66)
    x = data(np.uint8, shape, rng)
    a = ndl_amd.gaussian_downsample(x, sigma, f, ca, border)
    monkeypatch.setenv("NDL_DISABLE_BLUR_DOWN_MARCH", "1")
    b = ndl_amd.gaussian_downsample(x, sigma, f, ca, border)
    assert same_bits(a, b)
    assert same_bits(a, oracle.downsample(oracle.gaussian_blur(x, sigma, border), f, ca))


PAIR_CASES = [
    # the last two blurred axes with the block means in one launch (factor 2
    # or 4, every other axis the channel axis): colour images after their
    # channel pass, 2-D images with one axis kept
    ((30, 40, 3), 1.0, 2, 0, "clamp"),
    ((31, 45, 3), 1.0, 2, 0, "reflect"),
    ((70, 90, 3), 1.0, 4, 0, "wrap"),
    ((64, 80), 1.6, 2, 1, "reflect"),
    ((63, 81), 2.0, 4, 0, "clamp"),
    ((100, 47), 0.9, 2, 1, "wrap"),
]


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32, np.float64])
@pytest.mark.parametrize("case", PAIR_CASES, ids=lambda c: f"{c[0]}-f{c[2]}-ca{c[3]}-{c[4]}")
def test_two_pass_block_means_match(case, dtype, monkeypatch):
    shape, sigma, f, ca, border = case
    rng = np.random.default_rng(67)
    x = data(dtype, shape, rng)
    say("blur-downsample two-pass", np.dtype(dtype).name, shape, f, ca, border)
    a = ndl_amd.gaussian_downsample(x, sigma, f, ca, border)
    monkeypatch.setenv("NDL_XY_DOWN", "0")
    b = ndl_amd.gaussian_downsample(x, sigma, f, ca, border)
    assert same_bits(a, b)
    assert same_bits(a, oracle.downsample(oracle.gaussian_blur(x, sigma, border), f, ca))


def test_two_pass_block_means_strided_source():
    rng = np.random.default_rng(68)
    full = data(np.uint8, (90, 140), rng)
    view = (slice(None, None, -1), slice(3, 133))
    dense = np.ascontiguousarray(full[view])
    want = oracle.downsample(oracle.gaussian_blur(dense, 1.2, "reflect"), 2, 1)
    s = ndl_amd.DeviceArray.from_numpy(full).view(view)
    d = ndl_amd.DeviceArray(want.shape, np.uint8)
    ndl_amd.gaussian_downsample_dev(s, d, 1.2, 2, 1, "reflect", REF, NEAR)
    assert same_bits(d.to_numpy(), want)
