"""GPU parity, through the C ABI, for the filters next to the convolution
path: erode / dilate (processing/morphology.h:57-101; ndl_morph_separable,
ndl_morph_nd), downsample (convolution.h:241-305; ndl_downsample) and the
multi-sigma scale stack with its difference-of-Gaussians layers
(feature_detection.h:91-120; ndl_gaussian_scales, ndl_dog), against the CPU
oracle.  The bar is bit-exact for every element type (NaN payloads and signed
zeros included) except FP32 accumulation, which is held to 1e-6 relative."""
import itertools
import sys

import numpy as np
import pytest

import ndl_amd
import oracle

pytestmark = pytest.mark.gpu

GRID = np.arange(1, 26, dtype=np.int32).reshape(5, 5)
BOX3 = np.ones((3, 3))
CROSS3 = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=np.float64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_reference_unit_values():
    # image_morphology_tests.cpp:33-44 and :56-71
    say("morph reference values")
    assert ndl_amd.morph(GRID, BOX3, "erode")[2, 2] == 7
    assert ndl_amd.morph(GRID, BOX3, "dilate")[2, 2] == 19
    assert ndl_amd.morph(GRID, CROSS3, "erode")[2, 2] == 8
    assert ndl_amd.morph(GRID, CROSS3, "dilate")[2, 2] == 18
    dot = np.zeros((5, 5), dtype=np.int32)
    dot[2, 2] = 1
    assert int(ndl_amd.morph(dot, BOX3, "dilate").sum()) == 9
    assert int(ndl_amd.morph(dot, CROSS3, "dilate").sum()) == 5


def _data(dtype, shape, rng, specials=True):
    if np.issubdtype(dtype, np.floating):
        x = rng.standard_normal(shape).astype(dtype)
        if specials:
            flat = x.reshape(-1)
            idx = rng.choice(flat.size, size=max(3, flat.size // 15), replace=False)
            k = len(idx) // 3
            flat[idx[:k]] = np.nan
            flat[idx[k:2 * k]] = 0.0
            flat[idx[2 * k:]] = -0.0
        return x
    info = np.iinfo(dtype)
    return rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.int32, np.float32, np.float64,
                                   np.float16])
def test_separable_box_vs_oracle(dtype):
    rng = np.random.default_rng(41)
    cases = [((37, 53), (5, 7)), ((19, 23, 29), (3, 5, 7)), ((300,), (9,)), ((8, 70, 130), (1, 9, 1)),
             ((6, 40), (13, 3))]
    for (shape, ksize), border, op in itertools.product(cases, ["clamp", "wrap", "reflect"], ["erode", "dilate"]):
        if border != "clamp" and any(k // 2 > s for k, s in zip(ksize, shape)):
            continue
        x = _data(dtype, shape, rng)
        k = np.ones(ksize)
        say("morph separable", np.dtype(dtype).name, shape, ksize, border, op)
        want = oracle.morph(x, k, op, border)
        got = ndl_amd.morph(x, k, op, border)
        assert same_bits(got, want), (np.dtype(dtype).name, shape, ksize, border, op)


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float32, np.float64])
def test_nd_masks_vs_oracle(dtype):
    rng = np.random.default_rng(42)
    masks = [CROSS3, (rng.random((5, 3)) < 0.5).astype(float), np.ones((3, 3)), np.zeros((3, 3)),
             np.array([[0, 0, 1], [0, 0, 0], [0, 0, 0]], dtype=float)]
    for k, border, op in itertools.product(masks, ["clamp", "wrap", "reflect"], ["erode", "dilate"]):
        x = _data(dtype, (29, 31), rng)
        say("morph nd", np.dtype(dtype).name, k.shape, border, op)
        want = oracle.morph(x, k, op, border)
        assert same_bits(ndl_amd.morph(x, k, op, border, force_nd=True), want)
        # the same element through the separable path when it is a product
        assert same_bits(ndl_amd.morph(x, k, op, border), want)


def test_3d_cross_and_product_masks():
    rng = np.random.default_rng(43)
    jack = np.zeros((3, 3, 3))
    jack[1, 1, :] = jack[1, :, 1] = jack[:, 1, 1] = 1  # make_cross_kernel in 3-D: 7 taps
    prod = np.multiply.outer(np.multiply.outer([1, 0, 1], [0, 1, 1, 1]), [1, 1, 0, 1, 1]).astype(float)
    x = _data(np.float32, (12, 17, 21), rng)
    for k, op in itertools.product([jack, prod], ["erode", "dilate"]):
        say("morph 3d", k.shape, op)
        assert same_bits(ndl_amd.morph(x, k, op, "reflect"), oracle.morph(x, k, op, "reflect"))


def test_cast_to_other_destination_type():
    # static_cast<DstT>(best): float -> int32 truncates, int32 -> uint8 keeps the low bits
    rng = np.random.default_rng(44)
    x = (rng.standard_normal((31, 37)) * 1000).astype(np.float32)
    want = oracle.morph(x, np.ones((3, 5)), "dilate", dst=np.zeros(x.shape, np.int32))
    assert np.array_equal(ndl_amd.morph(x, np.ones((3, 5)), "dilate", out_dtype=np.int32), want)
    xi = rng.integers(-100000, 100000, size=(31, 37)).astype(np.int32)
    want = oracle.morph(xi, np.ones((5, 3)), "erode", dst=np.zeros(xi.shape, np.uint8))
    assert np.array_equal(ndl_amd.morph(xi, np.ones((5, 3)), "erode", out_dtype=np.uint8), want)


def test_views():
    # strided sources; per-axis borders (the oracle applies one border, so
    # the reference operator is two one-axis passes, exact without NaN)
    rng = np.random.default_rng(45)
    x = _data(np.float32, (40, 50), rng, specials=False)
    d = ndl_amd.DeviceArray.from_numpy(x)
    for view in [(slice(3, 33, 2), slice(None, None, -1)), (slice(None), slice(5, 45, 3))]:
        say("morph view", view)
        src = d.view(view)
        out = ndl_amd.DeviceArray(src.shape, np.float32)
        ndl_amd.morph_separable_dev(src, out, [0, 1], [np.ones(5), np.ones(3)], ["reflect", "clamp"], "erode")
        dense = np.ascontiguousarray(x[view])
        want = oracle.morph(oracle.morph(dense, np.ones((1, 5)), "erode", "reflect"), np.ones((3, 1)), "erode",
                            "clamp")
        assert same_bits(out.to_numpy(), want)


def test_open_close_idempotent():
    rng = np.random.default_rng(46)
    x = _data(np.uint8, (64, 80), rng)
    k = np.ones((5, 5))
    opened = ndl_amd.morph(ndl_amd.morph(x, k, "erode"), k, "dilate")
    closed = ndl_amd.morph(ndl_amd.morph(x, k, "dilate"), k, "erode")
    assert np.all(opened <= x) and np.all(closed >= x)
    assert np.array_equal(ndl_amd.morph(ndl_amd.morph(opened, k, "erode"), k, "dilate"), opened)
    assert np.array_equal(ndl_amd.morph(ndl_amd.morph(closed, k, "dilate"), k, "erode"), closed)


# ------------------------------------------------------------- downsample
@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.int32, np.float32, np.float64,
                                   np.float16])
def test_downsample_vs_oracle(dtype):
    rng = np.random.default_rng(53)
    for shape, f, ca in [((30, 40, 3), 2, 0), ((13, 17), 3, 1), ((5, 9, 11), 2, 0), ((3, 20, 21), 4, 2),
                         ((6, 7, 8, 9), 3, 1), ((2, 31), 5, 1), ((256, 257), 2, 1)]:
        if np.issubdtype(dtype, np.floating):
            x = (rng.standard_normal(shape) * 100).astype(dtype)
        else:
            info = np.iinfo(dtype)
            x = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
        say("downsample", np.dtype(dtype).name, shape, f, ca)
        want = oracle.downsample(x, f, ca)
        got = ndl_amd.downsample(x, f, ca)
        assert same_bits(got, want), (np.dtype(dtype).name, shape, f, ca)


@pytest.mark.parametrize("fast", ["1", "0"])
def test_downsample_unrolled_and_generic_kernels(fast, monkeypatch):
    # blocks of at most 16 positions take the unrolled kernel (NDL_DOWN_FAST=1),
    # larger ones the generic walk; both match the oracle bit for bit
    monkeypatch.setenv("NDL_DOWN_FAST", fast)
    rng = np.random.default_rng(58)
    for dtype in (np.uint8, np.int16, np.float32, np.float64):
        for shape, f, ca in [((31, 45, 3), 2, 0), ((9, 10, 11, 12), 2, 3), ((37, 41), 4, 1), ((5, 50, 51), 3, 2),
                             ((13, 1, 17), 2, 0), ((7, 8, 9, 10), 3, 0)]:
            if np.issubdtype(dtype, np.floating):
                x = (rng.standard_normal(shape) * 100).astype(dtype)
            else:
                info = np.iinfo(dtype)
                x = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
            np.testing.assert_array_equal(ndl_amd.downsample(x, f, ca).view(np.uint8),
                                          oracle.downsample(x, f, ca).view(np.uint8))


def test_downsample_reference_constant_colour():
    # image_convolution_tests.cpp:70-91
    x = np.full((30, 40, 3), 77, dtype=np.uint8)
    out = ndl_amd.downsample(x, 2)
    assert out.shape == (15, 20, 3) and np.all(out == 77)


def test_downsample_strided_source():
    rng = np.random.default_rng(54)
    x = rng.standard_normal((40, 50)).astype(np.float32)
    d = ndl_amd.DeviceArray.from_numpy(x)
    src = d.view((slice(None, None, -1), slice(3, 48, 2)))
    dense = np.ascontiguousarray(x[::-1, 3:48:2])
    want = oracle.downsample(dense, 3, 1)
    out = ndl_amd.DeviceArray(want.shape, np.float32)
    ndl_amd.downsample_dev(src, out, 3, 1)
    assert same_bits(out.to_numpy(), want)


# ----------------------------------------------------------- scale stacks
def _dog_sigmas(num_scales=5, sigma0=1.6):
    # detect_keypoints: sigma_i = sigma0 * 2^(i / (numScales - 3)), feature_detection.h:101-107
    return [sigma0 * 2.0 ** (i / (num_scales - 3)) for i in range(num_scales)]


@pytest.mark.parametrize("shape", [(70, 90), (20, 33, 41)])
def test_scale_stack_and_dog_exact(shape):
    rng = np.random.default_rng(55)
    x = rng.random(shape) * 255.0  # detect_keypoints blurs a double copy of the image
    sig = _dog_sigmas()
    say("scale stack", shape)
    s = ndl_amd.DeviceArray.from_numpy(x)
    levels = [ndl_amd.DeviceArray(shape, np.float64) for _ in sig]
    ndl_amd.gaussian_scales_dev(s, levels, sig, "reflect")
    want = [oracle.gaussian_blur(x, sg, "reflect") for sg in sig]
    for lv, w in zip(levels, want):
        assert same_bits(lv.to_numpy(), w)
    dogs = [ndl_amd.DeviceArray(shape, np.float64) for _ in sig[1:]]
    ndl_amd.dog_dev(levels, dogs)
    for i, d in enumerate(dogs):
        assert same_bits(d.to_numpy(), want[i + 1] - want[i])


def test_scale_stack_wide_sigmas_exact():
    # detect_keypoints' octaves reach large sigmas; 12.0 and 20.0 give 73 and
    # 121 taps (past the 64 carried in kernel arguments), beside a narrow level
    rng = np.random.default_rng(57)
    x = rng.random((70, 130)) * 255.0
    sig = [1.6, 12.0, 20.0]
    say("scale stack, wide sigmas")
    s = ndl_amd.DeviceArray.from_numpy(x)
    levels = [ndl_amd.DeviceArray(x.shape, np.float64) for _ in sig]
    ndl_amd.gaussian_scales_dev(s, levels, sig, "reflect")
    for lv, sg in zip(levels, sig):
        assert same_bits(lv.to_numpy(), oracle.gaussian_blur(x, sg, "reflect"))


def test_scale_stack_uint8_and_fp32():
    rng = np.random.default_rng(56)
    x = rng.integers(0, 256, size=(64, 80)).astype(np.uint8)
    sig = _dog_sigmas()[:4]
    s = ndl_amd.DeviceArray.from_numpy(x)
    levels = [ndl_amd.DeviceArray(x.shape, np.uint8) for _ in sig]
    ndl_amd.gaussian_scales_dev(s, levels, sig, "clamp", rounding=ndl_amd.abi.ROUND_NEAREST)
    for lv, sg in zip(levels, sig):
        assert np.array_equal(lv.to_numpy(), oracle.gaussian_blur(x, sg, "clamp"))
    xf = rng.random((40, 50, 60)).astype(np.float32)
    s = ndl_amd.DeviceArray.from_numpy(xf)
    levels = [ndl_amd.DeviceArray(xf.shape, np.float32) for _ in sig]
    ndl_amd.gaussian_scales_dev(s, levels, sig, "reflect", accum=ndl_amd.abi.ACCUM_F32)
    for lv, sg in zip(levels, sig):
        want = oracle.gaussian_blur(xf.astype(np.float64), sg, "reflect")
        assert np.max(np.abs(lv.to_numpy() - want) / np.maximum(np.abs(want), 1e-30)) <= 1e-6


MORPH_ONE_AXIS = [
    # (dtype, shape, box): the passes before the last (every pass for integer
    # images) run on the row kernel / line march
    (np.float32, (30, 20, 260), (5, 3, 7)),
    (np.float32, (300, 260), (5, 9)),
    (np.uint8, (3, 300, 700), (1, 7, 7)),
    (np.uint8, (40, 33, 96), (3, 11, 5)),
    # interleaved colour {3, x, y}: lanes over channel and pixel
    (np.uint8, (120, 200, 3), (5, 5, 1)),
    (np.float32, (90, 150, 3), (3, 7, 1)),
]


@pytest.mark.parametrize("op", ["erode", "dilate"])
@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
@pytest.mark.parametrize("case", range(len(MORPH_ONE_AXIS)))
def test_morph_box_one_axis_kernels(case, border, op, monkeypatch):
    dtype, shape, box = MORPH_ONE_AXIS[case]
    rng = np.random.default_rng(500 + case)
    x = _data(dtype, shape, rng)
    mask = np.ones(box)
    say("morph one axis", dtype.__name__, shape, box, border, op)
    want = oracle.morph(x, mask, op, border)
    got = ndl_amd.morph(x, mask, op, border)
    assert same_bits(got, want)
    monkeypatch.setenv("NDL_LINE_MARCH", "0")
    monkeypatch.setenv("NDL_ROW_PASS", "0")
    assert same_bits(ndl_amd.morph(x, mask, op, border), want)
