"""Pins the CPU oracle (oracle/ndl_oracle.c) -- CPU only.

Each check is chosen so that a plausible restatement mistake (a wrong border
fold, an off-by-one centre, a transposed axis, an FMA-contracted sum, a
dropped rounding step) fails at least one of them:
  * literal values from the reference's own tests and tutorial snapshots;
  * fixtures generated from the compiled reference (tests/golden);
  * brute force, scipy.ndimage.correlate1d and FFT identities on small inputs;
  * dense-Jacobian and dot-product identities for the adjoints;
  * the compiled reference itself, when oracle/_ref is present.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle
from ndl_amd import abi
from tests.golden.make_golden import all_fixtures, load


def grid5():
    # the tutorial's 5x5 grid: value = 5*y + x + 1 (convolution_tutorial.md:50)
    return np.arange(1, 26, dtype=np.float64).reshape(5, 5)


# ---------------------------------------------- reference tests and tutorial
def test_tutorial_sum_kernel_table():
    # convolution_tutorial.md:61-66 (out2, default Clamp)
    expect = np.array([[27, 33, 42, 51, 57], [57, 63, 72, 81, 87], [102, 108, 117, 126, 132],
                       [147, 153, 162, 171, 177], [177, 183, 192, 201, 207]], dtype=np.float64)
    np.testing.assert_array_equal(oracle.convolve(grid5(), np.ones((3, 3))), expect)


def test_tutorial_corner_breakdowns():
    # convolution_tutorial.md:83-85: clamp 27, wrap 99, reflect 27
    g = grid5()
    assert oracle.convolve(g, np.ones((3, 3)), "clamp")[0, 0] == 27
    assert oracle.convolve(g, np.ones((3, 3)), "wrap")[0, 0] == 99
    assert oracle.convolve(g, np.ones((3, 3)), "reflect")[0, 0] == 27


def test_tutorial_gaussian_impulse():
    # convolution_tutorial.md:138-149: sigma=1, 11x11 impulse of 1000, 2 decimals
    imp = np.zeros((11, 11))
    imp[5, 5] = 1000.0
    r = oracle.gaussian_blur(imp, 1.0)
    table = {(5, 5): 159.24, (5, 4): 96.58, (4, 4): 58.58, (5, 3): 21.55, (4, 3): 13.07,
             (3, 3): 2.92, (5, 2): 1.77, (4, 2): 1.07, (3, 2): 0.24, (2, 2): 0.02, (5, 1): 0.00}
    for (y, x), v in table.items():
        assert round(r[y, x], 2) == v, (y, x, r[y, x])
    np.testing.assert_allclose(r, r.T, rtol=0, atol=1e-12)


def test_unit_image_convolution():
    # image_convolution_tests.cpp:13-67 (4x4 ints 1..16)
    img = np.arange(1, 17, dtype=np.int32).reshape(4, 4)
    ident = np.zeros((3, 3))
    ident[1, 1] = 1
    np.testing.assert_array_equal(oracle.convolve(img, ident), img)
    s = oracle.convolve(img, np.ones((3, 3)), "clamp")
    assert s[1, 1] == 1 + 2 + 3 + 5 + 6 + 7 + 9 + 10 + 11
    assert s[0, 0] == 4 * img[0, 0] + 2 * img[0, 1] + 2 * img[1, 0] + img[1, 1] == 24
    assert oracle.convolve(img, np.ones((3, 3)), "wrap")[0, 0] != s[0, 0]
    c5 = oracle.convolve(img, np.ones((5, 5)), "clamp")[0, 0]
    r5 = oracle.convolve(img, np.ones((5, 5)), "reflect")[0, 0]
    assert c5 != r5


def test_unit_gradient_kernel():
    # image_convolution_tests.cpp:117-160: {-0.5,0,0.5} under Reflect
    ramp = np.tile(np.arange(10, dtype=np.float64), (4, 1))
    out = oracle.sepconv(ramp, [0], [np.array([-0.5, 0, 0.5])], ["reflect"])
    np.testing.assert_array_equal(out[:, 1:-1], 1.0)
    y, x = np.mgrid[0:6, 0:7].astype(np.float64)
    plane = 2 * x + 3 * y
    gx = oracle.sepconv(plane, [0], [np.array([-0.5, 0, 0.5])], ["reflect"])
    gy = oracle.sepconv(plane, [1], [np.array([-0.5, 0, 0.5])], ["reflect"])
    np.testing.assert_allclose(gx[1:-1, 1:-1], 2.0, atol=1e-9)
    np.testing.assert_allclose(gy[1:-1, 1:-1], 3.0, atol=1e-9)


def test_unit_wrap_sum_identity():
    # composite_tests.cpp:105-133: sum(out) == sum(kernel)*sum(img) under Wrap
    img = np.arange(1, 21, dtype=np.int32).reshape(4, 5)
    k = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], dtype=np.float64)
    out = oracle.convolve(img, k, "wrap")
    assert out.sum() == k.sum() * img.sum()


def test_unit_float_average():
    # type_coverage_tests.cpp:141-149: double accumulation, float output
    a = np.array([3, 5, 7, 9], dtype=np.float32)
    out = oracle.convolve(a, np.array([1 / 3, 1 / 3, 1 / 3]), "clamp")
    assert out[1] == np.float32((3.0 + 5.0 + 7.0) / 3.0)


def test_sat_boxblur_border_cases():
    # summed_area_table_tests.cpp:119-135: same border semantics in 1-D
    ramp = np.arange(20, dtype=np.float64)
    box = np.full(5, 1 / 5)
    c = oracle.sepconv(ramp, [0], [box], ["clamp"])
    w = oracle.sepconv(ramp, [0], [box], ["wrap"])
    assert abs(c[0] - (3 * ramp[0] + ramp[1] + ramp[2]) / 5) < 1e-9
    assert abs(w[0] - (ramp[18] + ramp[19] + ramp[0] + ramp[1] + ramp[2]) / 5) < 1e-9


def test_fft_circular_correlation():
    # fft_tests.cpp:164-223: wrap convolve == circular cross-correlation
    rng = np.random.default_rng(1)
    img = rng.standard_normal((8, 16))
    k = np.array([1, 2, 1, 0, 1, 0, -1, 0, 2], dtype=np.float64).reshape(3, 3)
    out = oracle.convolve(img, k, "wrap")
    kp = np.zeros_like(img)
    for (i, j), v in np.ndenumerate(k):
        kp[(i - 1) % 8, (j - 1) % 16] = v
    fft = np.real(np.fft.ifft2(np.fft.fft2(img) * np.conj(np.fft.fft2(kp))))
    np.testing.assert_allclose(out, fft, atol=1e-9)


# ------------------------------------------------------ compiled-reference
@pytest.mark.parametrize("path", all_fixtures(), ids=lambda p: p.split("/")[-1][:-4])
def test_golden_fixture(path):
    x, out, m = load(path)
    if m["kind"] == "gaussian":
        got = oracle.gaussian_blur(x, m["sigma"], m["border"])
    elif m["kind"] == "convolve":
        got = oracle.convolve(x, m["kernel"], m["border"])
    else:
        got = oracle.sepconv(x, m["axes"], m["taps"], m["borders"], m["rounding"])
    assert got.dtype == out.dtype
    np.testing.assert_array_equal(got.view(np.uint8), out.view(np.uint8))


@pytest.mark.skipif(oracle.ref_lib is None, reason="reference sources absent: oracle/_ref not built")
@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.int32, np.float32, np.float64])
def test_oracle_equals_compiled_reference(dtype):
    rng = np.random.default_rng(7)
    for trial in range(4):
        shape = tuple(int(s) for s in rng.integers(3, 14, size=trial % 3 + 1))
        if np.issubdtype(dtype, np.integer):
            info = np.iinfo(dtype)
            x = rng.integers(max(info.min, -5000), min(info.max, 5000), size=shape).astype(dtype)
        else:
            x = (rng.standard_normal(shape) * 10).astype(dtype)
        nd = len(shape)
        for b in ["clamp", "wrap", "reflect"]:
            sigma = 0.6 + 0.3 * trial
            r = oracle.kernel_size(sigma)
            if b != "clamp" and r > min(shape):
                continue
            np.testing.assert_array_equal(oracle.gaussian_blur(x, sigma, b),
                                          oracle.gaussian_blur(x, sigma, b, use_ref=True))
            taps = [rng.standard_normal(2 * int(rng.integers(0, 3)) + 1) for _ in range(nd)]
            if any(b != "clamp" and len(t) // 2 > s for t, s in zip(taps, reversed(shape))):
                continue
            for rnd in (abi.ROUND_TRUNC, abi.ROUND_NEAREST):
                np.testing.assert_array_equal(
                    oracle.sepconv(x, list(range(nd)), taps, [b] * nd, rnd),
                    oracle.sepconv(x, list(range(nd)), taps, [b] * nd, rnd, use_ref=True))
            k = rng.integers(-2, 3, size=tuple(3 for _ in range(nd))).astype(np.float64)
            np.testing.assert_array_equal(oracle.convolve(x, k, b), oracle.convolve(x, k, b, use_ref=True))


@pytest.mark.skipif(oracle.ref_lib is None, reason="reference sources absent: oracle/_ref not built")
def test_kernel_size_equals_reference():
    for s in [0.1, 0.5, 1 / 3, 2 / 3, 1.0, 1.5, 2.0, 3.0, 3.9999, 4.0, 0.33333333333]:
        assert oracle.kernel_size(s) == oracle.ref_lib.ref_kernel_size(s)


# ------------------------------------------------------ textbook identities
@pytest.mark.parametrize("border,mode", [("clamp", "nearest"), ("wrap", "wrap"), ("reflect", "reflect")])
def test_matches_scipy_correlate1d(border, mode):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((9, 23))
    for k in [3, 5, 9, 13]:
        w = rng.standard_normal(k)
        got = oracle.sepconv(x, [0], [w], [border])
        want = ndi.correlate1d(x, w, axis=1, mode=mode)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
        got = oracle.sepconv(x, [1], [w], [border])
        want = ndi.correlate1d(x, w, axis=0, mode=mode)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_brute_force_even_width_and_offset_center():
    rng = np.random.default_rng(4)
    x = rng.standard_normal(17)
    for K, c in [(4, 2), (4, 0), (6, 5), (2, 1)]:
        w = rng.standard_normal(K)
        for b, f in [("clamp", lambda i: min(max(i, 0), 16)), ("wrap", lambda i: i % 17),
                     ("reflect", lambda i: -i - 1 if i < 0 else (33 - i if i >= 17 else i))]:
            got = oracle.sepconv(x, [0], [w], [b], centers=[c])
            want = [sum(w[t] * x[f(i + t - c)] for t in range(K)) for i in range(17)]
            np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_separable_equals_outer_product_kernel():
    rng = np.random.default_rng(5)
    x = rng.random((12, 15))
    g = oracle.gaussian_taps(1.2)
    sep = oracle.sepconv(x, [0, 1], [g, g], ["reflect"] * 2)
    full = oracle.convolve(x, np.outer(g, g), "reflect")
    np.testing.assert_allclose(sep, full, rtol=1e-13)


def test_gaussian_taps_closed_form():
    for s in [0.5, 1.0, 1.5, 2.0, 3.0, 4.0]:
        t = oracle.gaussian_taps(s)
        r = len(t) // 2
        assert r == int(np.ceil(3 * s))
        d = np.arange(-r, r + 1)
        w = np.exp(-d * d / (2 * s * s))
        np.testing.assert_allclose(t, w / w.sum(), rtol=1e-15)
        assert abs(t.sum() - 1) < 1e-15 and np.array_equal(t, t[::-1])


def test_narrowing_matches_x86_cvttsd2si():
    # convolution.h:69 narrowing as compiled for x86-64 (SURVEY a10 probe values)
    x = np.array([[12495.0, 300.7, -3.5, 255.9]])
    out = oracle.sepconv(x, [0], [np.array([1.0])], ["clamp"], dst=np.zeros((1, 4), np.uint8))
    np.testing.assert_array_equal(out, [[207, 44, 253, 255]])
    assert oracle.lib.oracle_cvtt_i32(3e9) == -2**31
    assert oracle.lib.oracle_cvtt_i32(float("nan")) == -2**31
    assert oracle.lib.oracle_cvtt_i32(-2147483648.9) == -2**31


def test_u8_box_is_window_sum_mod_256():
    rng = np.random.default_rng(6)
    x = rng.integers(0, 256, size=(3, 20, 24), dtype=np.uint8)
    box = np.ones(7)
    out = oracle.sepconv(x, [0, 1], [box, box], ["wrap", "wrap"])
    xi = x.astype(np.int64)
    s = sum(np.roll(np.roll(xi, dy, axis=1), dx, axis=2) for dy in range(-3, 4) for dx in range(-3, 4))
    np.testing.assert_array_equal(out, (s % 256).astype(np.uint8))


def test_strided_view_equals_dense_copy():
    rng = np.random.default_rng(8)
    base = rng.standard_normal((10, 12, 14))
    for view in [base[::-1], base[:, ::2, :], base[2:9, 1:, ::-3], base.transpose(0, 2, 1), base[4]]:
        nd = view.ndim
        taps = [oracle.gaussian_taps(0.8)] * nd
        a = oracle.sepconv(view, list(range(nd)), taps, ["reflect"] * nd)
        b = oracle.sepconv(np.ascontiguousarray(view), list(range(nd)), taps, ["reflect"] * nd)
        np.testing.assert_array_equal(a, b)


def test_radius_beyond_extent_is_refused():
    x = np.zeros((4, 5))
    with pytest.raises(oracle.OracleError):
        oracle.sepconv(x, [0], [np.ones(13)], ["wrap"])
    oracle.sepconv(x, [0], [np.ones(13)], ["clamp"])  # clamp is defined for any reach


# ------------------------------------------------------------- adjoints
def dense_matrix(shape, axes, taps, borders, centers=None):
    n = int(np.prod(shape))
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1
        A[:, j] = oracle.sepconv(e.reshape(shape), axes, taps, borders, centers=centers).ravel()
    return A


@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
def test_adjoint_input_is_dense_transpose(border):
    rng = np.random.default_rng(9)
    shape = (5, 6, 7)
    taps = [rng.standard_normal(k) for k in (3, 5, 4)]
    centers = [1, 2, 1]
    axes = [0, 1, 2]
    A = dense_matrix(shape, axes, taps, [border] * 3, centers)
    g = rng.standard_normal(shape)
    got = oracle.adjoint_input(g, axes, taps, [border] * 3, centers)
    np.testing.assert_allclose(got.ravel(), A.T @ g.ravel(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
def test_adjoint_dot_product_identity(border):
    # <A x, w> == <x, A^T w> (projection_tests.cpp:149-210 pattern)
    rng = np.random.default_rng(10)
    shape = (9, 11, 8)
    taps = [rng.standard_normal(9) for _ in range(3)]
    x = rng.standard_normal(shape)
    w = rng.standard_normal(shape)
    ax = oracle.sepconv(x, [0, 1, 2], taps, [border] * 3)
    atw = oracle.adjoint_input(w, [0, 1, 2], taps, [border] * 3)
    lhs, rhs = np.vdot(ax, w), np.vdot(x, atw)
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
def test_adjoint_taps_central_differences(border):
    rng = np.random.default_rng(11)
    shape = (6, 7, 8)
    taps = [rng.standard_normal(k) for k in (3, 5, 3)]
    x = rng.standard_normal(shape)
    w = rng.standard_normal(shape)
    grad = oracle.adjoint_taps(x, w, [0, 1, 2], taps, [border] * 3)
    h = 1e-3
    i = 0
    for a in range(3):
        for t in range(len(taps[a])):
            tp = [v.copy() for v in taps]
            tm = [v.copy() for v in taps]
            tp[a][t] += h
            tm[a][t] -= h
            lp = np.vdot(w, oracle.sepconv(x, [0, 1, 2], tp, [border] * 3))
            lm = np.vdot(w, oracle.sepconv(x, [0, 1, 2], tm, [border] * 3))
            fd = (lp - lm) / (2 * h)  # exact up to rounding: L is linear in each tap
            assert abs(fd - grad[i]) <= 1e-8 * max(1.0, abs(fd)), (a, t, fd, grad[i])
            i += 1


def test_storage_codecs():
    for v in [0.0, -0.0, 1.0, 65504.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 1 / 3, -2.5e-6]:
        h = oracle.lib.oracle_double_to_half(v)
        assert h == np.array(v).astype(np.float16).view(np.uint16)
        assert oracle.lib.oracle_half_to_double(h) == float(np.float16(v))
    for v in [1.0, 1 / 3, 3.14159, -2.5e-6, 1e30]:
        b = oracle.lib.oracle_double_to_bf16(v)
        f = np.array(v, dtype=np.float32).view(np.uint32)
        rne = (int(f) + 0x7FFF + ((int(f) >> 16) & 1)) >> 16
        assert b == rne
