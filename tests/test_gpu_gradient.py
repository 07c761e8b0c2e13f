"""GPU parity of gradient() (processing/convolution.h:359-388;
ndl_gradient): every component in one launch, bit-identical to the oracle's
convolve() with {-0.5, 0, 0.5} along each axis (zero tap skipped, FP64,
truncating cast) and to the per-axis calls into component slices.  Element
types u8/i16/f32/f64 in, f32/f64/i32 out, 1-D to 4-D, every border, an
extent-1 axis, and a strided (mirrored, stepped) source.
"""
import sys

import numpy as np
import pytest

import ndl_amd
import oracle
from ndl_amd import abi

pytestmark = pytest.mark.gpu
G = np.array([-0.5, 0.0, 0.5])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def want_of(x, border, out_dtype):
    nd = x.ndim
    out = np.zeros(x.shape + (nd,), out_dtype)
    for a in range(nd):
        out[..., a] = oracle.sepconv(x, [a], [G], [border], abi.ROUND_TRUNC, dst=np.zeros(x.shape, out_dtype))
    return out


CASES = [
    ((100,), np.float64, np.float64),
    ((30, 40), np.uint8, np.float32),
    ((20, 30, 40), np.float32, np.float32),
    ((7, 9, 11, 5), np.int16, np.int32),
    ((1, 50, 60), np.float32, np.float64),
]


@pytest.mark.parametrize("border", ["reflect", "clamp", "wrap"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_gradient_vs_oracle(case, border):
    shape, dt, odt = CASES[case]
    rng = np.random.default_rng(120 + case)
    x = (rng.integers(0, 256, size=shape) if np.issubdtype(dt, np.integer) else rng.standard_normal(shape) * 10)
    x = x.astype(dt)
    say("gradient", shape, np.dtype(dt).name, np.dtype(odt).name, border)
    s = ndl_amd.DeviceArray.from_numpy(x)
    d = ndl_amd.DeviceArray(x.shape + (x.ndim,), odt)
    ndl_amd.gradient_dev(s, d, border)
    got = d.to_numpy()
    want = want_of(x, border, odt)
    np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))
    # the per-axis calls into component slices give the same bits
    d2 = ndl_amd.DeviceArray(x.shape + (x.ndim,), odt)
    for a in range(x.ndim):
        comp = d2.view(tuple([slice(None)] * x.ndim + [a]))
        ndl_amd.sepconv_dev(s, comp, [a], [G], [border], abi.ACCUM_REF, abi.ROUND_TRUNC)
    np.testing.assert_array_equal(d2.to_numpy().view(np.uint8), got.view(np.uint8))


def test_gradient_strided_source():
    rng = np.random.default_rng(130)
    full = rng.standard_normal((30, 40, 50)).astype(np.float32)
    view = (slice(None, None, -1), slice(2, 38, 3), slice(None))
    dense = np.ascontiguousarray(full[view])
    s = ndl_amd.DeviceArray.from_numpy(full).view(view)
    d = ndl_amd.DeviceArray(dense.shape + (3,), np.float32)
    ndl_amd.gradient_dev(s, d, "reflect")
    np.testing.assert_array_equal(d.to_numpy(), want_of(dense, "reflect", np.float32))


def test_gradient_argument_errors():
    s = ndl_amd.DeviceArray((10, 12), np.float32)
    with pytest.raises(ndl_amd.NdlError):
        ndl_amd.gradient_dev(s, ndl_amd.DeviceArray((10, 12, 3), np.float32))  # 3 components for 2-D
    with pytest.raises(ndl_amd.NdlError):
        ndl_amd.gradient_dev(s, ndl_amd.DeviceArray((10, 13, 2), np.float32))  # extent mismatch
