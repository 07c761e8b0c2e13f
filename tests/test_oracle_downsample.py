"""downsample oracle (oracle_downsample, processing/convolution.h:241-305)
pinned by the reference's own test (image_convolution_tests.cpp:70-91), the
compiled reference, and block means computed independently.  CPU only."""
import numpy as np
import pytest

import oracle


def test_reference_constant_colour():
    # OwnedImage<uint8_t,3> {3,40,30} = 77, downsample(., 2): {3,20,15}, all 77
    x = np.full((30, 40, 3), 77, dtype=np.uint8)  # numpy order of ndl {3,40,30}
    out = oracle.downsample(x, 2)
    assert out.shape == (15, 20, 3) and np.all(out == 77)
    assert np.array_equal(oracle.downsample(x, 2, 0), out)


@pytest.mark.parametrize("factor", [2, 3, 4])
def test_block_means(factor):
    # float64 blocks that divide the extents: the mean of each block, summed in
    # block order (dimension 0 fastest) -- numpy's pairwise sum may differ in
    # the last bits, so compare with a relative tolerance; integers exactly
    rng = np.random.default_rng(51)
    x = rng.standard_normal((4, 6 * factor, 5 * factor))  # ndl {5f, 6f, 4}, channel axis 2
    want = x.reshape(4, 6, factor, 5, factor).mean(axis=(2, 4))
    got = oracle.downsample(x, factor, channel_axis=2)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
    xi = rng.integers(0, 1000, size=(4, 6 * factor, 5 * factor)).astype(np.int32)
    s = xi.reshape(4, 6, factor, 5, factor).sum(axis=(2, 4)).astype(np.float64) / factor ** 2
    want_i = np.where(s >= 0, np.floor(s + 0.5), np.ceil(s - 0.5)).astype(np.int32)  # std::round
    assert np.array_equal(oracle.downsample(xi, factor, channel_axis=2), want_i)


def test_clipped_edge_blocks():
    # extent 7 with factor 3: blocks {0,1,2}, {3,4,5}, {6} -> mean of 1 element
    x = np.arange(7 * 2, dtype=np.float64).reshape(7, 2)  # ndl {2, 7}; channel 0 keeps 2
    out = oracle.downsample(x, 3, channel_axis=0)
    assert out.shape == (3, 2)
    assert np.array_equal(out[:, 0], [2.0, 8.0, 12.0]) and np.array_equal(out[:, 1], [3.0, 9.0, 13.0])


@pytest.mark.skipif(oracle.ref_lib is None or not hasattr(oracle.ref_lib, "ref_downsample"),
                    reason="reference sources absent (oracle/_ref not built)")
@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.int32, np.float32, np.float64])
def test_oracle_equals_compiled_reference(dtype):
    rng = np.random.default_rng(52)
    for shape, f, ca in [((13, 17), 2, 0), ((13, 17), 3, 1), ((5, 9, 11), 2, 0), ((3, 20, 21), 4, 2),
                         ((6, 7, 8, 9), 3, 1), ((2, 31), 5, 1)]:
        if np.issubdtype(dtype, np.floating):
            x = (rng.standard_normal(shape) * 100).astype(dtype)
        else:
            info = np.iinfo(dtype)
            x = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
        a = oracle.downsample(x, f, ca)
        b = oracle.downsample(x, f, ca, use_ref=True)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (shape, f, ca)
