"""Morphology oracle (oracle_morph, processing/morphology.h:57-101) pinned by
the reference's own test values, the compiled reference, scipy.ndimage and
the separability identity.  CPU only."""
import itertools

import numpy as np
import pytest

import oracle

GRID = np.arange(1, 26, dtype=np.int32).reshape(5, 5)
BOX3 = np.ones((3, 3))
CROSS3 = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=np.float64)


def test_reference_unit_values():
    # image_morphology_tests.cpp:33-44: 5x5 grid 1..25, centre (2,2)
    assert oracle.morph(GRID, BOX3, "erode")[2, 2] == 7
    assert oracle.morph(GRID, BOX3, "dilate")[2, 2] == 19
    assert oracle.morph(GRID, CROSS3, "erode")[2, 2] == 8
    assert oracle.morph(GRID, CROSS3, "dilate")[2, 2] == 18


def test_dilating_a_point_gives_the_element():
    # image_morphology_tests.cpp:56-71: box -> 9 ones, cross -> 5 ones
    dot = np.zeros((5, 5), dtype=np.int32)
    dot[2, 2] = 1
    assert int(oracle.morph(dot, BOX3, "dilate").sum()) == 9
    out = oracle.morph(dot, CROSS3, "dilate")
    assert int(out.sum()) == 5 and np.array_equal(out[1:4, 1:4], CROSS3.astype(np.int32))


def test_open_close_by_hand():
    # image_morphology_tests.cpp:73-98: open = dilate(erode), close = erode(dilate)
    opened = oracle.morph(oracle.morph(GRID, BOX3, "erode"), BOX3, "dilate")
    closed = oracle.morph(oracle.morph(GRID, BOX3, "dilate"), BOX3, "erode")
    # opening is anti-extensive and idempotent, closing extensive and idempotent
    assert np.all(opened <= GRID) and np.all(closed >= GRID)
    assert np.array_equal(oracle.morph(oracle.morph(opened, BOX3, "erode"), BOX3, "dilate"), opened)
    assert np.array_equal(oracle.morph(oracle.morph(closed, BOX3, "dilate"), BOX3, "erode"), closed)


@pytest.mark.parametrize("mode,border", [("nearest", "clamp"), ("wrap", "wrap"), ("reflect", "reflect")])
@pytest.mark.parametrize("size", [(3, 5), (7, 1), (1, 9), (5, 5)])
def test_box_matches_scipy(mode, border, size):
    from scipy import ndimage
    rng = np.random.default_rng(31)
    x = rng.integers(-1000, 1000, size=(23, 29)).astype(np.int32)
    k = np.ones(size)
    assert np.array_equal(oracle.morph(x, k, "erode", border), ndimage.minimum_filter(x, size=size, mode=mode))
    assert np.array_equal(oracle.morph(x, k, "dilate", border), ndimage.maximum_filter(x, size=size, mode=mode))


def test_box_is_separable():
    # min over a product set = nested minima, pass by pass
    rng = np.random.default_rng(32)
    x = rng.standard_normal((9, 11, 13))
    full = oracle.morph(x, np.ones((3, 5, 7)), "erode", "reflect")
    seq = x
    for shape in ((1, 1, 7), (1, 5, 1), (3, 1, 1)):
        seq = oracle.morph(seq, np.ones(shape), "erode", "reflect")
    assert np.array_equal(full, seq)


def test_empty_mask_writes_zero():
    # no included tap: best stays SrcT{} (morphology.h:66-73)
    assert np.array_equal(oracle.morph(GRID, np.zeros((3, 3)), "erode"), np.zeros_like(GRID))


def _with_specials(x, rng):
    x = x.copy()
    flat = x.reshape(-1)
    idx = rng.choice(flat.size, size=max(1, flat.size // 20), replace=False)
    flat[idx[: len(idx) // 3]] = np.nan
    flat[idx[len(idx) // 3: 2 * len(idx) // 3]] = 0.0
    flat[idx[2 * len(idx) // 3:]] = -0.0
    return x


@pytest.mark.skipif(oracle.ref_lib is None or not hasattr(oracle.ref_lib, "ref_morph"),
                    reason="reference sources absent (oracle/_ref not built)")
@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.int32, np.float32, np.float64])
def test_oracle_equals_compiled_reference(dtype):
    rng = np.random.default_rng(33)
    masks = [np.ones((3, 5)), CROSS3, (rng.random((5, 3)) < 0.6).astype(float), np.ones((1, 7)),
             np.ones((3, 3, 3)), np.zeros((3, 3))]
    for k, border, op in itertools.product(masks, ["clamp", "wrap", "reflect"], ["erode", "dilate"]):
        shape = (7, 9, 6) if k.ndim == 3 else (11, 13)
        if np.issubdtype(dtype, np.floating):
            x = _with_specials(rng.standard_normal(shape).astype(dtype), rng)
        else:
            info = np.iinfo(dtype)
            x = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
        a = oracle.morph(x, k, op, border)
        b = oracle.morph(x, k, op, border, use_ref=True)
        # bitwise: NaN payloads and signed zeros included
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (dtype, k.shape, border, op)
