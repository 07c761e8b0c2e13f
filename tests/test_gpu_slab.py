"""GPU parity of the whole-row fused kernel (kernels_slab.hip).

float32 volumes whose unit-stride rows fit one tile (x extent even, at most
256) with three convolved axes and an instance's tap counts take this
kernel: every column of 4 rows per workgroup, one long z chunk.  Each case
runs against the oracle (FP32 within 1e-6 relative of the FP64 reference on
U[0,1) data with positive taps) and against the 64 x 16 tile kernel
(NDL_SLAB=0): both sum the same products in the same order (y, x, z passes,
taps ascending, FP32 FMA), so the two are bit-identical.

Shapes cover full and partial rows, partial 4-row tiles, short and long z
chunks (NDL_SLAB_TZ), z chunks at both volume ends, every border per axis,
batch dimensions, strided and z-mirrored views, and the fallbacks (odd or
wide rows, an off-centre x kernel).
"""
import sys

import numpy as np
import pytest

import ndl_amd
import oracle
from ndl_amd import abi, workloads

pytestmark = pytest.mark.gpu

F32 = abi.ACCUM_F32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(got, want):
    want = want.astype(np.float64)
    return float(np.max(np.abs(got.astype(np.float64) - want) / np.maximum(np.abs(want), 1e-30)))


def run(x, axes, taps, borders, view=None, centers=None):
    d = ndl_amd.DeviceArray.from_numpy(x)
    src = d.view(view) if view is not None else d
    out = ndl_amd.DeviceArray(src.shape, np.float32)
    ndl_amd.sepconv_dev(src, out, axes, taps, borders, F32, centers=centers)
    return out.to_numpy()


def both_paths(monkeypatch, x, axes, taps, borders, view=None, centers=None):
    monkeypatch.setenv("NDL_SLAB", "1")
    a = run(x, axes, taps, borders, view, centers)
    monkeypatch.setenv("NDL_SLAB", "0")
    b = run(x, axes, taps, borders, view, centers)
    monkeypatch.delenv("NDL_SLAB")
    return a, b


def vol(shape, seed):
    return workloads.uniform(seed, 0, int(np.prod(shape))).astype(np.float32).reshape(shape)


G = {3: workloads.gaussian_taps(1.0 / 3.0), 7: workloads.gaussian_taps(1.0), 9: workloads.gaussian_taps(4.0 / 3.0),
     11: workloads.gaussian_taps(1.5), 13: workloads.gaussian_taps(2.0), 19: workloads.gaussian_taps(3.0)}
assert all(len(t) == k for k, t in G.items())

CASES = [
    # (numpy shape (z, y, x), taps x/y/z, borders x/y/z, slab z chunk or 0)
    ((70, 37, 256), (11, 11, 19), ("reflect", "reflect", "reflect"), 0),
    ((70, 37, 256), (11, 11, 19), ("reflect", "reflect", "reflect"), 23),
    ((45, 37, 200), (11, 11, 19), ("clamp", "wrap", "clamp"), 16),
    ((30, 22, 96), (11, 11, 19), ("wrap", "clamp", "wrap"), 7),
    ((19, 9, 12), (11, 11, 19), ("reflect", "reflect", "reflect"), 0),
    ((40, 33, 128), (7, 7, 7), ("clamp", "reflect", "wrap"), 0),
    ((40, 33, 250), (13, 13, 13), ("wrap", "wrap", "reflect"), 9),
    ((40, 33, 64), (9, 9, 9), ("reflect", "clamp", "clamp"), 0),
    ((40, 33, 256), (11, 11, 11), ("clamp", "clamp", "clamp"), 0),
    ((21, 17, 40), (3, 3, 3), ("reflect", "wrap", "reflect"), 5),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-{'/'.join(c[2])}-tz{c[3]}")
def test_slab_vs_oracle_and_tile_kernel(case, monkeypatch):
    shape, ks, borders, tz = case
    x = vol(shape, 11)
    taps = [G[k] for k in ks]
    say("slab", shape, ks, borders, tz)
    if tz:
        monkeypatch.setenv("NDL_SLAB_TZ", str(tz))
    a, b = both_paths(monkeypatch, x, [0, 1, 2], taps, list(borders))
    want = oracle.sepconv(x, [0, 1, 2], taps, list(borders))
    err = rel_err(a, want)
    say("  max rel err", err)
    assert err <= 1e-6
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_slab_batch_dimension(monkeypatch):
    # {x, y, z, b}: the fourth dimension is untouched (a batch of volumes)
    x = vol((3, 30, 20, 64), 12)
    taps = [G[11], G[11], G[19]]
    borders = ["reflect", "clamp", "wrap"]
    a, b = both_paths(monkeypatch, x, [0, 1, 2], taps, borders)
    want = oracle.sepconv(x, [0, 1, 2], taps, borders)
    assert rel_err(a, want) <= 1e-6
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("name,view", [
    ("row window", (slice(None), slice(None), slice(8, 208))),
    ("z mirrored", (slice(None, None, -1), slice(None), slice(None))),
    ("y window, z step 2", (slice(0, 60, 2), slice(3, 40), slice(0, 256))),
])
def test_slab_strided_views(name, view, monkeypatch):
    full = vol((60, 40, 256), 13)
    taps = [G[11], G[11], G[19]]
    borders = ["reflect"] * 3
    say("slab view", name)
    a, b = both_paths(monkeypatch, full, [0, 1, 2], taps, borders, view=view)
    sub = np.ascontiguousarray(full[view])
    want = oracle.sepconv(sub, [0, 1, 2], taps, borders)
    assert rel_err(a, want) <= 1e-6
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("shape,centers", [
    ((30, 20, 201), None),        # odd row: the tile kernel
    ((30, 20, 300), None),        # wider than one tile: the tile kernel
    ((30, 20, 128), [3, 5, 9]),   # off-centre x kernel: the tile kernel
])
def test_slab_fallbacks_stay_correct(shape, centers):
    x = vol(shape, 14)
    taps = [G[11], G[11], G[19]]
    borders = ["clamp"] * 3
    got = run(x, [0, 1, 2], taps, borders, centers=centers)
    want = oracle.sepconv(x, [0, 1, 2], taps, borders, centers=centers)
    assert rel_err(got, want) <= 1e-6


def test_slab_adjoint_of_self_adjoint_filter():
    # reflect with symmetric taps is its own transpose: the input adjoint
    # runs the forward kernels (here the slab) and matches the FP64 adjoint
    g = vol((50, 36, 256), 15)
    taps = [G[11], G[11], G[19]]
    borders = ["reflect"] * 3
    d = ndl_amd.DeviceArray.from_numpy(g)
    o = ndl_amd.DeviceArray(g.shape, np.float32)
    ndl_amd.adjoint_input_dev(d, o, [0, 1, 2], taps, borders, F32)
    want = oracle.adjoint_input(g.astype(np.float64), [0, 1, 2], taps, borders)
    assert rel_err(o.to_numpy(), want) <= 1e-6
