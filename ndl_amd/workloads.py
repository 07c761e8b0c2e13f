"""Seeded synthetic inputs for the benchmark configurations.

Counter-based: element i of stream s under seed k is a function of (k, s, i)
only (splitmix64 finalizer), so any sub-block can be regenerated without the
rest and every run sees identical bytes.

  uniform(seed, stream, n)      U[0,1) float64 from the top 24 bits (exact in f32)
  normal(seed, stream, n)       N(0,1) via Box-Muller on two uniform streams
  config_*()                    the five benchmark workloads at full or reduced size

Axis convention: numpy shape is the reversed ndl extent (ndl axis 0 = x is
the last numpy axis).
"""
import math

import numpy as np

from . import abi

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def hash64(seed, stream, n, start=0):
    with np.errstate(over="ignore"):
        key = _mix(np.uint64(seed) * _G + np.uint64(stream + 1) * np.uint64(0xD1B54A32D192ED03))
        i = np.arange(start, start + n, dtype=np.uint64)
        return _mix(key + (i + np.uint64(1)) * _G)


def uniform(seed, stream, n, start=0):
    h = hash64(seed, stream, n, start)
    return (h >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def normal(seed, stream, n, start=0):
    u1 = uniform(seed, 2 * stream, n, start)
    u2 = uniform(seed, 2 * stream + 1, n, start)
    u1 = np.maximum(u1, 1.0 / (1 << 25))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def uniform_ints(seed, stream, n, lo, hi):
    """integers uniform in [lo, hi]"""
    h = hash64(seed, stream, n)
    return (lo + (h % np.uint64(hi - lo + 1)).astype(np.int64))


def gaussian_taps(sigma):
    """The taps gaussian_blur builds (convolution.h:116-126), computed on the
    host like the C ABI's ndl_gaussian_taps."""
    num = 3.0 * float(sigma)
    t = int(num)
    radius = t if float(t) == num else t + (1 if num > 0 else 0)
    # math.exp is libm's exp, as std::exp in the reference (numpy's
    # vectorised exp can differ in the last bit)
    w = np.array([math.exp(-(float(d) * d) / (2 * sigma * sigma)) for d in range(-radius, radius + 1)])
    s = 0.0
    for v in w:  # ascending sum, as the reference
        s += v
    return w / s


# ---------------------------------------------------------------- configs
def config1(seed=1):
    """2-D f32 512x512, sigma=2 (13 taps) both axes, clamp."""
    x = uniform(seed, 0, 512 * 512).astype(np.float32).reshape(512, 512)
    t = gaussian_taps(2.0)
    return x, [0, 1], [t, t], ["clamp", "clamp"]


def config2a(n=256, seed=2):
    """3-D f32 n^3, sigma=(1.5,1.5,3.0) -> 11/11/19 taps, reflect."""
    x = uniform(seed, 0, n * n * n).astype(np.float32).reshape(n, n, n)
    taps = [gaussian_taps(1.5), gaussian_taps(1.5), gaussian_taps(3.0)]
    return x, [0, 1, 2], taps, ["reflect"] * 3


def config2b(w=2048, h=2048, interleaved=False, seed=3):
    """u8 color w x h x 3, box [1]*7 on x and y, wrap, channel untouched.
    interleaved: ndl extent {3,w,h} (x-stride 3); planar: {w,h,3}."""
    v = uniform_ints(seed, 0, w * h * 3, 0, 255).astype(np.uint8)
    box = np.ones(7)
    if interleaved:
        return v.reshape(h, w, 3), [1, 2], [box, box], ["wrap", "wrap"]
    return v.reshape(3, h, w), [0, 1], [box, box], ["wrap", "wrap"]


def config3(w=1920, h=1080, frames=32, seed=4):
    """4-D f32 {w,h,3,frames}: U[0,1) + N(0,0.05^2); x,y sigma=1 clamp; t binomial wrap."""
    n = w * h * 3 * frames
    x = (uniform(seed, 0, n) + 0.05 * normal(seed, 1, n)).astype(np.float32).reshape(frames, 3, h, w)
    g = gaussian_taps(1.0)
    binom = np.array([1, 4, 6, 4, 1], dtype=np.float64) / 16.0
    return x, [0, 1, 3], [g, g, binom], ["clamp", "clamp", "wrap"]


def config4(seed=5, count=256, n2d=192, max2d=2048, max3d=256):
    """Batched uneven instances: (shape, taps per axis, border) list; values are
    generated per instance from the seed."""
    items = []
    h = hash64(seed, 0, count * 8)
    u = (h >> np.uint64(40)).astype(np.float64) / (1 << 24)
    for i in range(count):
        r = u[i * 8:(i + 1) * 8]
        if i < n2d:
            shape = (64 + int(r[0] * (max2d - 63)), 64 + int(r[1] * (max2d - 63)))
        else:
            shape = tuple(32 + int(r[j] * (max3d - 31)) for j in range(3))
        nd = len(shape)
        sig = [0.5 + 3.5 * r[3 + j] for j in range(nd)]
        border = ["clamp", "wrap", "reflect"][int(r[6] * 3) % 3]
        taps = [gaussian_taps(s) for s in sig]
        items.append((shape, taps, border))
    return items


def config4_values(seed, i, shape):
    n = int(np.prod(shape))
    return uniform(seed, 100 + i, n).astype(np.float32).reshape(shape)


def config5(n=384, seed=6, storage="bf16"):
    """3-D n^3 N(0,1) clipped to [-4,4] in bf16/f16; 3 random 9-tap N(0,1) kernels;
    w N(0,1) f32.  Returns (x_storage, x_f64_of_storage, w, taps)."""
    v = np.clip(normal(seed, 0, n * n * n), -4, 4)
    if storage == "f16":
        xs = v.astype(np.float16)
        xd = xs.astype(np.float64)
    else:
        f = v.astype(np.float32)
        u = f.view(np.uint32).astype(np.uint64)
        u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
        xs = u.astype(np.uint16)
        xd = (xs.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    w = normal(seed, 1, n * n * n).astype(np.float32)
    taps = [normal(seed, 10 + a, 9) for a in range(3)]
    return xs.reshape(n, n, n), xd.reshape(n, n, n), w.reshape(n, n, n), taps


BORDER_CODE = abi.BORDERS
