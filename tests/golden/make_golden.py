"""Writes tests/golden/*.npz: input/output vectors produced by the REFERENCE's
own templates (oracle/_ref/libndl_ref.so, compiled from the reference sources
by oracle/Makefile in the build container).

Each fixture holds the input, the parameters and the reference output, so the
oracle restatement (tests/test_oracle.py) and the GPU path
(tests/test_gpu_parity.py) can be checked against the reference on machines
where the reference is absent.  Regenerate with

    make -C oracle && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ndl_amd import abi  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def save(name, x, out, kind, **meta):
    arrays = {"x": x, "out": out}
    for i, t in enumerate(meta.pop("taps", [])):
        arrays[f"taps_{i}"] = np.asarray(t, dtype=np.float64)
    if "kernel" in meta:
        arrays["kernel"] = np.asarray(meta.pop("kernel"), dtype=np.float64)
    meta["kind"] = kind
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)


def main():
    if oracle.ref_lib is None:
        raise SystemExit("oracle/_ref/libndl_ref.so is not built (needs the reference sources)")
    rng = np.random.default_rng(20261203)

    # gaussian_blur on uint8 (std::round path, convolution.h:180-185), every border
    for b in ["clamp", "wrap", "reflect"]:
        x = rng.integers(0, 256, size=(29, 37), dtype=np.uint8)
        out = oracle.gaussian_blur(x, 1.7, b, use_ref=True)
        save(f"u8_gauss2d_{b}", x, out, "gaussian", sigma=1.7, border=b)
    # 1-D gaussian_blur: straight convolve(), truncating (convolution.h:141-146)
    x = rng.integers(0, 256, size=(50,), dtype=np.uint8)
    save("u8_gauss1d_reflect", x, oracle.gaussian_blur(x, 1.2, "reflect", use_ref=True), "gaussian",
         sigma=1.2, border="reflect")
    # 3-D float32 Gaussian with per-axis sigma (config-2 taps), reflect
    x = rng.random((17, 19, 21), dtype=np.float64).astype(np.float32)
    taps = [oracle.gaussian_taps(1.5), oracle.gaussian_taps(1.5), oracle.gaussian_taps(3.0)]
    out = oracle.sepconv(x, [0, 1, 2], taps, ["reflect"] * 3, abi.ROUND_TRUNC, use_ref=True)
    save("f32_chain3d_reflect", x, out, "sepconv", axes=[0, 1, 2], taps=taps, borders=["reflect"] * 3,
         rounding=abi.ROUND_TRUNC)
    # uint8 box 7 on x and y with wrap (config-2 integer case): planar and interleaved
    box = np.ones(7)
    x = rng.integers(0, 256, size=(3, 24, 32), dtype=np.uint8)
    out = oracle.sepconv(x, [0, 1], [box, box], ["wrap", "wrap"], use_ref=True)
    save("u8_box7_wrap_planar", x, out, "sepconv", axes=[0, 1], taps=[box, box], borders=["wrap", "wrap"],
         rounding=abi.ROUND_TRUNC)
    x = rng.integers(0, 256, size=(24, 32, 3), dtype=np.uint8)
    out = oracle.sepconv(x, [1, 2], [box, box], ["wrap", "wrap"], use_ref=True)
    save("u8_box7_wrap_interleaved", x, out, "sepconv", axes=[1, 2], taps=[box, box],
         borders=["wrap", "wrap"], rounding=abi.ROUND_TRUNC)
    # int32 central difference (gradient()'s kernel, convolution.h:373-387), reflect
    x = rng.integers(-1000, 1000, size=(12, 16), dtype=np.int32)
    g = np.array([-0.5, 0.0, 0.5])
    out = oracle.sepconv(x, [0], [g], ["reflect"], use_ref=True)
    save("i32_gradient_reflect", x, out, "sepconv", axes=[0], taps=[g], borders=["reflect"],
         rounding=abi.ROUND_TRUNC)
    # float64 random taps, wrap
    x = rng.standard_normal((20, 33))
    t5 = rng.standard_normal(5)
    t3 = rng.standard_normal(3)
    out = oracle.sepconv(x, [0, 1], [t5, t3], ["wrap", "wrap"], use_ref=True)
    save("f64_chain2d_wrap", x, out, "sepconv", axes=[0, 1], taps=[t5, t3], borders=["wrap", "wrap"],
         rounding=abi.ROUND_TRUNC)
    # int16 out-of-range narrowing (the reference's de-facto low-bits result)
    x = rng.integers(-30000, 30000, size=(9, 40), dtype=np.int16)
    t = np.array([3.0, -7.5, 11.25, 2.0])
    out = oracle.sepconv(x, [0], [t], ["clamp"], use_ref=True)
    save("i16_overflow_clamp", x, out, "sepconv", axes=[0], taps=[t], borders=["clamp"],
         rounding=abi.ROUND_TRUNC)
    # dense 2-D kernels through convolve(): uint8 7x7 box (sums overflow uint8), wrap
    x = rng.integers(0, 256, size=(20, 26), dtype=np.uint8)
    k = np.ones((7, 7))
    save("u8_box7x7_convolve_wrap", x, oracle.convolve(x, k, "wrap", use_ref=True), "convolve",
         kernel=k, border="wrap")
    x = rng.random((11, 13)).astype(np.float32)
    k = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], dtype=np.float64) / 16
    save("f32_binomial3x3_convolve_reflect", x, oracle.convolve(x, k, "reflect", use_ref=True), "convolve",
         kernel=k, border="reflect")
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    taps = []
    i = 0
    while f"taps_{i}" in z:
        taps.append(z[f"taps_{i}"])
        i += 1
    meta["taps"] = taps
    if "kernel" in z:
        meta["kernel"] = z["kernel"]
    return z["x"], z["out"], meta


def all_fixtures():
    return sorted(os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".npz"))


if __name__ == "__main__":
    main()
