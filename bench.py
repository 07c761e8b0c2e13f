#!/usr/bin/env python3
"""Separable Gaussian forward + adjoint on one MI355X (Gvoxels/s, % of HBM roofline).

Workload (the largest single-GPU headline configuration): a 256^3 float32
volume, U[0,1) from a counter-based seed, separable Gaussian
sigma = (1.5, 1.5, 3.0) -> 11/11/19 taps, reflect border, FP32 accumulation.
One step = forward (ndl_sepconv_forward) + input adjoint
(ndl_sepconv_adjoint_input) of a seeded N(0,1) float32 field, inputs resident
in HBM.  Each launch is timed with HIP events (created with
hipEventDisableSystemFence, so the interval is the kernel plus its dispatch,
without a system-scope release) on the stream it runs on; by
default the Infinity Cache is flushed with a 1 GiB read before every step
(cold protocol; --warm skips the flush and is reported as warm_value).

roofline: algorithmic bytes per launch = (4 B read + 4 B write) per voxel of
the fused kernel, divided by its mean launch time, against 8 TB/s.
cpu_baseline: the reference's own convolve() chain (oracle/_ref, built from
the reference sources) on one host core over a quarter slab of the volume,
forward of x and input adjoint of w each timed (the reflect Gaussian is its
own transpose), median of three samples; falls back to the C restatement when
the reference build is absent.  An all-cores run of the restatement is
reported beside it, labelled as a port.

    python bench.py [--steps K] [--warmup W] [--n 256] [--warm] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvoxels/s + % HBM roofline, separable conv fwd/adjoint, 1 MI355X vs host CPU"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--warm", action="store_true", help="no cache flush between steps")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(sharding over several GPUs is designed, not built or measured)", file=sys.stderr)
        sys.exit(2)

    import ndl_amd
    from ndl_amd import abi, workloads

    n = args.n
    nvox = n ** 3
    hip = ndl_amd.hip()
    stream = hip.stream()
    lib = ndl_amd.lib()
    bu = C.CDLL(os.path.join(ROOT, "ndl_amd", "libndl_benchutil.so"))
    bu.ndl_bench_flush.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
    bu.ndl_bench_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]

    log(f"bench: building inputs {n}^3 f32")
    x, axes, taps, borders = workloads.config2a(n=n)
    w = workloads.normal(2, 7, nvox).astype(np.float32).reshape(x.shape)
    dx = ndl_amd.DeviceArray.from_numpy(x)
    dw = ndl_amd.DeviceArray.from_numpy(w)
    dy = ndl_amd.DeviceArray(x.shape, np.float32)
    dgx = ndl_amd.DeviceArray(x.shape, np.float32)
    flush_bytes = 1 << 30
    dflush = ndl_amd.DeviceArray((flush_bytes,), np.uint8)
    dsink = ndl_amd.DeviceArray((4096,), np.uint32)
    ndl_amd.check(lib.ndl_memset(dflush.ptr, 1, flush_bytes, stream))

    ks = abi.Kernels(taps, borders)
    ax = abi.Kernels.axes(axes)
    tx, ty, tw, tg = dx.tensor(), dy.tensor(), dw.tensor(), dgx.tensor()

    def fwd():
        ndl_amd.check(lib.ndl_sepconv_forward(C.byref(tx), C.byref(ty), len(axes), ax, ks.arr,
                                              abi.ACCUM_F32, abi.ROUND_TRUNC, stream))

    def adj():
        ndl_amd.check(lib.ndl_sepconv_adjoint_input(C.byref(tw), C.byref(tg), len(axes), ax, ks.arr,
                                                    abi.ACCUM_F32, stream))

    def flush():
        if bu.ndl_bench_flush(dflush.ptr, flush_bytes, dsink.ptr, stream) != 0:
            raise RuntimeError("flush launch failed")

    def run(steps, cold):
        ev = [[hip.event(hip.EVENT_NO_SYSTEM_FENCE) for _ in range(3)] for _ in range(steps)]
        ndl_amd.synchronize(stream)
        t0 = time.perf_counter()
        for i in range(steps):
            if cold:
                flush()
            hip.record(ev[i][0], stream)
            fwd()
            hip.record(ev[i][1], stream)
            adj()
            hip.record(ev[i][2], stream)
        ndl_amd.synchronize(stream)
        wall = time.perf_counter() - t0
        f = [hip.elapsed_ms(e[0], e[1]) for e in ev]
        a = [hip.elapsed_ms(e[1], e[2]) for e in ev]
        return np.array(f), np.array(a), wall

    log(f"bench: warmup {args.warmup}")
    run(max(args.warmup, 1), cold=not args.warm)
    cold = not args.warm
    log(f"bench: timing {args.steps} steps ({'cold' if cold else 'warm'} cache)")
    f_ms, a_ms, wall = run(args.steps, cold)
    step_ms = f_ms + a_ms
    ms_per_step = float(np.mean(step_ms))
    value = nvox / (ms_per_step * 1e-3) / 1e9
    log(f"bench: fwd {np.mean(f_ms)*1e3:.1f} us, adj {np.mean(a_ms)*1e3:.1f} us, wall {wall:.3f} s")
    # the other protocol, reported beside
    f2, a2, _ = run(max(args.steps // 2, 3), not cold)
    other_value = nvox / (float(np.mean(f2 + a2)) * 1e-3) / 1e9

    # measured streaming-copy ceiling (1 GiB -> 1 GiB... half each way)
    half = flush_bytes // 2
    e0, e1 = hip.event(hip.EVENT_NO_SYSTEM_FENCE), hip.event(hip.EVENT_NO_SYSTEM_FENCE)
    bu.ndl_bench_copy(dflush.ptr, dflush.ptr + half, half, stream)
    hip.record(e0, stream)
    for _ in range(5):
        bu.ndl_bench_copy(dflush.ptr, dflush.ptr + half, half, stream)
    hip.record(e1, stream)
    copy_gbps = 5 * 2 * half / (hip.elapsed_ms(e0, e1) * 1e-3) / 1e9

    launch_ms = float(np.mean(np.concatenate([f_ms, a_ms])))
    alg_bytes = 8 * nvox
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    # HBM bytes per launch: measured by two rocprofv3 --pmc runs of this bench
    # (FETCH_SIZE, WRITE_SIZE, calibrated against kernels of known size) and
    # committed with its commands; not measurable from inside this process
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "profiles", f"traffic_gauss3d_{n}_f32.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            traffic = tj.get("hbm_bytes_per_launch")
            traffic_src = f"profiles/traffic_gauss3d_{n}_f32.json, from {tj.get('source', 'rocprofv3 --pmc')}"
        except (OSError, ValueError):
            traffic = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, w, axes, taps, borders)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded counter-based U[0,1) volume, N(0,1) adjoint seed",
        "config": {
            "workload": f"separable Gaussian forward+adjoint, f32 {n}^3 volume",
            "extent": [n, n, n],
            "sigma": [1.5, 1.5, 3.0],
            "taps": [len(t) for t in taps],
            "border": "reflect",
            "accumulation": "fp32",
            "step": "ndl_sepconv_forward + ndl_sepconv_adjoint_input",
            "cache": "cold (1 GiB read flush before each step)" if cold else "warm",
        },
        "roofline": {
            # the roofline denominator is HBM; the SQ counters of the kernel
            # (profiles/) show it bound by instruction issue and memory
            # latency, not by HBM bandwidth (warm and cold times are equal)
            "bound": "hbm",
            "measured_limiter": "instruction issue and memory latency (SQ counters), not HBM bandwidth",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "slab_kernel (fwd and adjoint launches)",
            "alg_bytes_per_launch": alg_bytes,
            "mean_launch_us": round(launch_ms * 1e3, 2),
        },
        "cpu_baseline": cpu,
        ("warm_value" if cold else "cold_value"): round(other_value, 3),
        "fwd_us": round(float(np.mean(f_ms)) * 1e3, 2),
        "adj_us": round(float(np.mean(a_ms)) * 1e3, 2),
        "copy_ceiling_GBps": round(copy_gbps, 1),
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(x, w, axes, taps, borders):
    """The reference's CPU path on a bounded sample of the workload.

    Sample: a 256 x 256 x 64 slab of the volume (one quarter; the per-voxel
    cost of the serial loop does not depend on the slab's place), the same
    taps and borders.  One sample = the forward of x plus the input adjoint
    of w; this operator (reflect, symmetric taps) is its own transpose, so the
    adjoint is timed as the reference convolve() chain applied to w.  Median
    of three samples on one pinned core (the reference loop is serial,
    convolution.h:64-70).  Also timed: the oracle's restatement on every host
    core (bit-identical outputs), labelled separately as a port."""
    import oracle
    nz = max(1, x.shape[0] // 4)
    xs, ws = np.ascontiguousarray(x[:nz]), np.ascontiguousarray(w[:nz])
    nvox = xs.size
    y = np.empty_like(xs)
    try:
        saved = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {sorted(saved)[0]})
    except (AttributeError, OSError):
        saved = None
    use_ref = oracle.ref_lib is not None
    log(f"bench: CPU baseline ({'reference build' if use_ref else 'C restatement'}), one core, {xs.shape} slab x3")
    samples = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.sepconv(xs, axes, taps, borders, dst=y, use_ref=use_ref)
        t1 = time.perf_counter()
        oracle.sepconv(ws, axes, taps, borders, dst=y, use_ref=use_ref)
        samples.append((t1 - t0, time.perf_counter() - t1))
    fwd = float(np.median([a for a, _ in samples]))
    adj = float(np.median([b for _, b in samples]))
    step = float(np.median([a + b for a, b in samples]))
    if saved is not None:
        try:
            os.sched_setaffinity(0, saved)
        except OSError:
            pass
    ncores = len(saved) if saved is not None else (os.cpu_count() or 1)
    with oracle.threads(ncores):
        t0 = time.perf_counter()
        oracle.sepconv(xs, axes, taps, borders, dst=y)
        oracle.sepconv(ws, axes, taps, borders, dst=y)
        t_all = time.perf_counter() - t0
    log(f"bench: CPU fwd {fwd:.2f} s, adj {adj:.2f} s (median of 3), all-cores port {t_all:.2f} s")
    return {
        "value": round(nvox / step / 1e9, 5),
        "unit": "Gvoxels/s",
        "cores": 1,
        "kind": "reference" if use_ref else "port",
        "sample": (f"{xs.shape[2]}x{xs.shape[1]}x{xs.shape[0]} slab of the volume, forward of x + input adjoint "
                   f"of w (the reflect Gaussian is self-adjoint: the reference convolve() chain on w), median of "
                   f"3 samples: fwd {fwd:.3f} s, adj {adj:.3f} s; host: {cpu_model()}, nproc {os.cpu_count()}"),
        "all_cores_port": {"value": round(nvox / t_all / 1e9, 5), "unit": "Gvoxels/s", "cores": ncores,
                           "kind": "port", "note": "the oracle's C restatement on every core, not the reference"},
    }


# --------------------------------------------------------------------- suite
def _time_launches(hip, stream, fn, reps, flush=None):
    """mean ms of fn() over reps, each bracketed by events (flush outside).

    `flush` runs before each rep's first event: the cache flush for cold
    timings, or a spin kernel for warm ones. Either keeps the GPU busy
    while the host enqueues the events and fn's launches, so the event pair
    times GPU work only."""
    import ndl_amd
    ts = []
    for _ in range(reps):
        if flush:
            flush()
        e0, e1 = hip.event(hip.EVENT_NO_SYSTEM_FENCE), hip.event(hip.EVENT_NO_SYSTEM_FENCE)
        hip.record(e0, stream)
        fn()
        hip.record(e1, stream)
        ts.append((e0, e1))
    ndl_amd.synchronize(stream)
    return float(np.mean([hip.elapsed_ms(a, b) for a, b in ts]))


def suite(reps=10, configs=(1, 2, 3, 4, 5, 6)):
    """Secondary configurations; one JSON line each on stdout."""
    import ndl_amd
    import oracle
    from ndl_amd import abi, workloads

    hip = ndl_amd.hip()
    stream = hip.stream()
    lib = ndl_amd.lib()
    bu = C.CDLL(os.path.join(ROOT, "ndl_amd", "libndl_benchutil.so"))
    bu.ndl_bench_flush.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
    bu.ndl_bench_spin.argtypes = [C.c_double, C.c_void_p, C.c_void_p]
    dflush = ndl_amd.DeviceArray((1 << 30,), np.uint8)
    dsink = ndl_amd.DeviceArray((4096,), np.uint32)

    def flush():
        bu.ndl_bench_flush(dflush.ptr, 1 << 30, dsink.ptr, stream)

    def spin():
        # warm timings: 200 us of one busy wave ahead of the events
        bu.ndl_bench_spin(200.0, dsink.ptr, stream)

    def fwd_fn(src, dst, axes, taps, borders, accum):
        ks = abi.Kernels(taps, borders)
        ax = abi.Kernels.axes(axes)
        ts, td = src.tensor(), dst.tensor()

        def f():
            ndl_amd.check(lib.ndl_sepconv_forward(C.byref(ts), C.byref(td), len(axes), ax, ks.arr, accum,
                                                  abi.ROUND_TRUNC, stream))
        f.keep = (ks, ax, ts, td)
        return f

    def emit(d):
        print(json.dumps(d), flush=True)

    if 1 in configs:
        x, axes, taps, borders = workloads.config1()
        s = ndl_amd.DeviceArray.from_numpy(x)
        d = ndl_amd.DeviceArray(x.shape, np.float32)
        f = fwd_fn(s, d, axes, taps, borders, abi.ACCUM_F32)
        f()
        ms = _time_launches(hip, stream, f, reps, spin)
        emit({"config": "1: 2-D f32 512x512 sigma 2 clamp, forward, warm", "us": round(ms * 1e3, 2),
              "gvox_s": round(x.size / ms / 1e6, 2), "GBps_alg": round(8 * x.size / ms / 1e6, 1)})
    if 2 in configs:
        for inter in (False, True):
            x, axes, taps, borders = workloads.config2b(interleaved=inter)
            s = ndl_amd.DeviceArray.from_numpy(x)
            d = ndl_amd.DeviceArray(x.shape, np.uint8)
            f = fwd_fn(s, d, axes, taps, borders, abi.ACCUM_REF)
            f()
            ms = _time_launches(hip, stream, f, reps, spin)
            ms_host = _time_launches(hip, stream, f, reps)
            ok = np.array_equal(d.to_numpy(), oracle.sepconv(x, axes, taps, borders))
            emit({"config": f"2b: u8 2048x2048x3 box7 wrap ({'interleaved' if inter else 'planar'}), forward, warm",
                  "us": round(ms * 1e3, 2), "gvox_s": round(x.size / ms / 1e6, 2),
                  "GBps_alg": round(2 * x.size / ms / 1e6, 1), "bit_exact": bool(ok),
                  "us_with_host_launch": round(ms_host * 1e3, 2)})
        # the 256^3 Gaussian's axes one at a time (convolve() with a 1-D
        # kernel image per axis), cold cache: x on the row kernel, y and z
        # on the line march; FP32 accumulation, then the reference's FP64
        # arithmetic (bit-identical), then gaussian_blur()'s reference mode
        # (three passes through FP64 intermediates)
        x, axes, taps, borders = workloads.config2a()
        s = ndl_amd.DeviceArray.from_numpy(x)
        d = ndl_amd.DeviceArray(x.shape, np.float32)
        for accum, mode in ((abi.ACCUM_F32, "fp32"), (abi.ACCUM_REF, "reference fp64")):
            for a, name in zip(axes, "xyz"):
                f = fwd_fn(s, d, [a], [taps[a]], [borders[a]], accum)
                f()
                ms = _time_launches(hip, stream, f, reps, flush)
                want = oracle.sepconv(x, [a], [taps[a]], [borders[a]])
                got = d.to_numpy()
                emit({"config": f"2a one axis: 256^3 f32, {len(taps[a])} taps on {name} only, reflect, {mode}, cold",
                      "us": round(ms * 1e3, 2), "gvox_s": round(x.size / ms / 1e6, 2),
                      "GBps_alg": round(8 * x.size / ms / 1e6, 1),
                      "max_abs_err": float(np.max(np.abs(got.astype(np.float64) - want.astype(np.float64)))),
                      "bit_exact": bool(np.array_equal(got, want))})
        f = fwd_fn(s, d, axes, taps, borders, abi.ACCUM_REF)
        f()
        ms = _time_launches(hip, stream, f, reps, flush)
        emit({"config": "2a reference arithmetic: 256^3 f32 Gaussian 11/11/19 reflect, FP64 chain, cold",
              "us": round(ms * 1e3, 2), "gvox_s": round(x.size / ms / 1e6, 2),
              "bit_exact": bool(np.array_equal(d.to_numpy(), oracle.sepconv(x, axes, taps, borders)))})
        # the drop-in gaussian_blur() on a uint8 colour image (every axis,
        # the channel axis included, rounded to nearest), and one pyramid
        # step (blur then downsample by 2) of the 256^3 volume, both in the
        # reference arithmetic; the pyramid step against the two calls
        rng = np.random.default_rng(7)
        xc = rng.integers(0, 256, size=(2048, 2048, 3)).astype(np.uint8)
        sc = ndl_amd.DeviceArray.from_numpy(xc)
        dc = ndl_amd.DeviceArray(xc.shape, np.uint8)
        g1 = workloads.gaussian_taps(1.0)
        ks = abi.Kernels([g1] * 3, ["clamp"] * 3)
        ax3 = abi.Kernels.axes([0, 1, 2])
        tsc, tdc = sc.tensor(), dc.tensor()

        def fc():
            ndl_amd.check(lib.ndl_sepconv_forward(C.byref(tsc), C.byref(tdc), 3, ax3, ks.arr, abi.ACCUM_REF,
                                                  abi.ROUND_NEAREST, stream))
        fc()
        ms = _time_launches(hip, stream, fc, reps, flush)
        emit({"config": "gaussian_blur uint8 colour {3,2048,2048}, sigma 1, clamp, reference arithmetic, cold",
              "us": round(ms * 1e3, 2), "gvox_s": round(xc.size / ms / 1e6, 2),
              "bit_exact": bool(np.array_equal(dc.to_numpy(), oracle.gaussian_blur(xc, 1.0, "clamp")))})
        del sc, dc
        sig = np.full(3, 1.5)
        oshape = (128, 128, 256)
        dd = ndl_amd.DeviceArray(oshape, np.float32)
        mid = ndl_amd.DeviceArray(x.shape, np.float32)
        g15 = workloads.gaussian_taps(1.5)
        kb = abi.Kernels([g15] * 3, ["reflect"] * 3)
        ts_, tdd, tmid = s.tensor(), dd.tensor(), mid.tensor()

        def fpyr():
            ndl_amd.check(lib.ndl_gaussian_downsample(C.byref(ts_), C.byref(tdd), sig.ctypes.data_as(C.POINTER(C.c_double)),
                                                      abi.BORDERS["reflect"], 2, 0, abi.ACCUM_REF, abi.ROUND_TRUNC,
                                                      stream))

        def ftwo():
            ndl_amd.check(lib.ndl_sepconv_forward(C.byref(ts_), C.byref(tmid), 3, ax3, kb.arr, abi.ACCUM_REF,
                                                  abi.ROUND_TRUNC, stream))
            ndl_amd.check(lib.ndl_downsample(C.byref(tmid), C.byref(tdd), 2, 0, stream))
        fpyr()
        ms = _time_launches(hip, stream, fpyr, reps, flush)
        got = dd.to_numpy()
        ftwo()
        ms2 = _time_launches(hip, stream, ftwo, reps, flush)
        emit({"config": "pyramid step: 256^3 f32 gaussian_blur sigma 1.5 reflect then downsample by 2 (x kept), "
                        "reference arithmetic, cold",
              "us": round(ms * 1e3, 2), "us_blur_then_downsample": round(ms2 * 1e3, 2),
              "gvox_s": round(x.size / ms / 1e6, 2),
              "bit_exact_vs_two_calls": bool(np.array_equal(got, dd.to_numpy()))})
        del dd, mid
    if 3 in configs:
        x, axes, taps, borders = workloads.config3()
        s = ndl_amd.DeviceArray.from_numpy(x)
        d = ndl_amd.DeviceArray(x.shape, np.float32)
        f = fwd_fn(s, d, axes, taps, borders, abi.ACCUM_F32)
        f()
        ms = _time_launches(hip, stream, f, max(3, reps // 2), spin)
        emit({"config": "3: 4-D f32 1920x1080x3x32, x,y sigma 1 clamp + t binomial wrap, forward",
              "us": round(ms * 1e3, 2), "gvox_s": round(x.size / ms / 1e6, 2),
              "GBps_alg": round(8 * x.size / ms / 1e6, 1)})
        del s, d
    if 4 in configs:
        items = workloads.config4()
        srcs, dsts, bitems, keep = [], [], [], []
        nvox = 0
        for i, (shape, taps, border) in enumerate(items):
            x = workloads.config4_values(5, i, shape)
            nvox += x.size
            sd = ndl_amd.DeviceArray.from_numpy(x)
            dd = ndl_amd.DeviceArray(shape, np.float32)
            nd = len(shape)
            ks = abi.Kernels(taps, [border] * nd)
            it = abi.BatchItem()
            it.src, it.dst, it.naxes = sd.tensor(), dd.tensor(), nd
            for a in range(nd):
                it.axes[a] = a
                it.k[a] = ks.arr[a]
            srcs.append(sd)
            dsts.append(dd)
            bitems.append(it)
            keep.append(ks)
        arr = (abi.BatchItem * len(bitems))(*bitems)

        def time_plan(env=None):
            """one plan (autotuned, or pinned by an environment knob), ms per execute"""
            saved = {k: os.environ.get(k) for k in (env or {})}
            os.environ.update(env or {})
            try:
                plan = C.c_void_p()
                ndl_amd.check(lib.ndl_plan_create(arr, len(bitems), abi.ACCUM_F32, C.byref(plan)))
            finally:
                for k, v in saved.items():
                    if v is None:
                        os.environ.pop(k, None)
                    else:
                        os.environ[k] = v
            run_plan = lambda: ndl_amd.check(lib.ndl_plan_execute(plan, stream))  # noqa: E731
            run_plan()  # autotunes (and captures) on first execute
            ms = _time_launches(hip, stream, run_plan, max(3, reps // 2), spin)
            ndl_amd.check(lib.ndl_plan_destroy(plan))
            return ms

        ms_queues = time_plan({"NDL_PLAN_PER_AXIS": "0"})
        ms_axes = time_plan({"NDL_PLAN_PER_AXIS_ONLY": "1"})
        ms_plan = time_plan()
        singles = [fwd_fn(srcs[i], dsts[i], list(range(len(items[i][0]))), items[i][1],
                          [items[i][2]] * len(items[i][0]), abi.ACCUM_F32) for i in range(len(items))]

        def run_singles():
            for f in singles:
                f()
        run_singles()
        ms_single = _time_launches(hip, stream, run_singles, max(3, reps // 2), spin)
        emit({"config": "4: 256 uneven f32 instances (192 2-D, 64 3-D), one plan (autotuned, hipGraph)",
              "voxels": nvox, "plan_us": round(ms_plan * 1e3, 1), "plan_gvox_s": round(nvox / ms_plan / 1e6, 2),
              "work_queues_only_us": round(ms_queues * 1e3, 1), "per_axis_only_us": round(ms_axes * 1e3, 1),
              "single_launches_us": round(ms_single * 1e3, 1),
              "speedup_vs_single_launches": round(ms_single / ms_plan, 3),
              "GBps_alg_plan": round(8 * nvox / ms_plan / 1e6, 1)})
        del srcs, dsts
    if 5 in configs:
        n = 384
        for storage in ("bf16", "f16"):
            xs, xd, w, taps = workloads.config5(n=n, storage=storage)
            for border in ("clamp", "reflect"):
                axes = [0, 1, 2]
                borders = [border] * 3
                dx = ndl_amd.DeviceArray.from_numpy(xs, bf16=(storage == "bf16"))
                dy = ndl_amd.DeviceArray(xs.shape, xs.dtype, bf16=(storage == "bf16"))
                dw = ndl_amd.DeviceArray.from_numpy(w)
                dgx = ndl_amd.DeviceArray(xs.shape, np.float32)
                d32 = ndl_amd.DeviceArray.from_numpy(xd.astype(np.float32))
                y32 = ndl_amd.DeviceArray(xs.shape, np.float32)
                f_half = fwd_fn(dx, dy, axes, taps, borders, abi.ACCUM_F32)
                f_32 = fwd_fn(d32, y32, axes, taps, borders, abi.ACCUM_F32)

                def adj():
                    ndl_amd.adjoint_input_dev(dw, dgx, axes, taps, borders, abi.ACCUM_F32, stream=stream)

                dgt = ndl_amd.DeviceArray((27,), np.float64)

                def adj_taps():
                    ndl_amd.adjoint_taps_dev(d32, dw, axes, taps, borders, dgt, abi.ACCUM_F32, stream=stream)
                f_half()
                f_32()
                adj()
                ms_h = _time_launches(hip, stream, f_half, reps, flush)
                ms_32 = _time_launches(hip, stream, f_32, reps, flush)
                ms_adj = _time_launches(hip, stream, adj, reps, flush)
                adj_taps()
                ms_taps = _time_launches(hip, stream, adj_taps, max(3, reps // 2), flush)
                # errors vs the FP64 oracle fed the storage-rounded input
                want = oracle.sepconv(xd, axes, taps, borders, dst=np.zeros(xd.shape))
                got = dy.to_numpy()
                got = (got.astype(np.uint32) << 16).view(np.float32).astype(np.float64) if storage == "bf16" \
                    else got.astype(np.float64)
                err = got - want
                gxw = oracle.adjoint_input(w.astype(np.float64), axes, taps, borders)
                egx = dgx.to_numpy().astype(np.float64) - gxw
                emit({"config": f"5: {storage} 384^3, 3x9 random taps, {border}",
                      "fwd_us": round(ms_h * 1e3, 2), "fwd_f32_form_us": round(ms_32 * 1e3, 2),
                      "speedup_vs_f32": round(ms_32 / ms_h, 3), "adj_input_f32_us": round(ms_adj * 1e3, 2),
                      "adj_taps_f32_us": round(ms_taps * 1e3, 1),
                      "fwd_max_abs_err": float(np.max(np.abs(err))), "fwd_rms_err": float(np.sqrt(np.mean(err ** 2))),
                      "fwd_rms_ref": float(np.sqrt(np.mean(want ** 2))),
                      "adj_max_abs_err": float(np.max(np.abs(egx))), "adj_rms_err": float(np.sqrt(np.mean(egx ** 2)))})
                del dx, dy, dw, dgx, d32, y32, dgt
    if 6 in configs:
        # separable morphology (box structuring elements), bit-exact vs the oracle
        cases = [("6: u8 2048x2048 box 7x7 erode, clamp", (2048, 2048), np.uint8, [7, 7], "clamp", "erode"),
                 ("6: f32 256^3 box 5x5x5 dilate, reflect", (256, 256, 256), np.float32, [5, 5, 5], "reflect",
                  "dilate")]
        for name, shape, dt, ks_, border, op in cases:
            n = int(np.prod(shape))
            if dt == np.uint8:
                x = workloads.uniform_ints(3, 9, n, 0, 255).astype(np.uint8).reshape(shape)
            else:
                x = workloads.uniform(2, 9, n).astype(np.float32).reshape(shape)
            s_ = ndl_amd.DeviceArray.from_numpy(x)
            d_ = ndl_amd.DeviceArray(shape, dt)
            axes = list(range(len(shape)))
            kk = abi.Kernels([np.ones(k) for k in ks_], [border] * len(ks_))
            ax = abi.Kernels.axes(axes)
            ts_, td_ = s_.tensor(), d_.tensor()

            def f(ts_=ts_, td_=td_, kk=kk, ax=ax, nax=len(axes), op=op):
                ndl_amd.check(lib.ndl_morph_separable(C.byref(ts_), C.byref(td_), nax, ax, kk.arr,
                                                      ndl_amd.MORPH_OPS[op], stream))
            f()
            ms = _time_launches(hip, stream, f, reps, flush)
            ok = True
            if n <= 1 << 22:
                ok = bool(np.array_equal(d_.to_numpy(), oracle.morph(x, np.ones(ks_), op, border)))
            emit({"config": name, "us": round(ms * 1e3, 2), "gvox_s": round(n / ms / 1e6, 2),
                  "GBps_alg": round(2 * x.itemsize * n / ms / 1e6, 1), "passes": len(ks_),
                  "bit_exact": ok if n <= 1 << 22 else "checked in tests at smaller sizes"})
            del s_, d_


if __name__ == "__main__":
    if "--suite" in sys.argv:
        which = [int(a) for a in sys.argv[sys.argv.index("--suite") + 1:] if a.isdigit()] or [1, 2, 3, 4, 5, 6]
        suite(configs=tuple(which))
    else:
        main()
