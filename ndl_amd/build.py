"""Build libndl_gpu.so for gfx950 in-tree (ndl_amd/libndl_gpu.so).

    python -m ndl_amd.build        (or __graft_entry__.build())

hipcc cross-compiles without a GPU.  Every HIP source is compiled with
-ffp-contract=off (the reference's unfused FP64 arithmetic; fast paths call
fmaf explicitly) and with packed-FP32 instructions disabled.
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libndl_gpu.so")
BENCH_LIB = os.path.join(HERE, "libndl_benchutil.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["kernels_fused.hip", "kernels_fused_edge_a.hip", "kernels_fused_edge_b.hip", "kernels_fused_edge_c.hip",
           "kernels_fused_edge.hip", "kernels_fused_grad.hip", "kernels_generic.hip", "kernels_line.hip", "kernels_bytes.hip", "kernels_slab.hip", "kernels_xyref.hip", "kernels_short.hip", "kernels_gradient.hip", "kernels_downsample.hip", "kernels_wide.hip",
           "api.hip"]
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# the line march's ring is a register array indexed by fully unrolled steps:
# raise the size up to which `#pragma unroll` is honoured, or the largest tap
# counts keep a rolled loop and put the ring in scratch memory
EXTRA = {"kernels_line.hip": ["-mllvm", "-pragma-unroll-threshold=200000"]}


def _includes(path, seen):
    """path and every file it includes with quotes, recursively"""
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    for line in open(path):
        m = re.match(r'\s*#\s*include\s+"([^"]+)"', line)
        if m:
            _includes(os.path.join(os.path.dirname(path), m.group(1)), seen)
    return seen


def _stale(src, out):
    """NDL_BUILD_INCREMENTAL=1 skips objects newer than their source and every header it includes."""
    if os.environ.get("NDL_BUILD_INCREMENTAL") != "1" or not os.path.exists(out):
        return True
    deps = _includes(os.path.join(CSRC, src), set())
    return max(os.path.getmtime(d) for d in deps) > os.path.getmtime(out)


def _compile(src):
    out = os.path.join(OBJ, src.replace(".hip", ".o"))
    if not _stale(src, out):
        return out
    cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # measurement helpers for bench.py (cache flush, copy ceiling)
    bobj = _compile("bench_util.hip")
    tmp = BENCH_LIB + ".tmp"
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, bobj],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, BENCH_LIB)
    if verbose:
        print(f"built {LIB} and {BENCH_LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
