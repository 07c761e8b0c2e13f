"""The header-only C++ drop-in (include/ndl) on the GPU: runs the compiled
tests/cpp/test_dropin program, which mirrors the reference's own unit tests
(image_convolution_tests.cpp, composite_tests.cpp, type_coverage_tests.cpp,
the convolution tutorial tables) plus views, FP32 and adjoint checks."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_dropin")


@pytest.mark.gpu
def test_cpp_dropin_api():
    assert os.path.exists(EXE), "tests/cpp/test_dropin is not built (run __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert "Fail" not in r.stdout, r.stdout
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_dropin_compiles():
    """CPU: the header layer compiles cleanly against the C ABI."""
    src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
