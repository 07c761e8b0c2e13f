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


def test_non_strided_image_type_is_rejected_at_compile_time(tmp_path):
    """CPU: a type with the minimal interface but no linear strides (as the
    reference's PackedBitImage) stops at a static_assert that says why."""
    src = tmp_path / "packed.cpp"
    src.write_text('''
#include <array>
#include <vector>
#include <ndl/convolution.h>
struct Packed {  // at() returns a proxy value, no stride()
    typedef float value_type;
    std::array<int, 2> extent() const { return {4, 4}; }
    float at(const std::array<int, 2>&) const { return 0.f; }
    std::vector<std::array<int, 2>> coordinates() const { return {}; }
};
int main() {
    Packed p;
    ndl::OwnedImage<float, 2> out({4, 4});
    ndl::gaussian_blur(p, out, 1.0);
}
''')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode != 0
    assert "linear strides" in r.stderr, r.stderr[-2000:]
