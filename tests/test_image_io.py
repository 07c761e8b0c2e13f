"""Raw and NRRD files (include/ndl/image_io.h) on the CPU.

tests/cpp/test_image_io.cpp checks round trips, type conversion, header
variants and error cases on its own; here its files are compared with the
reference's own NRRD writer (imageIO/NRRD/nrrd.h:114-164, compiled into
oracle/_ref by oracle/Makefile) byte for byte, and every file it writes or
hand-crafts is read back with the reference's loader (nrrd.h:233-366), which
must see the same values (or fail where the header layer throws).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_image_io.cpp")
REF = os.path.join(ROOT, "oracle", "_ref", "libndl_ref.so")

# (tag, numpy dtype, extent dimension 0 first, reference writer) -- the
# patterns of test_image_io.cpp, value(i) over the dense index i
CASES = [
    ("f32", np.float32, (5, 4, 3), "ref_nrrd_save_f32"),
    ("u8", np.uint8, (7, 6), "ref_nrrd_save_u8"),
    ("i16", np.int16, (9, 2, 2), "ref_nrrd_save_i16"),
    ("u16", np.uint16, (3, 3), "ref_nrrd_save_u16"),
    ("i32", np.int32, (4, 5), "ref_nrrd_save_i32"),
    ("u32", np.uint32, (6,), "ref_nrrd_save_u32"),
    ("f64", np.float64, (2, 3, 4, 2), "ref_nrrd_save_f64"),
]


def pattern(tag, n):
    i = np.arange(n, dtype=np.int64)
    if tag == "f32":
        return (np.float32(0.25) * i.astype(np.float32) - np.float32(3.0)).astype(np.float32)
    if tag == "u8":
        return ((i * 7) % 256).astype(np.uint8)
    if tag == "i16":
        return (i * 37 - 300).astype(np.int16)
    if tag == "u16":
        return ((i * 1000) % 65536).astype(np.uint16)
    if tag == "i32":
        return (i * 100003 - 7).astype(np.int32)
    if tag == "u32":
        return ((i.astype(np.uint64) * np.uint64(123456789)) % np.uint64(1 << 32)).astype(np.uint32)
    return i.astype(np.float64) * 1e-3 + 1.0 / 3.0


@pytest.fixture(scope="module")
def io_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("image_io")
    exe = str(d / "test_image_io")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC,
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, str(d)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert "Fail" not in r.stdout, r.stdout
    assert r.returncode == 0, r.stdout + r.stderr
    return d


def test_image_io_program(io_dir):
    assert (io_dir / "ours_f32.nrrd").exists()


def _ref():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return C.CDLL(REF)


def ref_load(lib, path, cap=4096):
    out = np.zeros(cap, np.float64)
    sizes = (C.c_int * 10)()
    nd = C.c_int(0)
    rc = lib.ref_nrrd_load_f64(str(path).encode(), out.ctypes.data_as(C.POINTER(C.c_double)), C.c_long(cap),
                               sizes, C.byref(nd))
    if rc:
        return None, None
    shape = tuple(sizes[i] for i in range(nd.value))
    return out[:int(np.prod(shape))], shape


@pytest.mark.parametrize("tag,dt,ext,fn", CASES)
def test_nrrd_bytes_equal_reference_writer(io_dir, tag, dt, ext, fn):
    lib = _ref()
    data = pattern(tag, int(np.prod(ext)))
    assert data.dtype == dt
    size = (C.c_int * len(ext))(*ext)
    path = io_dir / f"ref_{tag}.nrrd"
    assert getattr(lib, fn)(str(path).encode(), data.ctypes.data_as(C.c_void_p), len(ext), size) == 0
    ours = (io_dir / f"ours_{tag}.nrrd").read_bytes()
    assert ours == path.read_bytes()


@pytest.mark.parametrize("tag,dt,ext,fn", CASES)
def test_reference_loader_reads_ours(io_dir, tag, dt, ext, fn):
    lib = _ref()
    vals, shape = ref_load(lib, io_dir / f"ours_{tag}.nrrd")
    assert shape == ext
    np.testing.assert_array_equal(vals, pattern(tag, int(np.prod(ext))).astype(np.float64))


def test_reference_loader_on_header_variants(io_dir):
    """The hand-crafted files load with the same values in the reference."""
    lib = _ref()
    vals, shape = ref_load(lib, io_dir / "v_big.nrrd")
    assert shape == (3, 2)
    np.testing.assert_array_equal(vals, 0.5 * np.arange(6) + 1.0)
    vals, shape = ref_load(lib, io_dir / "v_comments.nrrd")
    assert shape == (4,) and list(vals) == [10, 20, 30, 40]
    vals, shape = ref_load(lib, io_dir / "v_skip.nrrd")
    assert shape == (2, 2) and list(vals) == [0, -3, -6, -9]
    vals, shape = ref_load(lib, io_dir / "v_datafile.nrrd")
    assert shape == (3,) and list(vals) == [1000, 1001, 1002]
    # mirrored view: z reversed, written densely
    vals, shape = ref_load(lib, io_dir / "ours_f32_mirror.nrrd")
    base = pattern("f32", 60).reshape(3, 4, 5)
    np.testing.assert_array_equal(vals.reshape(3, 4, 5), base[::-1].astype(np.float64))


@pytest.mark.parametrize("name", ["v_gzip", "v_lineskip", "v_short", "v_type"])
def test_reference_loader_rejects_what_we_reject(io_dir, name):
    lib = _ref()
    vals, _ = ref_load(lib, io_dir / f"{name}.nrrd")
    assert vals is None
