"""The C ABI library loads and exports every entry point include/*.h
declares (CPU: no compute calls)."""
import ctypes
import os
import re

from fd23_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(fd23_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "fd23_ctx_create" in names and "fd23_run" in names and len(names) > 40


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    names = set(declared_functions())
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_invalid_descriptor_is_rejected_without_gpu_work():
    lib = _lib.load()
    d = _lib.Desc(7, 0, 100, 100, 10, 0, 1, 0, 0, None)
    h = ctypes.c_void_p()
    assert lib.fd23_ctx_create(ctypes.byref(d), ctypes.byref(h)) == _lib.EINVAL
    assert b"wave" in lib.fd23_last_error()
    d = _lib.Desc(0, 0, 100, 100, 60, 0, 1, 0, 0, None)
    assert lib.fd23_ctx_create(ctypes.byref(d), ctypes.byref(h)) == _lib.EINVAL


def test_device_bytes_estimate():
    from fd23_amd.core import desc_device_bytes

    one = desc_device_bytes(_lib.WAVE_PSV, 1024, 512, 32)
    many = desc_device_bytes(_lib.WAVE_PSV, 1024, 512, 32, nshots=64)
    half = desc_device_bytes(_lib.WAVE_PSV, 1024, 512, 32, nshots=64, storage=_lib.STORE_F16)
    assert 5 * 1024 * 512 * 4 < one < 20 * 1024 * 512 * 4
    # the material is shared by the shots of a context
    assert 30 * one < many < 64 * one and half < 0.6 * many


def test_reduction_keys_alignment_check(tmp_path):
    """The material kernels' 64-bit atomic keys must sit on an 8-byte
    boundary (a misaligned 64-bit atomic faults the device).  The host check
    accepts the layouts in use (keys at the start of an allocation) and
    rejects the one that faulted: keys placed after three FP32 arrays of an
    odd cell count (333 x 17 cells -> byte offset 67932)."""
    import shutil
    import subprocess

    if not shutil.which("g++"):
        import pytest

        pytest.skip("g++ not available")
    src = tmp_path / "keys_probe.cpp"
    src.write_text(
        '#include "keys.hpp"\n#include <cstdio>\n#include <cstdlib>\n'
        "int main() {\n"
        "  alignas(256) static unsigned char buf[1 << 18];\n"
        "  const long offs[] = {0, 8, 256, 4, 67932, 1, 3 * 333 * 17 * 8};\n"
        "  for (long o : offs) std::printf(\"%ld %d\\n\", o, (int)fd23_keys_aligned(buf + o));\n"
        "  std::printf(\"null %d\\n\", (int)fd23_keys_aligned(nullptr));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "keys_probe"
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "fd23_amd", "csrc"), str(src), "-o",
                           str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = dict(line.split() for line in out if line)
    assert got == {"0": "1", "8": "1", "256": "1", "4": "0", "67932": "0", "1": "0", str(3 * 333 * 17 * 8): "1",
                   "null": "0"}


def test_tile_table_kinds_and_block_shapes():
    """Every tile index reports a kind; block shapes report K steps per
    launch (seg_rows), rows per thread (rz) and region rows (block_rows, a
    multiple of rz), with a positive tile left inside the 64-column region
    after the 7*K-cell halos."""
    lib = _lib.load()
    n = lib.fd23_num_tiles()
    kinds = [lib.fd23_tile_kind(t) for t in range(n)]
    assert set(kinds) == {_lib.TILE_TILED, _lib.TILE_FUSED, _lib.TILE_BLOCK}
    assert lib.fd23_tile_kind(n) < 0 and lib.fd23_tile_kind(-1) < 0
    for t in range(n):
        if kinds[t] != _lib.TILE_BLOCK:
            continue
        v = [ctypes.c_int() for _ in range(4)]
        assert lib.fd23_tile_shape(t, *(ctypes.byref(x) for x in v)) == 0
        rz, spb, k, rows = (x.value for x in v)
        assert spb == 1 and k in (1, 2) and rows % rz == 0
        assert 64 - 14 * k > 0 and rows - 14 * k > 0


def test_calibration_kernels_reject_bad_arguments_without_gpu_work():
    """Argument checks of the copy and tiled-stream calibration entry
    points run before any launch."""
    lib = _lib.load()
    buf = (ctypes.c_float * 16)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert lib.fd23_device_copy(p, p, 20, 1, None) == _lib.EINVAL          # not a multiple of 16
    assert lib.fd23_device_copy(None, p, 16, 1, None) == _lib.EINVAL
    arr = (ctypes.c_void_p * 9)(*([p.value] * 9))
    assert lib.fd23_device_stream_tiles(arr, 0, arr, 1, 64, 16, None) == _lib.EINVAL   # no array read
    assert lib.fd23_device_stream_tiles(arr, 9, arr, 1, 64, 16, None) == _lib.EINVAL   # more than 8
    assert lib.fd23_device_stream_tiles(arr, 1, arr, 4, 64, 16, None) == _lib.EINVAL   # more than 3 written
    assert lib.fd23_device_stream_tiles(arr, 1, arr, 1, 96, 16, None) == _lib.EINVAL   # nx not a multiple of 64
    assert lib.fd23_device_stream_tiles(arr, 1, arr, 1, 64, 24, None) == _lib.EINVAL   # nz not a multiple of 16
    nul = (ctypes.c_void_p * 2)(p.value, None)
    assert lib.fd23_device_stream_tiles(nul, 2, arr, 1, 64, 16, None) == _lib.EINVAL
    assert b"null" in lib.fd23_last_error()
