"""The C-ABI library loads and exports every symbol include/mfsv.h declares.

CPU-only: loading libmfsv_hip.so needs no GPU; no compute entry point is
called here (device queries return an error or zero devices without one).
"""
import os
import re

from mfcm_amd import mfsv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mfsv.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mfsv_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert "mfsv_create" in syms and "mfsv_run" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol():
    L = mfsv.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert set(declared_symbols()) == set(mfsv.SIGNATURES)


def test_version_string():
    assert b"gfx950" in mfsv.lib().mfsv_version()


def test_argument_errors_without_device():
    # null arguments are rejected before any device call
    assert mfsv.lib().mfsv_create(None, 0, 0, None) == -1
    assert mfsv.lib().mfsv_destroy(None) == -1
    assert mfsv.lib().mfsv_device_count(None) == -1


def test_grid_limits_without_device():
    # sizes and spacings are checked before any device call: each field must
    # stay below 2 GiB (32-bit byte offsets in the kernels)
    import ctypes as C
    h = C.c_void_p()
    L = mfsv.lib()
    big64 = mfsv.Grid(16384, 16384, 1.0, 1.0, 0.0, 0.0)      # 2^28 cells x 8 B = 2 GiB
    assert L.mfsv_create(C.byref(big64), mfsv.F64, 0, C.byref(h)) == -1
    big32 = mfsv.Grid(32768, 16384, 1.0, 1.0, 0.0, 0.0)      # 2^29 cells x 4 B = 2 GiB
    assert L.mfsv_create(C.byref(big32), mfsv.F32, 0, C.byref(h)) == -1
    for bad in (mfsv.Grid(0, 5, 1.0, 1.0, 0.0, 0.0), mfsv.Grid(5, 5, 0.0, 1.0, 0.0, 0.0),
                mfsv.Grid(5, 5, 1.0, float("nan"), 0.0, 0.0)):
        assert L.mfsv_create(C.byref(bad), mfsv.F64, 0, C.byref(h)) == -1
    assert L.mfsv_create(C.byref(mfsv.Grid(5, 5, 1.0, 1.0, 0.0, 0.0)), 7, 0, C.byref(h)) == -1
    assert not h.value


def test_noalloc_control_number_rejects_bad_arguments_without_device():
    # sizes, spacings, dt, precision and null pointers are checked before any
    # device call; the cell limit is what the fixed scratch array holds
    import ctypes as C
    L = mfsv.lib()
    p = C.c_void_p(8)
    f = L.mfsv_control_number_noalloc
    assert f(mfsv.F64, 0, 8193, 8192, 1.0, 1.0, p, p, p, 0.1, p, None) == -1   # > 2^26 cells
    assert f(mfsv.F32, 0, 1 << 40, 1 << 40, 1.0, 1.0, p, p, p, 0.1, p, None) == -1
    assert f(mfsv.F64, 0, 0, 5, 1.0, 1.0, p, p, p, 0.1, p, None) == -1
    assert f(mfsv.F64, 0, 5, 5, 0.0, 1.0, p, p, p, 0.1, p, None) == -1
    assert f(mfsv.F64, 0, 5, 5, 1.0, 1.0, p, p, p, 0.0, p, None) == -1
    assert f(mfsv.F64, 0, 5, 5, 1.0, 1.0, None, p, p, 0.1, p, None) == -1
    assert f(mfsv.F64, 0, 5, 5, 1.0, 1.0, p, p, p, 0.1, None, None) == -1
    assert f(7, 0, 5, 5, 1.0, 1.0, p, p, p, 0.1, p, None) == -1
    assert f(mfsv.F64, -1, 5, 5, 1.0, 1.0, p, p, p, 0.1, p, None) == -1
