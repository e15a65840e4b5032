"""The C-ABI library loads and exports every entry point include/r2d.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "r2d.h")).read()
    return sorted(set(re.findall(r"^int\s+(r2d_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "r2d_process" in names and "r2d_process_batch" in names
    assert len(names) >= 30


def test_library_exports_all_declared_symbols():
    from resolve2d_amd import _native
    lib = _native.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers the header
    assert set(declared()) <= set(_native.SIGNATURES)


def test_no_device_without_gpu_is_an_error_not_a_fallback():
    """Creating a context needs a device; without one the call fails loudly."""
    import torch
    from resolve2d_amd import _native
    if torch.cuda.is_available():
        return
    lib = _native.load()
    h = ctypes.c_void_p()
    rc = lib.r2d_ctx_create(0, ctypes.byref(h))
    assert rc != 0
