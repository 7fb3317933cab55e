"""The C ABI library loads and exports every symbol include/deltagpu.h declares.

No compute calls here (no GPU in the CPU suite): only symbol presence, the
version and the argument-free helpers.
"""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "deltagpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "dg_subtract" in syms and "dg_subtract_device" in syms
    assert "dg_last_error" in syms


def test_library_exports_every_declared_symbol():
    from delta_amd import _core

    lib = ctypes.CDLL(_core.library_path())
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_host_helpers_without_device():
    from delta_amd import _core

    assert _core.version() >= 100
    # GaussHermiteBasis::default_radius (src/basis.cpp:9-12)
    assert _core.default_radius(1.0, 6) == 8
    assert _core.default_radius(0.1, 0) == 3


def test_label_noalloc_refuses_oversize_frames_before_any_device_call():
    """The allocation-free labelling checks the frame against its static
    scratch first: the pointers here are never touched (no GPU needed)."""
    from delta_amd import _core

    lib = ctypes.CDLL(_core.library_path())
    lib.dg_label_device_noalloc.restype = ctypes.c_int
    lib.dg_last_error.restype = ctypes.c_char_p
    n = ctypes.c_int(5)
    dummy = ctypes.c_void_p(16)
    rc = lib.dg_label_device_noalloc(16385, 16384, dummy, None, ctypes.c_double(5.0), 1, dummy,
                                     ctypes.byref(n), None)
    assert rc != 0 and b"2^28" in lib.dg_last_error()
    rc = lib.dg_label_device_noalloc(-1, 10, dummy, None, ctypes.c_double(5.0), 1, dummy,
                                     ctypes.byref(n), None)
    assert rc != 0
