"""The C ABI library loads and exports every entry point include/*.h
declares (no compute calls: this runs without a GPU)."""

from __future__ import annotations

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY = os.path.join(ROOT, "dials-metal-spotfinder_amd", "libdext_hip.so")


def declared_functions():
    names = set()
    for header in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(header).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for match in re.finditer(r"\b(sf_hip_\w+)\s*\(", text):
            names.add(match.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "sf_hip_find_u16" in names and "sf_hip_find_u32" in names
    assert len(names) > 30


@pytest.mark.skipif(not os.path.exists(LIBRARY), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIBRARY)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIBRARY), reason="library not built")
def test_backend_name_without_device():
    import spotfinder

    assert spotfinder.backend() == "HIP"


@pytest.mark.skipif(not os.path.exists(LIBRARY), reason="library not built")
def test_find_refuses_frames_over_int32_max_pixels():
    """gpu::find's bound (src/dext_cuda.cu:588-593): -1 for a frame of more
    than INT32_MAX pixels, checked before anything touches the device or the
    image (a 16-byte buffer stands in for a 65536 x 65536 frame), with
    *found cleared. Also -1 for a null image or a zero dimension."""
    lib = ctypes.CDLL(LIBRARY)
    tiny = (ctypes.c_uint8 * 16)()
    out = (ctypes.c_uint8 * 16)()
    found = ctypes.c_size_t(123)
    for name in ("sf_hip_find_u16", "sf_hip_find_u32"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                       ctypes.c_void_p, ctypes.c_size_t,
                       ctypes.POINTER(ctypes.c_size_t)]
        for h, w in ((65536, 65536), (2, 0x40000000), (0x7fffffff, 2),
                     (1 << 40, 1)):
            found.value = 123
            assert fn(ctypes.addressof(tiny), h, w, ctypes.addressof(out), 1,
                      ctypes.byref(found)) == -1, (name, h, w)
            assert found.value == 0
        for h, w in ((0, 16), (16, 0)):
            assert fn(ctypes.addressof(tiny), h, w, ctypes.addressof(out), 1,
                      ctypes.byref(found)) == -1
        assert fn(None, 4, 4, ctypes.addressof(out), 1,
                  ctypes.byref(found)) == -1
    lib.sf_hip_last_error.restype = ctypes.c_char_p
    assert b"INT32_MAX" in lib.sf_hip_last_error() or b"null" in \
        lib.sf_hip_last_error()
