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
