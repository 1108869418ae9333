"""The C++ drop-in headers (dext_gpu.hh, dials_spots_gpu.hh) used as the
reference's callers use them, checked by their own parity program."""

from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
PROGRAM = os.path.join(os.path.dirname(__file__), "test_dropin.bin.exe")


def test_dropin_program(sf):
    if not os.path.exists(PROGRAM):
        pytest.fail("tests/test_dropin.bin.exe is not built")
    result = subprocess.run([PROGRAM], capture_output=True, text=True,
                            timeout=120)
    print(result.stdout)
    assert result.returncode == 0, result.stdout + result.stderr
    assert "PASS" in result.stdout
