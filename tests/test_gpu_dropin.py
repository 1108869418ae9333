"""The drop-in boundary exercised the way the reference's callers use it.

test_dropin_program: tests/test_dropin.cc through the C++ shims
(host/dext_gpu.hh, host/dials_spots_gpu.hh) against the oracle.

test_reference_device_test: the reference's own tests/test_dext_gpu.cc,
unchanged, linked with its src/dext.cc and src/dext_gpu.cc and with
backend/dext_hip_backend.cc in place of src/dext_cuda.cu (oracle/Makefile,
target dropin). It diffs gpu::find against the reference's CPU dext on its
own frames under all four window combinations and exits 0 only on exact
agreement."""

from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PROGRAM = os.path.join(HERE, "test_dropin.bin.exe")
REFERENCE_TEST = os.path.join(os.path.dirname(HERE), "oracle", "_ref",
                              "test_dext_gpu")


def _run(path):
    result = subprocess.run([path], capture_output=True, text=True,
                            timeout=120)
    print(result.stdout)
    print(result.stderr)
    return result


def test_dropin_program(sf):
    if not os.path.exists(PROGRAM):
        pytest.fail("tests/test_dropin.bin.exe is not built")
    result = _run(PROGRAM)
    assert result.returncode == 0, result.stdout + result.stderr
    assert "PASS" in result.stdout


def test_reference_device_test(sf):
    if not os.path.exists(REFERENCE_TEST):
        pytest.skip("the reference's device test was not built (its sources "
                    "were absent where the tree was built)")
    result = _run(REFERENCE_TEST)
    assert result.returncode == 0, result.stdout + result.stderr
