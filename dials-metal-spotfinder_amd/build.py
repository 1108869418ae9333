"""Build the gfx950 library and the CPU checkers, in-tree.

    python dials-metal-spotfinder_amd/build.py

Outputs (all git-ignored, all travel with the tree):
    dials-metal-spotfinder_amd/libdext_hip.so   the product: HIP kernels + C ABI
    tests/test_dropin.bin.exe                   the C++ drop-in headers' parity
                                                program (links the .so and,
                                                as its checker, the oracle)
    oracle/liboracle.so                         the CPU oracle (checker only)
    oracle/_ref/libdext_ref.so                  the reference's own CPU path,
                                                when its sources are present
    oracle/_ref/test_dext_gpu, bench_dext_gpu   the reference's own device
                                                test and benchmark linked
                                                with backend/dext_hip_backend.cc
                                                over the library, likewise

Every HIP translation unit is compiled -ffp-contract=off (no fused
multiply-add may change a threshold test) and without packed FP32.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libdext_hip.so")
DROPIN = os.path.join(ROOT, "tests", "test_dropin.bin.exe")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    result = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    noise = [
        line
        for line in (result.stdout + result.stderr).splitlines()
        if "packed-fp32-ops" not in line and line.strip()
    ]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def _compile(source: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(source) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", source, "-o", obj])
    return obj


def build_library(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    sources = sorted(
        os.path.join(CSRC, name)
        for name in os.listdir(CSRC)
        if name.endswith(".hip")
    )
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = LIBRARY + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objects,
          "-o", tmp])
    os.replace(tmp, LIBRARY)
    return LIBRARY


def build_dropin() -> str:
    """The parity program of the C++ drop-in headers (host/), a test."""
    source = os.path.join(ROOT, "tests", "test_dropin.cc")
    _run([
        "g++", "-O2", "-std=c++20", "-ffp-contract=off", "-Wall",
        "-I", os.path.join(ROOT, "include"),
        "-I", os.path.join(HERE, "host"),
        "-I", os.path.join(ROOT, "oracle"),
        source,
        os.path.join(ROOT, "oracle", "dext_oracle.c"),
        "-o", DROPIN,
        "-L", HERE, "-ldext_hip",
        "-Wl,-rpath,$ORIGIN/../dials-metal-spotfinder_amd",
        "-lpthread", "-lm",
    ])
    return DROPIN


def build_checkers() -> None:
    """The oracle always; the reference's own CPU path when it is present."""
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-B", "-s", "-C", oracle, "oracle"])
    if os.path.isdir(os.path.join(REFERENCE, "src")):
        _run(["make", "-B", "-s", "-C", oracle, "ref",
              f"REFERENCE={REFERENCE}"])


def build_reference_dropin() -> None:
    """The reference's tests/test_dext_gpu.cc and tests/bench_dext_gpu.cc,
    unchanged, with this library as their gpu:: backend (needs the library
    and the reference sources)."""
    if os.path.isdir(os.path.join(REFERENCE, "src")):
        _run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle"),
              "dropin", f"REFERENCE={REFERENCE}"])


def build_all() -> None:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_checkers()
    build_library()
    build_dropin()
    build_reference_dropin()


if __name__ == "__main__":
    build_all()
    print("built", LIBRARY)
