"""sf_hip_find_static, the threshold entry point that allocates nothing,
against sf_hip_find_u16 / sf_hip_find_u32 (the entry points whose kernels it
shares) through the C ABI: the same list, byte for byte, and the same
count, on two calls in a row, the second on a smaller frame (the static
scratch keeps the first call's tables, which must not leak into the
second). Also: the count of an overflowing frame, and the refusals."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import frames as fr
from oracle import cpu

pytestmark = pytest.mark.gpu


def entry(sf):
    fn = sf.lib().sf_hip_find_static
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                   ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                   ctypes.c_void_p, ctypes.c_void_p]
    return fn


def find_static(sf, frame: np.ndarray, capacity: int | None = None):
    """(status, list, found) from sf_hip_find_static on the null stream."""
    h, w = frame.shape
    pb = frame.dtype.itemsize
    if capacity is None:
        capacity = h * w * pb // 16
    src = sf.DeviceBuffer(frame.nbytes)
    src.upload(np.ascontiguousarray(frame))
    out = sf.DeviceBuffer(16 * capacity)
    found = sf.DeviceBuffer(4)
    status = entry(sf)(pb, src.ptr, h, w, out.ptr, capacity, found.ptr, None)
    if status != 0:
        return status, None, None
    assert sf.lib().sf_hip_synchronize() == 0
    n = np.zeros(1, np.uint32)
    found.download(n)
    n = int(n[0])
    lst = np.zeros(min(n, capacity), dtype=sf.SIGNAL)
    if n and n <= capacity:
        out.download(lst)
    for b in (src, out, found):
        b.close()
    return status, lst, n


def model_frame(sf, pb, h, w, seed):
    buf = sf.DeviceBuffer(pb * h * w)
    spots = fr.diffraction_spots(seed, h, w, 1, 1500.0 * h * w / (4362 * 4148),
                                 peak=20000.0 if pb == 2 else 1.0e6)
    sf.synth_frames(buf, pb, h, w, 1, seed, 0.3 if pb == 2 else 50.0, True,
                    0.0005, spots)
    host = np.zeros((h, w), dtype=np.uint16 if pb == 2 else np.uint32)
    buf.download(host)
    buf.close()
    return host


@pytest.mark.parametrize("pb", [2, 4])
def test_static_equals_find_two_calls_second_smaller(sf, pb):
    first = model_frame(sf, pb, 4362, 4148, 61 + pb)
    second = (fr.reference_test_frames()[5][1] if pb == 2
              else fr.u32_frame())
    for frame in (first, second):
        status, want = sf.find(frame)
        assert status == 0
        status, got, n = find_static(sf, frame)
        assert status == 0
        assert n == len(want), (frame.shape, n, len(want))
        assert got.tobytes() == want.tobytes(), frame.shape
        assert len(want) > 100


def test_static_overflow_count_and_refusals(sf):
    rng = np.random.default_rng(3)
    dense = rng.integers(0, 5000, size=(96, 128)).astype(np.uint16)
    capacity = dense.size * 2 // 16
    found = cpu.dext(dense)[2]
    assert found > capacity
    status, lst, n = find_static(sf, dense)
    assert status == 0 and n == found and len(lst) == capacity
    # then a frame that fits, on the same scratch
    ok = fr.reference_test_frames()[2][1]
    status, lst, n = find_static(sf, ok)
    assert status == 0 and lst.tobytes() == cpu.dext(ok)[1].tobytes()
    fn = entry(sf)
    buf = sf.DeviceBuffer(4096)
    cases = [
        (2, buf.ptr, 8192, 8192, buf.ptr, 8192 * 8192 // 8, buf.ptr),
        (4, buf.ptr, 6000, 6000, buf.ptr, 6000 * 6000 // 4, buf.ptr),
        (2, None, 64, 64, buf.ptr, 512, buf.ptr),
        (2, buf.ptr, 64, 64, buf.ptr, 512, None),
        (2, buf.ptr, 0, 64, buf.ptr, 512, buf.ptr),
        (2, buf.ptr, 64, 64, buf.ptr, 0, buf.ptr),
        (3, buf.ptr, 64, 64, buf.ptr, 512, buf.ptr),
        (2, buf.ptr, 65536, 65536, buf.ptr, 512, buf.ptr),
    ]
    sf.lib().sf_hip_last_error.restype = ctypes.c_char_p
    for k, case in enumerate(cases):
        assert fn(*case, None) == -1, k
        if k < 2:  # beyond the static array: says so
            assert b"scratch" in sf.lib().sf_hip_last_error()
    buf.close()
