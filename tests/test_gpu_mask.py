"""The detector pixel mask stamped into device frames (sf_hip_stamp_mask,
the step the reference names as its one omission, README.md:227-234): the
stamped frames equal the host's np.where(mask, max, frame), and the
threshold of the stamped frames equals the oracle's on the same frames."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu
import frames as fr

pytestmark = pytest.mark.gpu


def same(a, b):
    return len(a) == len(b) and a.tobytes() == b.tobytes()


def bad_pixels(rng, h, w):
    mask = rng.random((h, w)) < 0.01
    mask[100:103, :] = True  # whole rows and columns of bad pixels
    mask[:, 200:205] = True
    return mask


def test_stamp_then_threshold_u16(sf):
    h, w, n = 300, 412, 6  # h * w is not a multiple of 32 or 64
    buf = sf.DeviceBuffer(2 * h * w * n)
    # the series' 1500 spots a 4362 x 4148 frame, ten times denser by area
    spots = fr.diffraction_spots(11, h, w, n,
                                 1500.0 * 10 * h * w / (4362 * 4148))
    sf.synth_frames(buf, 2, h, w, n, 11, 0.3, True, 0.0005, spots)
    host = np.zeros((n, h, w), dtype=np.uint16)
    buf.download(host)
    mask = bad_pixels(np.random.default_rng(12), h, w)
    pm = sf.PixelMask(mask.astype(np.uint8))
    ptrs = [buf.ptr + 2 * h * w * f for f in range(n)]
    pm.stamp(2, ptrs)
    sf.lib().sf_hip_synchronize()
    got = np.zeros_like(host)
    buf.download(got)
    want = np.where(mask[None], np.uint16(0xFFFF), host)
    assert np.array_equal(got, want)
    plan = sf.Plan(2, [(h, w)] * n)
    plan.run(ptrs)
    plan.wait()
    counts = plan.counts()
    for f in range(n):
        assert 0 < counts[f] <= h * w * 2 // 16  # the list fits
        status, expected, _ = cpu.dext(want[f])
        assert status == 0
        assert same(plan.output(f, int(counts[f])), expected), f
        # nothing masked is reported
        idx = plan.output(f, int(counts[f]))["index"]
        assert not mask.reshape(-1)[idx].any()
    plan.close()
    pm.close()
    buf.close()


def test_stamp_then_threshold_u32(sf):
    base = fr.u32_frame()
    h, w = base.shape
    n = 3
    host = np.stack([np.roll(base, 7 * f, axis=1) for f in range(n)])
    buf = sf.DeviceBuffer(host.nbytes)
    buf.upload(host)
    mask = bad_pixels(np.random.default_rng(13), h, w)
    pm = sf.PixelMask(mask)
    ptrs = [buf.ptr + 4 * h * w * f for f in range(n)]
    pm.stamp(4, ptrs)
    sf.lib().sf_hip_synchronize()
    got = np.zeros_like(host)
    buf.download(got)
    want = np.where(mask[None], np.uint32(0xFFFFFFFF), host)
    assert np.array_equal(got, want)
    plan = sf.Plan(4, [(h, w)] * n)
    plan.run(ptrs)
    plan.wait()
    counts = plan.counts()
    for f in range(n):
        assert 0 < counts[f] <= h * w * 4 // 16  # the list fits
        status, expected, _ = cpu.dext(want[f])
        assert status == 0
        assert same(plan.output(f, int(counts[f])), expected), f
    plan.close()
    pm.close()
    buf.close()


@pytest.mark.parametrize("pixel_bytes", [2, 4])
def test_stamp_unaligned_frames(sf, pixel_bytes):
    """Frames one pixel past 16-byte alignment take the per-pixel stores
    even for fully masked bytes; the pixels between and after the frames
    (h * w is not a multiple of 8) stay untouched."""
    dt = np.uint16 if pixel_bytes == 2 else np.uint32
    h, w, n = 37, 211, 5
    stride = h * w + 9  # frame f starts at element 1 + f * stride
    rng = np.random.default_rng(pixel_bytes)
    host = rng.integers(0, 1000, size=1 + n * stride).astype(dt)
    buf = sf.DeviceBuffer(host.nbytes)
    buf.upload(host)
    mask = bad_pixels(rng, h, w) if h > 103 else rng.random((h, w)) < 0.05
    mask[3:5, :] = True
    mask[:, 17:41] = True
    pm = sf.PixelMask(mask)
    ptrs = [buf.ptr + pixel_bytes * (1 + f * stride) for f in range(n)]
    pm.stamp(pixel_bytes, ptrs)
    sf.lib().sf_hip_synchronize()
    got = np.zeros_like(host)
    buf.download(got)
    want = host.copy()
    for f in range(n):
        at = 1 + f * stride
        view = want[at:at + h * w].reshape(h, w)
        view[mask] = np.iinfo(dt).max
    assert np.array_equal(got, want)
    pm.close()
    buf.close()


def test_pack_on_the_device(sf):
    for h, w in ((300, 412), (160, 203), (1, 33), (7, 9)):
        mask = (np.random.default_rng(h * w).random((h, w)) < 0.3)
        mask = mask.astype(np.uint8) * np.uint8(3)  # any nonzero is masked
        dev = sf.DeviceBuffer(h * w)
        dev.upload(mask)
        words = (h * w + 31) // 32
        bits = sf.DeviceBuffer(4 * words)
        sf.pack_mask(dev.ptr, h, w, bits.ptr)
        sf.lib().sf_hip_synchronize()
        got = np.zeros(words, dtype=np.uint32)
        bits.download(got)
        want = np.zeros(4 * words, dtype=np.uint8)
        packed = np.packbits(mask.reshape(-1) != 0, bitorder="little")
        want[:len(packed)] = packed
        assert np.array_equal(got, want.view(np.uint32)), (h, w)
        dev.close()
        bits.close()


def test_bad_arguments(sf):
    pm = sf.PixelMask(np.ones((4, 4), np.uint8))
    with pytest.raises(ValueError):
        pm.stamp(3, [pm.bits.ptr])
    with pytest.raises(ValueError):
        pm.stamp(2, [0])
    with pytest.raises(ValueError):
        sf.PixelMask(np.ones(5, np.uint8))
    pm.close()
