"""Frames the tests share.

* `reference_test_frames()` -- the seven frames of the reference's device
  parity test (tests/test_dext_gpu.cc:153-177), made by its own generator
  (tests/synthetic_frame.hh), restated in the oracle.
* `nxmx_fixture()` -- the 12-frame, 256 x 384 series of
  tests/make_test_nxmx.py:45-98 (a deterministic numpy formula), whose
  published answer is 8,928 signal pixels and 48 spots (README.md:95-107).
* `planted_series()` -- the 12-frame CPU workload: Poisson(1) background
  with 48 Gaussian
  reflections spanning frames (sigma_xy 1.0-1.5, sigma_z 0.7, 500-5000
  counts), seeded.
* `diffraction_spots()` -- a spot list for the device generator (the value
  model of the device workloads): ~spots_per_frame spots a frame, sigma 0.8-2 px,
  spanning 1-5 frames, peak heights log-uniform up to `peak`.
"""

from __future__ import annotations

import numpy as np

from oracle import cpu


def reference_test_frames():
    specs = [
        ("exact tile", 32, 32, 1, 12.0, False),
        ("under one tile", 19, 27, 2, 8.0, False),
        ("ragged", 200, 301, 3, 15.0, False),
        ("ragged, masked", 200, 301, 4, 15.0, True),
        ("dark", 128, 128, 5, 1.0, False),
        ("module", 512, 1030, 6, 20.0, True),
    ]
    frames = [(name, cpu.synthetic_frame(h, w, seed, bg, mask))
              for name, h, w, seed, bg, mask in specs]
    frames.append(("entirely masked", np.full((64, 64), 0xFFFF, np.uint16)))
    return frames


def nxmx_fixture(frames: int = 12) -> np.ndarray:
    """tests/make_test_nxmx.py: HEIGHT 256, WIDTH 384, BACKGROUND 2, SIGMA 1.5,
    SPAN 3, CURVE (0.35, 1, 0.35), 3-sigma truncation, rint, clip 65533."""
    height, width, sigma, span = 256, 384, 1.5, 3
    curve = (0.35, 1.0, 0.35)
    peaks = (250.0, 400.0, 700.0, 1000.0, 1500.0, 2000.0, 2500.0)
    spots = []
    index = 0
    for slow in range(20, height - 20, 40):
        for fast in range(20, width - 20, 48):
            first = 1 + (index % max(1, frames - span - 1))
            spots.append((slow, fast, peaks[index % len(peaks)], first))
            index += 1
    stack = np.full((frames, height, width), 2.0, dtype=np.float64)
    rows = np.arange(height)[:, None]
    columns = np.arange(width)[None, :]
    for slow, fast, peak, first in spots:
        profile = np.exp(-((rows - slow) ** 2 + (columns - fast) ** 2)
                         / (2.0 * sigma * sigma))
        profile[profile < np.exp(-4.5)] = 0.0
        for step in range(span):
            frame = first + step
            if frame >= frames:
                continue
            stack[frame] += peak * curve[step] * profile
    return np.clip(np.rint(stack), 0, 65533).astype(np.uint16)


def planted_series(seed: int = 2024, frames: int = 12, height: int = 256,
                   width: int = 384, reflections: int = 48):
    """The 12-frame CPU workload: returns (stack, planted list of (x, y, z))."""
    rng = np.random.default_rng(seed)
    expected = np.ones((frames, height, width), dtype=np.float64)
    rows = np.arange(height)[:, None] + 0.5
    columns = np.arange(width)[None, :] + 0.5
    planted = []
    for _ in range(reflections):
        x = rng.uniform(8, width - 8)
        y = rng.uniform(8, height - 8)
        z = rng.uniform(1.5, frames - 1.5)
        sigma = rng.uniform(1.0, 1.5)
        total = rng.uniform(500, 5000)
        profile = np.exp(-((rows - y) ** 2 + (columns - x) ** 2)
                         / (2 * sigma * sigma)) / (2 * np.pi * sigma * sigma)
        for f in range(frames):
            weight = np.exp(-((f + 0.5 - z) ** 2) / (2 * 0.7 * 0.7)) / (
                np.sqrt(2 * np.pi) * 0.7)
            if weight > 1e-4:
                expected[f] += total * weight * profile
        planted.append((x, y, z))
    stack = rng.poisson(expected).clip(0, 65533).astype(np.uint16)
    return stack, planted


def diffraction_spots(seed: int, height: int, width: int, frames: int,
                      spots_per_frame: float, peak: float = 20000.0,
                      min_peak: float = 30.0) -> np.ndarray:
    """(n, 6) float64 rows of (x, y, sigma, total, first frame, span)."""
    rng = np.random.default_rng(seed)
    n = int(round(spots_per_frame * frames / 3.0))
    x = rng.uniform(0, width, n)
    y = rng.uniform(0, height, n)
    sigma = rng.uniform(0.8, 2.0, n)
    span = rng.integers(1, 6, n)
    first = rng.integers(-2, frames, n)
    height_px = np.exp(rng.uniform(np.log(min_peak), np.log(peak), n))
    # total counts so that the brightest pixel of the brightest frame is
    # about height_px: 2 pi sigma^2 / (rocking weight of the central frame)
    central = np.where(span == 1, 1.0, 1.0 / np.sqrt(np.minimum(span, 3)))
    total = height_px * 2 * np.pi * sigma ** 2 / central
    return np.stack([x, y, sigma, total, first.astype(float),
                     span.astype(float)], axis=1)


def wrap_frame() -> np.ndarray:
    """A u16 frame whose busy windows wrap the 32-bit sum of squares: two
    rows of 20,000-count pixels (11 of them already exceed 2^32)."""
    rng = np.random.default_rng(7)
    frame = rng.poisson(0.3, size=(96, 128)).astype(np.uint16)
    frame[40:42, 30:60] = 20000
    frame[70, 10:20] = 40000
    frame[10:13, 100:103] = rng.integers(3000, 9000, size=(3, 3))
    return frame


def u32_frame() -> np.ndarray:
    rng = np.random.default_rng(11)
    frame = rng.poisson(50, size=(160, 203)).astype(np.uint32)
    for (y, x, peak) in ((20, 30, 1e6), (80, 100, 3e8), (140, 190, 4e9),
                         (5, 5, 1e4), (100, 20, 2e5)):
        yy, xx = np.mgrid[-3:4, -3:4]
        stamp = peak * np.exp(-(yy ** 2 + xx ** 2) / 3.0)
        region = frame[max(y - 3, 0):y + 4, max(x - 3, 0):x + 4]
        s = stamp[3 - (y - max(y - 3, 0)):, 3 - (x - max(x - 3, 0)):]
        s = s[:region.shape[0], :region.shape[1]]
        frame[max(y - 3, 0):y + 4, max(x - 3, 0):x + 4] = np.minimum(
            region.astype(np.float64) + s, 0xFFFFFFFD).astype(np.uint32)
    frame[60, :] = 0xFFFFFFFF
    frame[61, 7] = 0xFFFFFFFE
    return frame
