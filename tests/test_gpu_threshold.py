"""Device threshold parity: every SignalPixel field equal to the oracle and to
the reference's golden vectors (tests/test_dext_gpu.cc's bar: no tolerance).

Sizes stay where the CPU oracle finishes in seconds; full-size frames are
covered by the full-size test (two frames, oracle ~1 s each).
"""

from __future__ import annotations

import os
import threading

import numpy as np
import pytest

from oracle import cpu
import frames as fr

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def same(a, b):
    return len(a) == len(b) and a.tobytes() == b.tobytes()


def describe(a, b):
    if len(a) != len(b):
        return f"{len(a)} vs {len(b)} pixels"
    for k in range(len(a)):
        if a[k].tobytes() != b[k].tobytes():
            return f"entry {k}: {a[k]} vs {b[k]}"
    return "equal"


@pytest.mark.parametrize("window0", [0, 1])
@pytest.mark.parametrize("window2", [0, 1])
def test_reference_test_frames_every_window(sf, window0, window2):
    """tests/test_dext_gpu.cc:141-207: seven frames x four windows."""
    sf.stage0_window(window0)
    sf.stage2_window(window2)
    try:
        for name, frame in fr.reference_test_frames():
            expected = cpu.dext(frame)[1]
            status, got = sf.find(frame)
            assert status == 0, name
            assert same(got, expected), (name, describe(got, expected))
    finally:
        sf.stage0_window(1)
        sf.stage2_window(1)


def test_golden_vectors(sf):
    data = np.load(os.path.join(GOLDEN, "threshold.npz"))
    cases = dict(fr.reference_test_frames())
    cases["wrap"] = fr.wrap_frame()
    cases["u32"] = fr.u32_frame()
    for k in range(int(data["n_threshold"])):
        name = str(data[f"t{k}_name"])
        expected = data[f"t{k}_signal"]
        for form in (1, 0):
            sf.stage0_window(form)
            sf.stage2_window(form)
            status, got = sf.find(cases[name])
            assert status == 0
            assert same(got, expected), (name, form, describe(got, expected))
    sf.stage0_window(1)
    sf.stage2_window(1)


def random_frame(rng, shape, dtype, mean):
    frame = rng.poisson(mean, size=shape).astype(dtype)
    top = np.iinfo(dtype).max
    for _ in range(max(1, frame.size // 250)):
        y, x = rng.integers(0, shape[0]), rng.integers(0, shape[1])
        r = int(rng.integers(0, 3))
        bump = int(rng.integers(20, 20000 if dtype == np.uint16 else 10 ** 8))
        region = frame[max(y - r, 0):y + r + 1, max(x - r, 0):x + r + 1]
        region[...] = np.minimum(region.astype(np.int64) + bump, top - 2)
    if shape[0] > 8:
        frame[int(rng.integers(0, shape[0])), :] = top
        frame[:, int(rng.integers(0, shape[1]))] = top - 1
    return frame


def test_random_frames_edges_and_both_paths(sf):
    """Odd widths take the per-pixel load path, multiples of the lane width
    the vector path; tiny frames run every window off every edge."""
    rng = np.random.default_rng(12)
    shapes = [(1, 1), (1, 9), (9, 1), (2, 3), (5, 5), (7, 300), (300, 7),
              (33, 65), (64, 64), (129, 257), (250, 232), (250, 233),
              (251, 464), (257, 108), (31, 109)]
    for shape in shapes:
        for dtype, mean in ((np.uint16, 0.3), (np.uint16, 30.0),
                            (np.uint32, 500.0)):
            frame = random_frame(rng, shape, dtype, mean)
            expected = cpu.dext(frame)[1]
            status, got = sf.find(frame)
            # the device list holds height*width*sizeof(T)/16 entries, which
            # on a tiny frame can be fewer than the signal: then -2
            if len(expected) > frame.nbytes // 16:
                assert status == -2, shape
                continue
            assert status == 0, shape
            assert same(got, expected), (shape, dtype, describe(got, expected))


@pytest.mark.parametrize("form", [0, 2])
def test_random_frames_device_resident(sf, form):
    """The same edge cases through device-resident plans: the sparse passes
    (form 0) and the register-sweep kernel (form 2), including widths that
    are not a multiple of 4 and an unaligned frame start."""
    rng = np.random.default_rng(40 + form)
    shapes = [(1, 1), (2, 3), (9, 1), (5, 5), (7, 300), (300, 7), (33, 65),
              (64, 64), (129, 257), (250, 232), (250, 233), (251, 464),
              (97, 105), (128, 104), (65, 209)]
    for shape in shapes:
        for dtype, mean in ((np.uint16, 0.3), (np.uint16, 30.0),
                            (np.uint32, 500.0)):
            frame = random_frame(rng, shape, dtype, mean)
            expected = cpu.dext(frame)[1]
            pb = frame.dtype.itemsize
            for offset in (0, pb):  # aligned, and one pixel in
                buf = sf.DeviceBuffer(frame.nbytes + 16)
                buf.upload(frame, offset=offset)
                plan = sf.Plan(pb, [shape], form=form)
                plan.run([buf.ptr + offset])
                plan.wait()
                n = int(plan.counts()[0])
                if len(expected) > frame.nbytes // 16:
                    assert n > plan.capacity(0), shape
                else:
                    got = plan.output(0, n)
                    assert same(got, expected), (shape, dtype, offset,
                                                 describe(got, expected))
                plan.close()
                buf.close()


def test_bad_arguments_and_empty(sf):
    assert sf.find(np.zeros((0, 4), np.uint16))[0] == -1
    assert sf.find(np.zeros((4, 0), np.uint32))[0] == -1
    status, got = sf.find(np.full((64, 64), 0xFFFF, np.uint16))
    assert status == 0 and len(got) == 0
    status, got = sf.find(np.full((64, 64), 0xFFFFFFFE, np.uint32))
    assert status == 0 and len(got) == 0
    status, got = sf.find(np.zeros((300, 301), np.uint16))
    assert status == 0 and len(got) == 0


def test_overflow_returns_minus_two(sf):
    """More signal than height*width*sizeof(T)/16 entries -> -2
    (src/dext_gpu.hh:148-151), exactly when the oracle's count says so.
    Noise up to 5000 counts (below the 32-bit square-sum wrap) flags every
    pixel, the erode then leaves no background, and every p > 0 survives."""
    rng = np.random.default_rng(3)
    frame = rng.integers(0, 5000, size=(96, 128)).astype(np.uint16)
    capacity = frame.size * 2 // 16
    assert cpu.dext(frame)[2] > capacity
    for form in (1, 0):
        sf.stage0_window(form)
        sf.stage2_window(form)
        assert sf.find(frame)[0] == -2
        # one side dense, the rest a Poisson background: just under and just
        # over the list's capacity
        for columns, expect in ((14, 0), (18, -2)):
            part = frame.copy()
            part[:, columns:] = rng.poisson(2.0, size=part[:, columns:].shape)
            n = cpu.dext(part)[2]
            assert (n > capacity) == (expect == -2)
            status, got = sf.find(part)
            assert status == expect
            if expect == 0:
                assert same(got, cpu.dext(part)[1])
    sf.stage0_window(1)
    sf.stage2_window(1)


def model_frames(sf, n_frames, height, width, seed, pixel_bytes=2,
                   background=0.3, spots_per_frame=1500.0):
    buf = sf.DeviceBuffer(pixel_bytes * height * width * n_frames)
    spots = fr.diffraction_spots(seed, height, width, n_frames,
                                 spots_per_frame * (height * width) /
                                 (4362 * 4148),
                                 peak=20000.0 if pixel_bytes == 2 else 1e7)
    sf.synth_frames(buf, pixel_bytes, height, width, n_frames, seed,
                    background, True, 0.0005, spots)
    dtype = np.uint16 if pixel_bytes == 2 else np.uint32
    host = np.zeros((n_frames, height, width), dtype=dtype)
    buf.download(host)
    return buf, host


def test_full_size_model_frames_fused_and_plain(sf):
    """Two 4362 x 4148 frames of the 1800-frame series value model (gaps, hot pixels,
    20,000-count peaks that wrap the 32-bit sum of squares)."""
    h, w = 4362, 4148
    buf, host = model_frames(sf, 2, h, w, seed=101)
    expected = [cpu.dext(host[f], capacity=h * w // 8)[1] for f in range(2)]
    assert all(len(e) > 1000 for e in expected)
    for form in (0, 1, 2):
        plan = sf.Plan(2, [(h, w)] * 2, form=form)
        plan.run([buf.ptr, buf.ptr + 2 * h * w])
        plan.wait()
        counts = plan.counts()
        for f in range(2):
            got = plan.output(f, int(counts[f]))
            assert same(got, expected[f]), (form, f,
                                             describe(got, expected[f]))
        plan.close()


def test_batched_uneven_sizes_graph_and_persistent(sf):
    """The uneven-chunk shape mix (2162x2068, 3269x3110, 4362x4148, scaled
    to keep the oracle quick) in one launch, replayed from a captured graph
    with eager work between replays, through a persistent grid, and after
    the autotuner has picked its launch shape (sparse and sweep
    kernels)."""
    shapes = [(216, 206), (327, 311), (436, 414), (216, 206), (101, 3)]
    bufs, hosts = [], []
    for k, (h, w) in enumerate(shapes):
        b, host = model_frames(sf, 1, h, w, seed=200 + k,
                                 spots_per_frame=30000.0)
        bufs.append(b)
        hosts.append(host[0])
    expected = [cpu.dext(x)[1] for x in hosts]
    for x, e in zip(hosts, expected):
        assert len(e) <= x.nbytes // 16  # every frame fits its list
    sweep = sf.Plan(2, shapes, form=2)
    sweep.run([b.ptr for b in bufs])
    sweep.wait()
    counts = sweep.counts()
    for f in range(len(shapes)):
        assert same(sweep.output(f, int(counts[f])), expected[f]), ("sweep", f)
    sweep.close()
    plan = sf.Plan(2, shapes)
    ptrs = [b.ptr for b in bufs]
    plan.run(ptrs)
    plan.wait()
    counts = plan.counts()
    for f in range(len(shapes)):
        assert same(plan.output(f, int(counts[f])), expected[f]), f
    plan.capture(ptrs)
    for replay in range(3):
        plan.replay()
        plan.wait()
        # eager work between replays
        assert sf.find(hosts[0])[0] == 0
        counts = plan.counts()
        for f in range(len(shapes)):
            assert same(plan.output(f, int(counts[f])), expected[f]), (replay, f)
    plan.set_grid(7)
    plan.run(ptrs)
    plan.wait()
    counts = plan.counts()
    for f in range(len(shapes)):
        assert same(plan.output(f, int(counts[f])), expected[f]), ("grid", f)
    for form in (-1, 3):  # no such forms
        with pytest.raises(ValueError, match="set_form"):
            plan.set_form(form)
    for form in (0, 2):
        plan.set_form(form)
        band, blocks = plan.autotune(ptrs, 1)
        assert blocks >= 0 and band > 0
        plan.capture(ptrs)
        plan.replay()
        plan.wait()
        counts = plan.counts()
        for f in range(len(shapes)):
            assert same(plan.output(f, int(counts[f])), expected[f]), (
                "autotuned", form, band, blocks, f)
    plan.close()


def test_band_heights(sf):
    rng = np.random.default_rng(8)
    frame = random_frame(rng, (700, 520), np.uint16, 5.0)
    expected = cpu.dext(frame)[1]
    buf = sf.DeviceBuffer(frame.nbytes)
    buf.upload(frame)
    for band in (1, 12, 44, 236, 492, 1000):
        plan = sf.Plan(2, [frame.shape], band_rows=band, form=2)
        plan.run([buf.ptr])
        plan.wait()
        assert same(plan.output(0), expected), band
        plan.close()


def test_profiling_and_threads(sf):
    frames = [f for _, f in fr.reference_test_frames()]
    expected = [cpu.dext(f)[1] for f in frames]
    sf.profile_stages(True)
    status, _ = sf.find(frames[5])
    times = sf.last_stage_times()
    sf.profile_stages(False)
    assert status == 0 and times["stage0"] > 0 and times["upload"] > 0
    failures = []

    def work(offset):
        for r in range(3):
            for k in range(len(frames)):
                j = (k + offset) % len(frames)
                status, got = sf.find(frames[j])
                if status != 0 or not same(got, expected[j]):
                    failures.append((offset, r, j))

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not failures
