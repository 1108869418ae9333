"""The CPU oracle, pinned: against the reference's compiled CPU path, its
golden vectors, its published known answers and hand-derived cases, an
independent direct-window summation, and brute force on tiny inputs.

Reference tests mirrored here: tests/test_dext_squares.cc,
tests/test_signal_order.cc, tests/test_dials_spots.cc,
tests/test_dials_end_to_end.cc, and the published fixture answer of
README.md:95-107.
"""

from __future__ import annotations

import hashlib
import os
from collections import deque

import numpy as np
import pytest

from oracle import cpu
import frames as fr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
needs_ref = pytest.mark.skipif(not cpu.have_ref(),
                               reason="oracle/_ref not built here")


def sha(array):
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def golden_threshold_cases():
    data = np.load(os.path.join(GOLDEN, "threshold.npz"))
    frames = dict(fr.reference_test_frames())
    frames["wrap"] = fr.wrap_frame()
    frames["u32"] = fr.u32_frame()
    for k in range(int(data["n_threshold"])):
        name = str(data[f"t{k}_name"])
        yield name, frames[name], str(data[f"t{k}_sha"]), data[f"t{k}_signal"]


# ---------------------------------------------------------------------------
# Threshold
# ---------------------------------------------------------------------------

def test_oracle_matches_reference_golden_vectors():
    seen = 0
    for name, frame, digest, expected in golden_threshold_cases():
        assert sha(frame) == digest, f"{name}: input generator drifted"
        status, got, found = cpu.dext(frame)
        assert status == 0 and found == len(expected), name
        assert got.tobytes() == expected.tobytes(), name
        status, got, found = cpu.dext_direct(frame)
        assert got.tobytes() == expected.tobytes(), f"{name} (direct form)"
        seen += 1
    assert seen == 9


@needs_ref
def test_oracle_matches_compiled_reference_on_random_frames():
    rng = np.random.default_rng(5)
    shapes = [(1, 1), (1, 17), (23, 1), (2, 2), (7, 7), (11, 13), (64, 64),
              (97, 130), (130, 97)]
    for shape in shapes:
        for dtype, mean in ((np.uint16, 3.0), (np.uint16, 40.0),
                            (np.uint32, 200.0)):
            frame = rng.poisson(mean, size=shape).astype(dtype)
            top = np.iinfo(dtype).max
            # bright blobs, masked lines and both sentinels
            for _ in range(max(1, frame.size // 300)):
                y, x = rng.integers(0, shape[0]), rng.integers(0, shape[1])
                frame[y, x] = min(int(frame[y, x]) + int(rng.integers(50, 3000)),
                                  top - 2)
            if shape[0] > 5:
                frame[shape[0] // 2, :] = top
                frame[rng.integers(0, shape[0]), rng.integers(0, shape[1])] = top - 1
            a = cpu.dext(frame)
            b = cpu.ref_dext(frame)
            assert a[0] == b[0] == 0
            assert a[1].tobytes() == b[1].tobytes(), (shape, dtype)


def test_published_fixture_known_answer():
    """README.md:95-107: the 12-frame make_test_nxmx.py series gives 8928
    signal pixels and 48 spots, none removed."""
    stack = fr.nxmx_fixture()
    lab = cpu.Labeller(256, 384)
    total = 0
    for z in range(12):
        status, signal, found = cpu.dext(stack[z])
        assert status == 0
        total += found
        assert lab.add(z, signal) == 0
    lab.finish()
    spots, pixels, counts = lab.results()
    assert total == 8928
    assert len(spots) == 48
    assert counts == {"signal_pixels": 8928, "groups": 48, "too_small": 0,
                      "too_large": 0, "separated": 0, "accepted": 48}


def largest_window_square(frame, radius=3):
    """Brute-force 64-bit maximum of the clipped window sum of squares."""
    f = frame.astype(np.uint64)
    masked = frame >= np.iinfo(frame.dtype).max - 1
    f[masked] = 0
    sq = f * f
    h, w = frame.shape
    best = 0
    padded = np.zeros((h + 2 * radius, w + 2 * radius), dtype=np.uint64)
    padded[radius:radius + h, radius:radius + w] = sq
    acc = np.zeros((h, w), dtype=np.uint64)
    for dy in range(2 * radius + 1):
        for dx in range(2 * radius + 1):
            acc += padded[dy:dy + h, dx:dx + w]
    best = int(acc.max())
    return best


def test_sum_of_squares_headroom_known_answers():
    """tests/test_dext_squares.cc:119-193."""
    limit = 1 << 32
    assert 65533 * 65533 < limit
    uniform = int(np.sqrt(limit // 49))
    assert uniform == 9362 and uniform > 5000
    frame = np.full((120, 160), 2, dtype=np.uint16)
    peaks = (250, 400, 700, 1000, 1400, 1800, 2200, 2500, 3000)

    def plant(slow, fast, peak):
        for di in range(-3, 4):
            for dj in range(-3, 4):
                i, j = slow + di, fast + dj
                if 0 <= i < 120 and 0 <= j < 160:
                    v = peak * np.exp(-(di * di + dj * dj) / (2 * 1.5 * 1.5))
                    total = frame[i, j] + v
                    frame[i, j] = int(min(total, 65533.0))

    for at, peak in enumerate(peaks):
        plant(10 + 12 * at, 10 + 17 * at, peak)
    plant(60, 100, 2500.0)
    plant(60, 102, 2500.0)
    largest = largest_window_square(frame)
    assert largest * 10 < limit
    status, signal, found = cpu.dext(frame)
    assert status == 0 and found > 0
    sentinels = np.full((120, 160), 0xFFFF, dtype=np.uint16)
    sentinels.reshape(-1)[::2] = 0xFFFE
    assert largest_window_square(sentinels) == 0
    assert cpu.dext(sentinels)[2] == 0


def test_wrap_fixture_does_wrap():
    """The golden 'wrap' frame must really overflow the 32-bit window sum of
    squares, so that its golden vector pins the wrapping arithmetic."""
    assert largest_window_square(fr.wrap_frame()) > (1 << 32)


def test_threshold_invariants():
    for name, frame in fr.reference_test_frames():
        status, signal, found = cpu.dext(frame)
        assert status == 0
        idx = signal["index"].astype(np.int64)
        assert np.all(np.diff(idx) > 0), name
        flat = frame.reshape(-1)
        assert np.all(signal["value"] > 0)
        assert np.all(signal["value"] == flat[idx])
        assert np.all(flat[idx] < 0xFFFE)
        assert np.all(signal["population"] <= 121)
        assert np.all(signal["background"] >= 0)
        assert np.all(signal["reserved"] == 0)


def test_bad_arguments():
    for fn in (cpu.dext, cpu.dext_direct) + ((cpu.ref_dext,) if cpu.have_ref() else ()):
        assert fn(np.zeros((0, 5), np.uint16))[0] == -1
        assert fn(np.zeros((5, 0), np.uint16))[0] == -1


# ---------------------------------------------------------------------------
# Ordering (tests/test_signal_order.cc)
# ---------------------------------------------------------------------------

def make(index):
    p = np.zeros(len(index), dtype=cpu.SIGNAL)
    index = np.asarray(index, dtype=np.uint32)
    p["index"] = index
    p["value"] = index * np.uint32(7) + np.uint32(1)
    p["background"] = (index % 1000).astype(np.float32) + np.float32(0.5)
    p["population"] = (index % 121).astype(np.uint16)
    return p


def emit_order(height, width, n, seed):
    rng = np.random.default_rng(seed)
    taken = set()
    out = []
    while len(out) < n:
        r = int(rng.integers(0, height))
        t = int(rng.integers(0, width // 32))
        length = int(rng.integers(1, 13))
        for c in range(length):
            k = r * width + t * 32 + c
            if len(out) >= n:
                break
            if k >= height * width or k in taken:
                continue
            taken.add(k)
            out.append(k)
    return make(out)


def ordering_cases():
    h, w = 4362, 4148
    yield "empty", make([])
    yield "one", make([17])
    yield "two reversed", make([9, 4])
    for n in (1023, 1024, 1025):
        yield f"{n} emit order", emit_order(h, w, n, 3)
    yield "126002", emit_order(h, w, 126002, 7)
    both = [(n // 2) if n % 2 == 0 else h * w - 1 - n // 2 for n in range(2000)]
    yield "both ends", make(both)
    yield "5000 reversed", make(list(range(4999, -1, -1)))
    yield "duplicates", make([n % 40 for n in range(3000)])
    yield "4000", emit_order(h, w, 4000, 5)


def test_by_index_equals_stable_sort():
    for name, pixels in ordering_cases():
        expected = pixels[np.argsort(pixels["index"], kind="stable")]
        got = cpu.by_index(pixels)
        assert got.tobytes() == expected.tobytes(), name
        if cpu.have_ref():
            assert cpu.ref_by_index(pixels).tobytes() == expected.tobytes(), name


# ---------------------------------------------------------------------------
# Grouping (tests/test_dials_spots.cc)
# ---------------------------------------------------------------------------

def pixels(pairs):
    p = np.zeros(len(pairs), dtype=cpu.SIGNAL)
    for k, (i, v) in enumerate(pairs):
        p[k]["index"] = i
        p[k]["value"] = v
    return p


def flat(indices, value):
    return pixels([(i, value) for i in indices])


PLAIN = dict(max_separation=0.0)

labellers = [cpu.Labeller] + ([cpu.RefLabeller] if cpu.have_ref() else [])


@pytest.mark.parametrize("Lab", labellers)
def test_labeller_known_answers(Lab):
    H, W = 8, 10
    close = lambda a, b: abs(a - b) < 1e-9  # noqa: E731

    lab = Lab(H, W, **PLAIN)
    lab.add(0, pixels([(0, 1), (1, 2), (2, 1)]))
    lab.finish()
    spots, kept, counts = lab.results()
    assert len(spots) == 1
    s = spots[0]
    assert s["n_signal"] == 3 and close(s["intensity"], 4.0)
    assert close(s["intensity_variance"], 4.0)
    assert close(s["position"][0], 1.5) and close(s["position"][1], 0.5)
    assert close(s["position"][2], 0.5)
    assert close(s["variance"][0], 0.8 / 4 + 1 / 12)
    assert close(s["variance"][1], 1 / 12) and close(s["variance"][2], 1 / 12)
    assert list(s["bbox"]) == [0, 3, 0, 1, 0, 1]
    assert len(kept) == 3 and counts["groups"] == 1 and counts["too_small"] == 0

    lab = Lab(H, W, **PLAIN)
    for z in range(3):
        lab.add(z, flat([0, 1, 2], 1))
    lab.finish()
    spots, _, _ = lab.results()
    assert len(spots) == 1
    s = spots[0]
    assert s["n_signal"] == 9 and close(s["intensity"], 9.0)
    assert close(s["position"][0], 1.5) and close(s["position"][2], 1.5)
    assert close(s["variance"][0], 0.75 / 9 + 1 / 12)
    assert close(s["variance"][2], 0.75 / 9 + 1 / 12)
    assert s["bbox"][4] == 0 and s["bbox"][5] == 3

    lab = Lab(H, W, two_d=True, **PLAIN)
    for z in range(3):
        lab.add(z, flat([0, 1, 2], 1))
    lab.finish()
    spots, _, _ = lab.results()
    assert len(spots) == 3
    assert close(spots[1]["position"][2], 1.5)
    assert spots[1]["bbox"][5] - spots[1]["bbox"][4] == 1

    lab = Lab(H, W, **PLAIN)
    for z in (0, 1, 3, 4):
        lab.add(z, flat([0, 1, 2], 1))
    lab.finish()
    spots, _, _ = lab.results()
    assert len(spots) == 2
    assert spots[0]["n_signal"] == 6 and spots[1]["n_signal"] == 6
    assert spots[1]["bbox"][4] == 3

    for idx, expect in (([9, 10], 2), ([0, 10], 1), ([0, 11], 2)):
        lab = Lab(H, W, min_spot_size=1, **PLAIN)
        lab.add(0, flat(idx, 1))
        lab.finish()
        assert len(lab.results()[0]) == expect, idx
    lab = Lab(H, W, min_spot_size=1, **PLAIN)
    lab.add(0, flat([0], 1))
    lab.add(1, flat([10], 1))
    lab.finish()
    assert len(lab.results()[0]) == 2

    lab = Lab(H, W, **PLAIN)
    lab.add(0, flat([0, 1], 1))
    lab.add(1, flat([40, 41, 42], 1))
    lab.finish()
    spots, _, counts = lab.results()
    assert len(spots) == 1 and counts["too_small"] == 1
    assert counts["groups"] == 2 and counts["signal_pixels"] == 5

    lab = Lab(H, W, max_spot_size=5, **PLAIN)
    lab.add(0, flat([0, 1, 2, 3, 4, 5], 1))
    lab.finish()
    spots, _, counts = lab.results()
    assert len(spots) == 0 and counts["too_large"] == 1

    lab = Lab(H, W, max_separation=2.0)
    lab.add(0, pixels([(0, 1000)] + [(x, 1) for x in range(1, 10)]))
    lab.finish()
    spots, _, counts = lab.results()
    assert len(spots) == 1 and counts["separated"] == 0

    lab = Lab(H, W, max_separation=2.0)
    lab.add(0, pixels([(0, 100), (1, 100), (2, 1), (3, 1), (4, 1), (5, 1),
                       (6, 90), (7, 90)]))
    lab.finish()
    spots, _, counts = lab.results()
    assert len(spots) == 0 and counts["separated"] == 1 and counts["groups"] == 1

    lab = Lab(H, W, min_spot_size=1, **PLAIN)
    lab.add(0, flat([35], 700))
    lab.finish()
    spots, _, _ = lab.results()
    assert len(spots) == 1
    s = spots[0]
    assert close(s["position"][0], 5.5) and close(s["position"][1], 3.5)
    assert close(s["variance"][0], 1 / 12) and close(s["variance"][1], 1 / 12)
    assert close(s["intensity"], 700.0)

    lab = Lab(H, W, **PLAIN)
    assert lab.add(5, flat([0, 1, 2], 1)) == 0
    assert lab.add(4, flat([0], 1)) != 0
    assert lab.add(5, flat([0], 1)) != 0
    lab = Lab(H, W, **PLAIN)
    assert lab.add(0, flat([2, 1], 1)) != 0
    lab = Lab(H, W, **PLAIN)
    assert lab.add(0, flat([H * W], 1)) != 0

    lab = Lab(H, W, **PLAIN)
    for z in range(200):
        lab.add(z, flat([44, 45, 46, 70, 71, 72], 3))
    lab.finish()
    spots, kept, _ = lab.results()
    assert len(spots) == 2
    assert spots[0]["n_signal"] == 600
    assert spots[0]["bbox"][5] - spots[0]["bbox"][4] == 200
    assert len(kept) == 1200


def test_labeller_rejects_bad_options():
    with pytest.raises(ValueError):
        cpu.Labeller(0, 5)
    with pytest.raises(ValueError):
        cpu.Labeller(5, 5, min_spot_size=5, max_spot_size=4)
    with pytest.raises(ValueError):
        cpu.Labeller(5, 5, min_spot_size=0)


def random_sequence(rng, height, width, frames, density, gap_every=0):
    seq = []
    z = int(rng.integers(-3, 3))
    for f in range(frames):
        mask = rng.random(height * width) < density
        idx = np.flatnonzero(mask).astype(np.uint32)
        p = np.zeros(len(idx), dtype=cpu.SIGNAL)
        p["index"] = idx
        p["value"] = rng.integers(1, 5000, len(idx))
        seq.append((z, p))
        z += 2 if (gap_every and f % gap_every == gap_every - 1) else 1
    return seq


def run_labeller(Lab, seq, height, width, **options):
    lab = Lab(height, width, **options)
    for z, p in seq:
        assert lab.add(z, p) == 0
    lab.finish()
    return lab.results()


@needs_ref
def test_labeller_matches_reference_on_random_sequences():
    rng = np.random.default_rng(9)
    option_sets = [{}, {"two_d": True}, {"max_separation": 0.0},
                   {"min_spot_size": 1, "max_spot_size": 20},
                   {"max_separation": 1.0}]
    for trial in range(24):
        h, w = int(rng.integers(3, 40)), int(rng.integers(3, 40))
        seq = random_sequence(rng, h, w, int(rng.integers(1, 9)),
                              float(rng.uniform(0.05, 0.45)),
                              gap_every=int(rng.integers(0, 4)))
        options = option_sets[trial % len(option_sets)]
        a = run_labeller(cpu.Labeller, seq, h, w, **options)
        b = run_labeller(cpu.RefLabeller, seq, h, w, **options)
        assert a[2] == b[2], (trial, options)
        assert a[0].tobytes() == b[0].tobytes(), (trial, options)
        assert a[1].tobytes() == b[1].tobytes(), (trial, options)


def flood_fill_components(seq, height, width, two_d):
    """Brute-force six-connected components over (frame, y, x)."""
    present = {}
    for z, p in seq:
        for i in p["index"]:
            present[(z, int(i) // width, int(i) % width)] = True
    seen = set()
    comps = []
    for key in sorted(present):
        if key in seen:
            continue
        comp = []
        queue = deque([key])
        seen.add(key)
        while queue:
            z, y, x = queue.popleft()
            comp.append((z, y, x))
            steps = [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0)]
            if not two_d:
                steps += [(1, 0, 0), (-1, 0, 0)]
            for dz, dy, dx in steps:
                nb = (z + dz, y + dy, x + dx)
                if 0 <= nb[1] < height and 0 <= nb[2] < width and \
                        nb in present and nb not in seen:
                    seen.add(nb)
                    queue.append(nb)
        comps.append(sorted(comp))
    return sorted(comps)


def test_labeller_components_match_flood_fill():
    rng = np.random.default_rng(4)
    for trial in range(30):
        h, w = int(rng.integers(2, 12)), int(rng.integers(2, 12))
        frames = int(rng.integers(1, 6))
        seq = random_sequence(rng, h, w, frames, float(rng.uniform(0.1, 0.6)),
                              gap_every=int(rng.integers(0, 3)))
        two_d = bool(trial % 3 == 0)
        spots, kept, counts = run_labeller(
            cpu.Labeller, seq, h, w, min_spot_size=1, max_spot_size=10 ** 6,
            max_separation=0.0, two_d=two_d)
        got = []
        for s in spots:
            members = kept[int(s["first"]): int(s["first"]) + int(s["n_signal"])]
            got.append(sorted((int(m["frame"]), int(m["index"]) // w,
                               int(m["index"]) % w) for m in members))
        assert sorted(got) == flood_fill_components(seq, h, w, two_d), trial
        assert counts["groups"] == len(got)


def test_labeller_matches_reference_golden_tables():
    data = np.load(os.path.join(GOLDEN, "labeller.npz"))
    nx = fr.nxmx_fixture()
    series, _ = fr.planted_series()
    stacks = {"nxmx": nx, "nxmx_2d": nx, "nxmx_holey": np.delete(nx, 5, axis=0),
              "planted": series, "planted_nosep": series}
    options = {"nxmx": {}, "nxmx_2d": {"two_d": True}, "nxmx_holey": {},
               "planted": {}, "planted_nosep": {"max_separation": 0.0,
                                                "min_spot_size": 1}}
    for k in range(int(data["n_labeller"])):
        name = str(data[f"l{k}_name"])
        stack = stacks[name]
        assert sha(stack) == str(data[f"l{k}_sha"]), name
        numbers = [int(z) for z in data[f"l{k}_numbers"]]
        lab = cpu.Labeller(stack.shape[1], stack.shape[2], **options[name])
        for z, frame in zip(numbers, stack):
            status, signal, _ = cpu.dext(frame)
            assert status == 0 and lab.add(z, signal) == 0
        lab.finish()
        spots, kept, counts = lab.results()
        assert spots.tobytes() == data[f"l{k}_spots"].tobytes(), name
        assert kept.tobytes() == data[f"l{k}_pixels"].tobytes(), name
        assert [counts[c] for c in cpu.COUNTS] == list(data[f"l{k}_counts"])


def test_end_to_end_planted_positions():
    """tests/test_dials_end_to_end.cc: 10 reflections rocking over frames
    1-3 come back where they were planted."""
    height, width = 240, 320
    planted = [(40, 50, 250.0), (40, 150, 500.0), (40, 250, 900.0),
               (120, 50, 1200.0), (120, 150, 1600.0), (120, 250, 2000.0),
               (200, 50, 2500.0), (200, 150, 700.0), (200, 250, 350.0),
               (5, 5, 1400.0)]
    curve = (0.0, 0.35, 1.0, 0.35, 0.0)
    lab = cpu.Labeller(height, width)
    for z in range(5):
        img = np.full((height, width), 2.0)
        for slow, fast, peak in planted:
            for di in range(-3, 4):
                for dj in range(-3, 4):
                    i, j = slow + di, fast + dj
                    if 0 <= i < height and 0 <= j < width:
                        v = peak * curve[z] * np.exp(-(di * di + dj * dj) / 4.5)
                        img[i, j] = min(img[i, j] + v, 65533.0)
        frame = img.astype(np.uint16)  # truncation, as static_cast does
        status, signal, _ = cpu.dext(frame)
        assert status == 0 and lab.add(z, signal) == 0
    lab.finish()
    spots, _, _ = lab.results()
    assert len(spots) == len(planted)
    for slow, fast, _ in planted:
        d = np.hypot(spots["position"][:, 0] - (fast + 0.5),
                     spots["position"][:, 1] - (slow + 0.5))
        best = spots[int(np.argmin(d))]
        assert abs(best["position"][0] - (fast + 0.5)) <= 0.5
        assert abs(best["position"][1] - (slow + 0.5)) <= 0.5
        assert abs(best["position"][2] - 2.5) <= 0.3
        assert best["bbox"][5] - best["bbox"][4] == 3
        assert best["intensity"] > best["n_signal"]
