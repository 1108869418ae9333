"""Device grouping parity: Spot and Pixel tables and Counts bit-identical to
the reference Labeller's (tests/test_dials_spots.cc's cases, the golden
tables written by the compiled reference, random sequences against the
oracle), through host lists and through device lists carried across chunks.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import cpu
import frames as fr
from test_oracle import random_sequence, flat, pixels, PLAIN

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def same_tables(a, b):
    return (a[0].tobytes() == b[0].tobytes() and
            a[1].tobytes() == b[1].tobytes() and a[2] == b[2])


def label_host(Lab, seq, h, w, **options):
    lab = Lab(h, w, **options)
    for z, p in seq:
        lab.add(z, p)
    lab.finish()
    return lab.results()


def test_known_answers(sf):
    """tests/test_dials_spots.cc:71-412 on the device labeller."""
    H, W = 8, 10
    close = lambda a, b: abs(a - b) < 1e-9  # noqa: E731
    spots, kept, counts = label_host(sf.Labeller, [(0, pixels([(0, 1), (1, 2),
                                                              (2, 1)]))],
                                     H, W, **PLAIN)
    s = spots[0]
    assert len(spots) == 1 and s["n_signal"] == 3
    assert close(s["position"][0], 1.5) and close(s["variance"][0],
                                                  0.8 / 4 + 1 / 12)
    assert list(s["bbox"]) == [0, 3, 0, 1, 0, 1] and len(kept) == 3
    seq = [(z, flat([0, 1, 2], 1)) for z in range(3)]
    spots, _, _ = label_host(sf.Labeller, seq, H, W, **PLAIN)
    assert len(spots) == 1 and spots[0]["n_signal"] == 9
    assert close(spots[0]["variance"][2], 0.75 / 9 + 1 / 12)
    spots, _, _ = label_host(sf.Labeller, seq, H, W, two_d=True, **PLAIN)
    assert len(spots) == 3
    seq = [(z, flat([0, 1, 2], 1)) for z in (0, 1, 3, 4)]
    spots, _, _ = label_host(sf.Labeller, seq, H, W, **PLAIN)
    assert [int(x) for x in spots["n_signal"]] == [6, 6]
    for idx, expect in (([9, 10], 2), ([0, 10], 1), ([0, 11], 2)):
        spots, _, _ = label_host(sf.Labeller, [(0, flat(idx, 1))], H, W,
                                 min_spot_size=1, **PLAIN)
        assert len(spots) == expect, idx
    # nothing emitted before it is finished; live() counts what is held
    lab = sf.Labeller(H, W, **PLAIN)
    lab.add(0, flat([0, 1, 2], 1))
    assert len(lab.results()[0]) == 0 and lab.live() == 3
    lab.add(1, flat([40, 41, 42], 1))
    assert len(lab.results()[0]) == 1 and lab.live() == 3
    lab.finish()
    assert len(lab.results()[0]) == 2 and lab.live() == 0
    # filters
    spots, _, counts = label_host(
        sf.Labeller, [(0, flat([0, 1], 1)), (1, flat([40, 41, 42], 1))], H, W,
        **PLAIN)
    assert len(spots) == 1 and counts["too_small"] == 1 and counts["groups"] == 2
    spots, _, counts = label_host(sf.Labeller, [(0, flat(range(6), 1))], H, W,
                                  max_spot_size=5, **PLAIN)
    assert len(spots) == 0 and counts["too_large"] == 1
    spots, _, counts = label_host(
        sf.Labeller, [(0, pixels([(0, 100), (1, 100), (2, 1), (3, 1), (4, 1),
                                  (5, 1), (6, 90), (7, 90)]))], H, W,
        max_separation=2.0)
    assert len(spots) == 0 and counts["separated"] == 1
    # refusals
    lab = sf.Labeller(H, W, **PLAIN)
    lab.add(5, flat([0, 1, 2], 1))
    for frame, p in ((4, flat([0], 1)), (5, flat([0], 1))):
        with pytest.raises(ValueError):
            lab.add(frame, p)
    with pytest.raises(ValueError):
        sf.Labeller(H, W, **PLAIN).add(0, flat([2, 1], 1))
    with pytest.raises(ValueError):
        sf.Labeller(H, W, **PLAIN).add(0, flat([H * W], 1))
    with pytest.raises(ValueError):
        sf.Labeller(H, W, min_spot_size=0)
    # a 200-frame streak, carried frame after frame
    seq = [(z, flat([44, 45, 46, 70, 71, 72], 3)) for z in range(200)]
    spots, kept, _ = label_host(sf.Labeller, seq, H, W, **PLAIN)
    assert len(spots) == 2 and spots[0]["n_signal"] == 600 and len(kept) == 1200


def test_random_sequences_match_oracle(sf):
    rng = np.random.default_rng(21)
    option_sets = [{}, {"two_d": True}, {"max_separation": 0.0},
                   {"min_spot_size": 1, "max_spot_size": 20},
                   {"max_separation": 1.0}]
    for trial in range(40):
        h, w = int(rng.integers(3, 60)), int(rng.integers(3, 60))
        seq = random_sequence(rng, h, w, int(rng.integers(1, 12)),
                              float(rng.uniform(0.02, 0.5)),
                              gap_every=int(rng.integers(0, 4)))
        options = option_sets[trial % len(option_sets)]
        a = label_host(cpu.Labeller, seq, h, w, **options)
        b = label_host(sf.Labeller, seq, h, w, **options)
        assert same_tables(a, b), (trial, options, a[2], b[2])


def test_large_values_and_large_components(sf):
    """Values where the double sums stop being exact (so the device takes
    the sequential member-order path) next to components of hundreds of
    members (several wave-sized chunks of the variance loop), against the
    oracle."""
    rng = np.random.default_rng(33)
    for trial, top in enumerate((65535, 2 ** 26, 2 ** 32 - 3, 3000)):
        h, w = 40, 50
        seq = []
        for z in range(6):
            mask = rng.random(h * w) < (0.45 if trial % 2 else 0.25)
            mask[5 * w:25 * w] = True  # a slab that joins into one big spot
            idx = np.flatnonzero(mask).astype(np.uint32)
            p = np.zeros(len(idx), dtype=cpu.SIGNAL)
            p["index"] = idx
            p["value"] = rng.integers(1, top, len(idx), dtype=np.uint64)
            seq.append((z, p))
        for options in ({}, {"max_spot_size": 100000}, {"two_d": True}):
            a = label_host(cpu.Labeller, seq, h, w, **options)
            b = label_host(sf.Labeller, seq, h, w, **options)
            assert same_tables(a, b), (trial, options, a[2], b[2])
            assert a[2]["accepted"] > 0 or a[2]["too_large"] > 0


def test_golden_tables_end_to_end(sf):
    """Device threshold then device grouping against the tables the
    compiled reference wrote."""
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
        numbers = [int(z) for z in data[f"l{k}_numbers"]]
        lab = sf.Labeller(stack.shape[1], stack.shape[2], **options[name])
        for z, frame in zip(numbers, stack):
            status, signal = sf.find(frame)
            assert status == 0
            lab.add(z, signal)
        lab.finish()
        spots, kept, counts = lab.results()
        assert spots.tobytes() == data[f"l{k}_spots"].tobytes(), name
        assert kept.tobytes() == data[f"l{k}_pixels"].tobytes(), name
        assert [counts[c] for c in cpu.COUNTS] == list(data[f"l{k}_counts"])
        if name == "nxmx":
            assert counts["signal_pixels"] == 8928 and len(spots) == 48


def test_device_lists_across_chunks(sf):
    """Frames thresholded in plans of 4, labelled straight from the device
    lists with the frontier carried between chunks, against the oracle fed
    the same frames one at a time."""
    h, w, n = 300, 412, 12
    buf = sf.DeviceBuffer(2 * h * w * n)
    spots = fr.diffraction_spots(5, h, w, n, 1500.0 * 40)
    sf.synth_frames(buf, 2, h, w, n, 5, 0.3, True, 0.0005, spots)
    host = np.zeros((n, h, w), dtype=np.uint16)
    buf.download(host)
    for two_d in (False, True):
        oracle = cpu.Labeller(h, w, two_d=two_d)
        for z in range(n):
            assert oracle.add(100 + z, cpu.dext(host[z])[1]) == 0
        oracle.finish()
        expected = oracle.results()
        plan = sf.Plan(2, [(h, w)] * 4)
        lab = sf.Labeller(h, w, two_d=two_d)
        for c in range(n // 4):
            plan.run([buf.ptr + 2 * h * w * (4 * c + k) for k in range(4)])
            lab.add_plan(plan, 100 + 4 * c)
        lab.finish()
        got = lab.results()
        assert got[2] == expected[2], (two_d, got[2], expected[2])
        assert same_tables(got, expected), two_d
        assert expected[2]["accepted"] > 20
        plan.close()


def test_reserved_tables(sf):
    """Result rows spread over table segments in every way reserve() can
    leave them: reserved before the first chunk (too small, so the chunks
    then grow the table by several segments), reserved in the middle of a
    series (a new segment while the last still has room), reserved twice
    (the second a no-op); the tables read back equal the oracle's."""
    h, w, n, per = 300, 412, 16, 4
    buf = sf.DeviceBuffer(2 * h * w * n)
    spots = fr.diffraction_spots(7, h, w, n, 1500.0 * 40)
    sf.synth_frames(buf, 2, h, w, n, 7, 0.3, True, 0.0005, spots)
    host = np.zeros((n, h, w), dtype=np.uint16)
    buf.download(host)
    oracle = cpu.Labeller(h, w)
    for z in range(n):
        assert oracle.add(z, cpu.dext(host[z])[1]) == 0
    oracle.finish()
    expected = oracle.results()
    assert len(expected[1]) > 4 * 1024
    plan = sf.Plan(2, [(h, w)] * per)
    for pattern in ("before", "middle", "twice"):
        lab = sf.Labeller(h, w)
        if pattern == "before":
            lab.reserve(1, 1)
        for c in range(n // per):
            plan.run([buf.ptr + 2 * h * w * (per * c + k) for k in range(per)])
            lab.add_plan(plan, per * c)
            if c == 1 and pattern != "before":
                lab.reserve(10_000, 100_000)
                assert lab.sizes()[1] > 0
                if pattern == "twice":
                    lab.reserve(10, 10)
        lab.finish()
        got = lab.results()
        assert got[2] == expected[2], (pattern, got[2], expected[2])
        assert same_tables(got, expected), pattern
        lab.close()
    plan.close()
    buf.close()


def test_chunks_past_the_merge_sort_limit(sf):
    """Chunks of more than 65,536 runs, where the component sort takes the
    radix (onesweep) path instead of rocPRIM's merge sort: device lists of
    24-frame plans against the oracle fed the same lists frame by frame."""
    h, w, n, per = 2048, 2048, 48, 24
    buf = sf.DeviceBuffer(2 * h * w * n)
    spots = fr.diffraction_spots(9, h, w, n, 1500.0 * h * w / (4362 * 4148))
    sf.synth_frames(buf, 2, h, w, n, 9, 0.3, True, 0.0005, spots)
    plan = sf.Plan(2, [(h, w)] * per)
    oracle = cpu.Labeller(h, w)
    lab = sf.Labeller(h, w)
    for c in range(n // per):
        plan.run([buf.ptr + 2 * h * w * (per * c + k) for k in range(per)])
        plan.wait()
        counts = plan.counts()
        runs = 0
        for k in range(per):
            lst = plan.output(k, int(counts[k]))
            idx = lst["index"].astype(np.int64)
            # a run starts where the index does not follow on the same row
            runs += int(np.count_nonzero(
                (np.diff(idx, prepend=-2) != 1) | (idx % w == 0)))
            assert oracle.add(300 + per * c + k, lst) == 0
        assert runs > 65536, runs
        lab.add_plan(plan, 300 + per * c)
    oracle.finish()
    lab.finish()
    expected, got = oracle.results(), lab.results()
    assert got[2] == expected[2], (got[2], expected[2])
    assert same_tables(got, expected)
    assert expected[2]["accepted"] > 1000
    plan.close()
    buf.close()


def test_plan_rerun_and_closed_right_after_add_plan(sf):
    """add_plan returns before the chunk is labelled: the plan is run again
    on other frames at once, added again, then closed before the labeller
    has finished. The tables still equal the oracle's over both chunks."""
    h, w, n = 256, 300, 8
    buf = sf.DeviceBuffer(2 * h * w * n)
    spots = fr.diffraction_spots(11, h, w, n, 1500.0 * 60)
    sf.synth_frames(buf, 2, h, w, n, 11, 0.3, True, 0.0005, spots)
    host = np.zeros((n, h, w), dtype=np.uint16)
    buf.download(host)
    oracle = cpu.Labeller(h, w)
    for z in range(n):
        assert oracle.add(z, cpu.dext(host[z])[1]) == 0
    oracle.finish()
    expected = oracle.results()
    lab = sf.Labeller(h, w)
    plan = sf.Plan(2, [(h, w)] * 4)
    for c in range(2):
        plan.run([buf.ptr + 2 * h * w * (4 * c + k) for k in range(4)])
        lab.add_plan(plan, 4 * c)
    plan.close()
    lab.finish()
    got = lab.results()
    assert got[2] == expected[2], (got[2], expected[2])
    assert same_tables(got, expected)
    assert expected[2]["accepted"] > 3
    lab.close()
    buf.close()


def test_add_lists_from_device_memory(sf):
    """Lists already in device memory (sf_hip_labeller_add_lists) over two
    calls with a gap in the frame numbers, against the oracle fed the same
    lists: blobs of 21 pixels spanning three frames."""
    h, w, n = 400, 500, 12
    rng = np.random.default_rng(21)
    frames = [[] for _ in range(n)]
    for _ in range(120):
        y, x = int(rng.integers(3, h - 3)), int(rng.integers(3, w - 3))
        first = int(rng.integers(0, n - 3))
        for step in range(3):
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    if dy * dy + dx * dx <= 5:
                        v = 50 + int(1500 * np.exp(-(dy * dy + dx * dx) / 4)
                                     * np.exp(-(step - 1) ** 2 / 4))
                        frames[first + step].append(((y + dy) * w + x + dx, v))
    lists = []
    for f in range(n):
        pts = dict(sorted(frames[f]))  # one value an index
        sig = np.zeros(len(pts), sf.SIGNAL)
        sig["index"] = list(pts.keys())
        sig["value"] = list(pts.values())
        lists.append(sig)
    numbers = [10 + f for f in range(6)] + [20 + f for f in range(6)]
    oracle = cpu.Labeller(h, w)
    for z, sig in zip(numbers, lists):
        assert oracle.add(z, sig) == 0
    oracle.finish()
    want = oracle.results()
    counts = [len(x) for x in lists]
    buf = sf.DeviceBuffer(16 * sum(counts))
    buf.upload(np.concatenate(lists))
    cbuf = sf.DeviceBuffer(4 * n)
    cbuf.upload(np.array(counts, np.uint32))
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    ptrs = [buf.ptr + 16 * int(s) for s in starts]
    cptrs = [cbuf.ptr + 4 * k for k in range(n)]
    lab = sf.Labeller(h, w)
    for a, b in ((0, 6), (6, 12)):
        lab.add_lists(numbers[a:b], ptrs[a:b], cptrs[a:b], counts[a:b],
                      counts[a:b])
    with pytest.raises(ValueError):  # not above the last frame
        lab.add_lists([5], ptrs[:1], cptrs[:1], counts[:1], counts[:1])
    lab.finish()
    got = lab.results()
    assert got[2] == want[2]
    assert same_tables(got, want)
    assert want[2]["accepted"] > 20
    lab.close()
    buf.close()
    cbuf.close()
