"""Parity at the sizes every performance figure is quoted at.

The reference's bar is byte-identical whole-run tables, device against CPU
(tests/regression.sh:230-245; tests/test_dext_gpu.cc:141-207). Here, on
frames generated by the same value models as the measured workloads:

* a full 4362 x 4148 uint32 frame pair (Poisson 50, spots up to ~1e6
  counts a pixel, masked gaps, hot pixels) through the plan and the drop-in
  find(), against the oracle and the reference's own dext;
* one plan of the real uneven chunk shapes (2162 x 2068, 3269 x 3110,
  4362 x 4148), captured and replayed, against the oracle;
* a 60-frame chunk of 4362 x 4148 uint16 frames, the measured step, then
  an 8-frame chunk, through the device labeller: every list against the
  oracle, the Spot/Pixel/Counts tables against the oracle Labeller and the
  reference's own Labeller;
* the measured steps themselves (pipelined, captured, the ordering left to
  the labeller) over three 60-frame chunks: tables against a labeller fed
  the same chunks' ordered lists;
* the four serving series at their real shapes and pixel types, together
  in one server, against the oracle.

No tolerance: lists, tables and counts compare byte for byte.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import frames as fr
from oracle import cpu

pytestmark = pytest.mark.gpu

H, W = 4362, 4148
THREADS = 16


def oracle_lists(host, capacity_div):
    """The oracle's list for every frame of a (n, h, w) stack, frames on
    host threads (the oracle releases the GIL)."""
    def one(frame):
        status, lst, found = cpu.dext(frame, capacity=frame.size //
                                      capacity_div)
        assert status == 0, (status, found)
        return lst
    with ThreadPoolExecutor(THREADS) as pool:
        return list(pool.map(one, host))


def series(sf, pixel_bytes, h, w, n, seed, background, peak, min_peak=30.0,
           spots_per_frame=1500.0):
    buf = sf.DeviceBuffer(pixel_bytes * h * w * n)
    spots = fr.diffraction_spots(seed, h, w, n,
                                 spots_per_frame * h * w / (H * W),
                                 peak=peak, min_peak=min_peak)
    sf.synth_frames(buf, pixel_bytes, h, w, n, seed, background, True,
                    0.0005, spots)
    host = np.zeros((n, h, w), dtype=np.uint16 if pixel_bytes == 2
                    else np.uint32)
    buf.download(host)
    return buf, host


def same_tables(a, b):
    return (a[0].tobytes() == b[0].tobytes() and
            a[1].tobytes() == b[1].tobytes() and a[2] == b[2])


def first_difference(got, expected):
    if len(got) != len(expected):
        return f"{len(got)} vs {len(expected)} entries"
    bad = np.nonzero(got != expected)[0]
    return f"{len(bad)} entries differ, first at {bad[0] if len(bad) else -1}"


def test_full_size_u32_frames(sf):
    """Two 4362 x 4148 uint32 frames of the 32-bit series' model: the plan
    (the sweep kernel with u64 windows and FP64 tests) and the drop-in
    find_u32 against the oracle and the reference's own dext."""
    buf, host = series(sf, 4, H, W, 2, 300, 50.0, peak=1.0e6,
                       min_peak=100.0)
    assert host.max() > 65535  # pixels beyond 16 bits
    assert (host == 0xFFFFFFFF).any() and (host == 0xFFFFFFFE).any()
    expected = oracle_lists(host, 4)
    assert all(len(e) > 10000 for e in expected)
    if cpu.have_ref():
        for f in range(2):
            ref = cpu.ref_dext(host[f], capacity=host[f].size // 4)[1]
            assert ref.tobytes() == expected[f].tobytes(), f
    plan = sf.Plan(4, [(H, W)] * 2)
    plan.run([buf.ptr, buf.ptr + 4 * H * W])
    plan.wait()
    counts = plan.counts()
    for f in range(2):
        got = plan.output(f, int(counts[f]))
        assert got.tobytes() == expected[f].tobytes(), (
            f, first_difference(got, expected[f]))
    plan.close()
    status, got = sf.find(host[1])
    assert status == 0
    assert got.tobytes() == expected[1].tobytes(), first_difference(
        got, expected[1])
    buf.close()


def test_real_size_uneven_chunk(sf):
    """One plan of the uneven chunk's real shapes (two 2162 x 2068, one
    3269 x 3110, one 4362 x 4148, H x W), with the measured series' value
    model scaled by area: run eagerly, then captured and replayed twice
    with the frames bound again between replays."""
    shapes = [(2162, 2068), (3269, 3110), (2162, 2068), (4362, 4148)]
    bufs, hosts = [], []
    for k, (h, w) in enumerate(shapes):
        b, host = series(sf, 2, h, w, 1, 200 + k, 0.3, peak=20000.0)
        bufs.append(b)
        hosts.append(host[0])
    expected = []
    with ThreadPoolExecutor(THREADS) as pool:
        for status, lst, found in pool.map(
                lambda x: cpu.dext(x, capacity=x.size // 8), hosts):
            assert status == 0
            expected.append(lst)
    assert all(len(e) > 1000 for e in expected)
    ptrs = [b.ptr for b in bufs]
    plan = sf.Plan(2, shapes)
    plan.run(ptrs)
    plan.wait()
    counts = plan.counts()
    for f in range(len(shapes)):
        got = plan.output(f, int(counts[f]))
        assert got.tobytes() == expected[f].tobytes(), (
            "eager", f, first_difference(got, expected[f]))
    plan.capture(ptrs)
    for replay in range(2):
        plan.bind(ptrs)
        plan.replay()
        plan.wait()
        counts = plan.counts()
        for f in range(len(shapes)):
            got = plan.output(f, int(counts[f]))
            assert got.tobytes() == expected[f].tobytes(), (
                "replay", replay, f, first_difference(got, expected[f]))
    plan.close()
    for b in bufs:
        b.close()


def test_full_size_chunk_through_the_device_labeller(sf):
    """A 60-frame chunk of 4362 x 4148 uint16 frames (the measured step),
    then an 8-frame chunk, so that components cross from one chunk to the
    next: every device list equals the oracle's, and the device labeller's
    tables equal the oracle Labeller's and the reference's own Labeller's
    over those lists, row for row."""
    chunks = (60, 8)
    n = sum(chunks)
    buf, host = series(sf, 2, H, W, n, 77, 0.3, peak=20000.0)
    expected = oracle_lists(host, 8)
    total = sum(len(e) for e in expected)
    assert total > 60 * 20000, total
    lab = sf.Labeller(H, W)
    first = 0
    for size in chunks:
        plan = sf.Plan(2, [(H, W)] * size)
        plan.run([buf.ptr + 2 * H * W * (first + k) for k in range(size)])
        plan.wait()
        counts = plan.counts()
        for k in range(size):
            got = plan.output(k, int(counts[k]))
            assert got.tobytes() == expected[first + k].tobytes(), (
                first + k, first_difference(got, expected[first + k]))
        lab.add_plan(plan, 1000 + first)
        lab.wait()
        plan.close()
        first += size
    lab.finish()
    got = lab.results()
    lab.close()
    buf.close()
    del host
    oracle = cpu.Labeller(H, W)
    for k in range(n):
        assert oracle.add(1000 + k, expected[k]) == 0
    oracle.finish()
    want = oracle.results()
    assert got[2] == want[2], (got[2], want[2])
    assert same_tables(got, want)
    assert want[2]["accepted"] > 10000
    if cpu.have_ref():
        ref = cpu.RefLabeller(H, W)
        for k in range(n):
            assert ref.add(1000 + k, expected[k]) == 0
        ref.finish()
        assert same_tables(got, ref.results())


def test_full_size_steps_leaving_the_ordering_to_the_labeller(sf):
    """The measured path as bench.py runs it: two plans in turn, captured
    pipelined steps that leave each chunk's ordering pending, and the
    labeller taking each chunk's survivors by the ordering's places (no
    list is written). Over three 60-frame chunks of 4362 x 4148 uint16
    frames and an 8-frame one, its tables equal, row for row, those of a
    labeller fed the same chunks' ordered lists by plain runs (the path the
    test above holds to the oracle); a chunk ordered after the hand-over
    gives the same lists as a plain run."""
    size, n_chunks = 60, 3
    n = size * n_chunks + 8
    buf = sf.DeviceBuffer(2 * H * W * n)
    spots = fr.diffraction_spots(91, H, W, n, 1500.0, peak=20000.0,
                                 min_peak=30.0)
    sf.synth_frames(buf, 2, H, W, n, 91, 0.3, True, 0.0005, spots)
    ptr = [buf.ptr + 2 * H * W * f for f in range(n)]
    chunks = [ptr[c * size:(c + 1) * size] for c in range(n_chunks)]
    # the reference: plain runs, ordered lists, add_plan
    ref = sf.Labeller(H, W)
    plain = sf.Plan(2, [(H, W)] * size)
    lists = []
    for c, frames in enumerate(chunks):
        plain.run(frames)
        plain.wait()
        counts = plain.counts()
        if c == 1:
            lists = [plain.output(k, int(counts[k])) for k in range(size)]
        ref.add_plan(plain, size * c)
        ref.wait()
    tail = sf.Plan(2, [(H, W)] * 8)
    tail.run(ptr[size * n_chunks:])
    ref.add_plan(tail, size * n_chunks)
    ref.finish()
    want = ref.results()
    ref.close()
    # the measured path
    plans = [sf.Plan(2, [(H, W)] * size) for _ in range(2)]
    for p in plans:
        p.set_deferred_order(True)
    sf.capture_pipeline(plans[0], plans[1])
    sf.capture_pipeline(plans[1], plans[0])
    lab = sf.Labeller(H, W)
    plans[0].bind(chunks[0])
    sf.pipeline(None, plans[0])
    for c in range(n_chunks):
        cur, nxt = plans[c % 2], plans[(c + 1) % 2]
        if c + 1 < n_chunks:
            nxt.bind(chunks[c + 1])
            sf.replay_pipeline(cur, nxt)
        else:
            sf.pipeline(cur, None)
        lab.add_plan(cur, size * c)
        if c == 1:
            cur.order()
            cur.wait()
            counts = cur.counts()
            for k in range(size):
                got = cur.output(k, int(counts[k]))
                assert got.tobytes() == lists[k].tobytes(), (
                    k, first_difference(got, lists[k]))
    tail.run(ptr[size * n_chunks:])
    lab.add_plan(tail, size * n_chunks)
    lab.finish()
    got = lab.results()
    lab.close()
    for p in plans + [plain, tail]:
        p.close()
    buf.close()
    assert got[2] == want[2], (got[2], want[2])
    assert same_tables(got, want)
    assert want[2]["accepted"] > 10000


def test_serving_series_at_real_sizes(sf):
    """The four serving series at their real shapes and pixel types in one
    server under a budget that makes frames wait for slots: every list and
    each series' tables against the oracle."""
    specs = [  # pixel bytes, h, w, frames, seed, background, peak
        (2, 2162, 2068, 6, 41, 0.3, 20000.0),
        (2, 3269, 3110, 4, 42, 0.3, 20000.0),
        (2, 4362, 4148, 3, 43, 0.3, 20000.0),
        (4, 4362, 4148, 2, 44, 50.0, 1.0e6),
    ]
    data = [series(sf, pb, h, w, n, seed, bg, peak)
            for pb, h, w, n, seed, bg, peak in specs]
    budget = sum(sf.server_slot_bytes(pb, h, w)
                 for pb, h, w, *_ in specs) + 4096
    server = sf.Server(budget, max_batch=4, keep_lists=True)
    ids = [server.open(pb, h, w) for pb, h, w, *_ in specs]
    rings = [[b.ptr + pb * h * w * k for k in range(n)]
             for (b, _), (pb, h, w, n, *_) in zip(data, specs)]
    server.play(ids, [750.0, 500.0, 250.0, 100.0],
                [spec[3] for spec in specs], rings, first_frame=5)
    results = []
    for sid in ids:
        server.close_series(sid)
        fm = server.frames(sid)
        lists = [server.frame_list(sid, k) for k in range(len(fm["frame"]))]
        results.append((fm, lists, server.results(sid)))
    stats = server.stats()
    server.close()
    assert stats["overflowed"] == 0
    for (pb, h, w, n, *_), (b, host), (fm, lists, tables) in zip(
            specs, data, results):
        b.close()
        assert list(fm["frame"]) == list(range(5, 5 + n))
        expected = oracle_lists(host, 16 // pb)
        oracle = cpu.Labeller(h, w)
        for k in range(n):
            assert lists[k].tobytes() == expected[k].tobytes(), (
                h, w, pb, k, first_difference(lists[k], expected[k]))
            assert oracle.add(5 + k, expected[k]) == 0
        oracle.finish()
        want = oracle.results()
        assert same_tables(tables, want), (h, w, pb, tables[2], want[2])
        assert want[2]["accepted"] > 0
