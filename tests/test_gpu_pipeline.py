"""Pipelined steps (sf_hip_plan_pipeline / capture / replay): a chunk's
scout runs beside the previous chunk's erosion and Poisson passes. The lists
each step finishes must equal the oracle's (and run()'s) byte for byte,
whatever the order of plans, shapes and rebinding; the labeller fed from
the finished plans must give the oracle's tables."""

from __future__ import annotations

import numpy as np
import pytest

import frames as fr
from oracle import cpu

pytestmark = pytest.mark.gpu


def chunk_frames(sf, shapes, seed, density=30000.0):
    """Device frames of the measured series' value model, one per shape;
    returns (buffers, pointers, host frames)."""
    bufs, ptrs, hosts = [], [], []
    for k, (h, w) in enumerate(shapes):
        b = sf.DeviceBuffer(2 * h * w)
        spots = fr.diffraction_spots(seed + k, h, w, 1,
                                     density * h * w / (4362 * 4148))
        sf.synth_frames(b, 2, h, w, 1, seed + k, 0.3, True, 0.0005, spots)
        host = np.zeros((h, w), np.uint16)
        b.download(host)
        bufs.append(b)
        ptrs.append(b.ptr)
        hosts.append(host)
    return bufs, ptrs, hosts


def chunk_frames32(sf, shapes, seed, density=30000.0):
    """Device u32 frames of the 32-bit series' value model (Poisson 50,
    peaks up to 10^6, gaps, hot pixels), one per shape."""
    bufs, ptrs, hosts = [], [], []
    for k, (h, w) in enumerate(shapes):
        b = sf.DeviceBuffer(4 * h * w)
        spots = fr.diffraction_spots(seed + k, h, w, 1,
                                     density * h * w / (4362 * 4148),
                                     peak=1.0e6, min_peak=100.0)
        sf.synth_frames(b, 4, h, w, 1, seed + k, 50.0, True, 0.0005, spots)
        host = np.zeros((h, w), np.uint32)
        b.download(host)
        bufs.append(b)
        ptrs.append(b.ptr)
        hosts.append(host)
    return bufs, ptrs, hosts


def check(plan, hosts, what):
    plan.wait()
    counts = plan.counts()
    for f, host in enumerate(hosts):
        want = cpu.dext(host, capacity=host.size // 8)[1]
        got = plan.output(f, int(counts[f]))
        assert got.tobytes() == want.tobytes(), (what, f, len(got),
                                                 len(want))


SHAPES_A = [(216, 206), (327, 311), (101, 3), (436, 414)]
SHAPES_B = [(300, 517), (64, 64), (560, 1040)]


def test_eager_steps_over_chunks_of_other_shapes(sf):
    chunks = [chunk_frames(sf, SHAPES_A if c % 2 == 0 else SHAPES_B, 10 * c)
              for c in range(4)]
    plans = [sf.Plan(2, SHAPES_A), sf.Plan(2, SHAPES_B)]
    plans[0].bind(chunks[0][1])
    sf.pipeline(None, plans[0])
    for c in range(1, 4):
        cur, nxt = plans[(c - 1) % 2], plans[c % 2]
        nxt.bind(chunks[c][1])
        sf.pipeline(cur, nxt)
        check(cur, chunks[c - 1][2], ("chunk", c - 1))
    sf.pipeline(plans[1], None)
    check(plans[1], chunks[3][2], ("chunk", 3))
    for p in plans:
        p.close()
    for bufs, _, _ in chunks:
        for b in bufs:
            b.close()


def test_u32_steps_over_chunks_of_other_shapes(sf):
    """The 32-bit form's pipelined steps (its scout beside its erosion and
    Poisson passes) over chunks of two shape lists in turn, eager and
    replayed: every list equals the oracle's."""
    chunks = [chunk_frames32(sf, SHAPES_A if c % 2 == 0 else SHAPES_B,
                             500 + 10 * c) for c in range(4)]
    plans = [sf.Plan(4, SHAPES_A), sf.Plan(4, SHAPES_B)]
    plans[0].bind(chunks[0][1])
    sf.pipeline(None, plans[0])
    for c in range(1, 4):
        cur, nxt = plans[(c - 1) % 2], plans[c % 2]
        nxt.bind(chunks[c][1])
        sf.pipeline(cur, nxt)
        check(cur, chunks[c - 1][2], ("u32 chunk", c - 1))
    sf.pipeline(plans[1], None)
    check(plans[1], chunks[3][2], ("u32 chunk", 3))
    for p in plans:
        p.close()
    for bufs, _, _ in chunks:
        for b in bufs:
            b.close()


def test_u32_full_size_steps_equal_run(sf):
    """Two 4-frame chunks of 4362 x 4148 u32 frames through pipelined steps
    (graphs replayed): every list equals the plan's own run() on the same
    frames, byte for byte."""
    h, w, n = 4362, 4148, 8
    buf = sf.DeviceBuffer(4 * h * w * n)
    spots = fr.diffraction_spots(7, h, w, n, 1500.0, peak=1.0e6,
                                 min_peak=100.0)
    sf.synth_frames(buf, 4, h, w, n, 7, 50.0, True, 0.0005, spots)
    ptr = [buf.ptr + 4 * h * w * f for f in range(n)]
    ref = sf.Plan(4, [(h, w)] * 4)
    want = []
    for c in range(2):
        ref.run(ptr[4 * c:4 * c + 4])
        ref.wait()
        counts = ref.counts()
        want.append([ref.output(k, int(counts[k])) for k in range(4)])
    ref.close()
    p = [sf.Plan(4, [(h, w)] * 4), sf.Plan(4, [(h, w)] * 4)]
    p[0].bind(ptr[:4])
    sf.pipeline(None, p[0])
    p[1].bind(ptr[4:])
    sf.capture_pipeline(p[0], p[1])
    sf.replay_pipeline(p[0], p[1])
    sf.pipeline(p[1], None)
    for c in range(2):
        p[c].wait()
        counts = p[c].counts()
        for k in range(4):
            got = p[c].output(k, int(counts[k]))
            assert got.tobytes() == want[c][k].tobytes(), (c, k)
        assert len(want[c][0]) > 20000
    for q in p:
        q.close()
    buf.close()


def test_replayed_steps_and_the_labeller(sf):
    """Two captured steps (P0 -> P1 and P1 -> P0) replayed over six chunks
    of one shape list, frames rebound before each replay, eager work between
    replays; the finished plans feed one device labeller, whose tables equal
    the oracle Labeller's over the oracle's lists."""
    shapes = [(256, 384)] * 3
    chunks = [chunk_frames(sf, shapes, 100 + 10 * c, density=60000.0)
              for c in range(6)]
    p = [sf.Plan(2, shapes), sf.Plan(2, shapes)]
    p[0].bind(chunks[0][1])
    sf.pipeline(None, p[0])
    p[1].bind(chunks[1][1])
    sf.capture_pipeline(p[0], p[1])
    lab = sf.Labeller(256, 384)
    oracle = cpu.Labeller(256, 384)
    for c in range(6):
        cur, nxt = p[c % 2], p[(c + 1) % 2]
        if c + 1 < 6:
            if c == 1:  # the second pair's graph, captured on first use
                nxt.bind(chunks[c + 1][1])
                sf.capture_pipeline(cur, nxt)
            nxt.bind(chunks[c + 1][1])
            sf.replay_pipeline(cur, nxt)
        else:
            sf.pipeline(cur, None)
        check(cur, chunks[c][2], ("chunk", c))
        lab.add_plan(cur, 3 * c)
        for k, host in enumerate(chunks[c][2]):
            assert oracle.add(3 * c + k, cpu.dext(host)[1]) == 0
        assert sf.find(chunks[c][2][0])[0] == 0  # eager work in between
    lab.finish()
    oracle.finish()
    got, want = lab.results(), oracle.results()
    assert got[2] == want[2]
    assert got[0].tobytes() == want[0].tobytes()
    assert got[1].tobytes() == want[1].tobytes()
    assert want[2]["accepted"] > 10
    lab.close()
    for q in p:
        q.close()
    for bufs, _, _ in chunks:
        for b in bufs:
            b.close()


def test_full_size_steps_equal_run(sf):
    """Two 8-frame chunks of 4362 x 4148 frames of the measured series
    through pipelined steps: every list equals the plan's own run() on the
    same frames, byte for byte."""
    h, w, n = 4362, 4148, 16
    buf = sf.DeviceBuffer(2 * h * w * n)
    spots = fr.diffraction_spots(5, h, w, n, 1500.0)
    sf.synth_frames(buf, 2, h, w, n, 5, 0.3, True, 0.0005, spots)
    ptr = [buf.ptr + 2 * h * w * f for f in range(n)]
    ref = sf.Plan(2, [(h, w)] * 8)
    want = []
    for c in range(2):
        ref.run(ptr[8 * c:8 * c + 8])
        ref.wait()
        counts = ref.counts()
        want.append([ref.output(k, int(counts[k])) for k in range(8)])
    ref.close()
    p = [sf.Plan(2, [(h, w)] * 8), sf.Plan(2, [(h, w)] * 8)]
    p[0].bind(ptr[:8])
    sf.pipeline(None, p[0])
    p[1].bind(ptr[8:])
    sf.pipeline(p[0], p[1])
    sf.pipeline(p[1], None)
    for c in range(2):
        p[c].wait()
        counts = p[c].counts()
        for k in range(8):
            got = p[c].output(k, int(counts[k]))
            assert got.tobytes() == want[c][k].tobytes(), (c, k)
        assert len(want[c][0]) > 20000
    for q in p:
        q.close()
    buf.close()


def test_refusals(sf):
    a, b = sf.Plan(2, [(64, 64)]), sf.Plan(2, [(64, 64)])
    c32 = sf.Plan(4, [(64, 64)])
    buf = sf.DeviceBuffer(4 * 64 * 64)
    with pytest.raises(ValueError):
        sf.pipeline(None, None)
    with pytest.raises(ValueError):
        sf.pipeline(a, b)  # a was not scouted
    with pytest.raises(ValueError):
        sf.pipeline(None, a)  # a has no frames bound
    a.bind([buf.ptr])
    with pytest.raises(ValueError):
        sf.pipeline(a, a)
    c32.bind([buf.ptr])
    with pytest.raises(ValueError):
        sf.pipeline(b, c32)  # pixels of two sizes
    sf.pipeline(None, a)
    with pytest.raises(ValueError):
        sf.replay_pipeline(a, None)  # nothing captured
    a.bind([buf.ptr])  # rebinding drops the scout
    with pytest.raises(ValueError):
        sf.pipeline(a, None)
    for q in (a, b, c32):
        q.close()
    buf.close()


def test_labeller_takes_a_pending_ordering(sf):
    """Two plans in turn with deferred ordering, each chunk handed to the
    labeller with its ordering still pending: the labeller enqueues the
    ordering's places itself and takes the survivors from staging by them,
    so its tables equal the oracle Labeller's although no list is written;
    a plan ordered after the hand-over (only the copy is left) still gives
    the oracle's lists, at 16 and at 32 bits."""
    for pixel_bytes, (h, w) in ((2, (256, 384)), (4, (300, 517))):
        shapes = [(h, w)] * 3
        n = 6
        if pixel_bytes == 2:
            chunks = [chunk_frames(sf, shapes, 500 + 10 * c, density=60000.0)
                      for c in range(n)]
        else:
            chunks = [chunk_frames32(sf, shapes, 500 + 10 * c)
                      for c in range(n)]
        p = [sf.Plan(pixel_bytes, shapes) for _ in range(2)]
        for q in p:
            q.set_deferred_order(True)
        p[0].bind(chunks[0][1])
        sf.pipeline(None, p[0])
        lab = sf.Labeller(h, w)
        oracle = cpu.Labeller(h, w)
        for c in range(n):
            cur, nxt = p[c % 2], p[(c + 1) % 2]
            if c + 1 < n:
                nxt.bind(chunks[c + 1][1])
                sf.pipeline(cur, nxt)
            else:
                sf.pipeline(cur, None)
            lab.add_plan(cur, 3 * c)
            for k, host in enumerate(chunks[c][2]):
                assert oracle.add(3 * c + k, cpu.dext(host)[1]) == 0
            if c % 3 == 1:
                cur.order()
                check(cur, chunks[c][2], ("ordered after", pixel_bytes, c))
        lab.finish()
        oracle.finish()
        got, want = lab.results(), oracle.results()
        assert got[2] == want[2], (pixel_bytes, got[2], want[2])
        assert got[0].tobytes() == want[0].tobytes()
        assert got[1].tobytes() == want[1].tobytes()
        assert want[2]["accepted"] > 5
        lab.close()
        for q in p:
            q.close()
        for bufs, _, _ in chunks:
            for b in bufs:
                b.close()


def test_deferred_ordering_three_plans_and_the_labeller(sf):
    """Three plans in turn with deferred ordering: each captured step
    (P[c] -> P[c+1]) leaves P[c]'s ordering to order(), which runs while the
    next step does; every list equals the oracle's, and the labeller fed
    from the ordered plans gives the oracle Labeller's tables. At 32 bits
    the same rotation, eager, gives the oracle's lists."""
    shapes = [(256, 384)] * 3
    n = 7
    chunks = [chunk_frames(sf, shapes, 300 + 10 * c, density=60000.0)
              for c in range(n)]
    p = [sf.Plan(2, shapes) for _ in range(3)]
    for q in p:
        q.set_deferred_order(True)
    p[0].bind(chunks[0][1])
    sf.pipeline(None, p[0])
    lab = sf.Labeller(256, 384)
    oracle = cpu.Labeller(256, 384)
    for c in range(n):
        cur, nxt = p[c % 3], p[(c + 1) % 3]
        if c + 1 < n:
            nxt.bind(chunks[c + 1][1])
            if c < 3:  # each pair's graph, captured on first use
                sf.capture_pipeline(cur, nxt)
            sf.replay_pipeline(cur, nxt)
        else:
            sf.pipeline(cur, None)
        cur.order()
        lab.add_plan(cur, 3 * c)
        for k, host in enumerate(chunks[c][2]):
            assert oracle.add(3 * c + k, cpu.dext(host)[1]) == 0
        if c % 2:
            check(cur, chunks[c][2], ("chunk", c))
    lab.finish()
    oracle.finish()
    got, want = lab.results(), oracle.results()
    assert got[2] == want[2]
    assert got[0].tobytes() == want[0].tobytes()
    assert got[1].tobytes() == want[1].tobytes()
    assert want[2]["accepted"] > 10
    lab.close()
    for q in p:
        q.close()
    for bufs, _, _ in chunks:
        for b in bufs:
            b.close()
    shapes32 = [(216, 206), (101, 3), (300, 517)]
    chunks32 = [chunk_frames32(sf, shapes32, 700 + 10 * c) for c in range(4)]
    p32 = [sf.Plan(4, shapes32) for _ in range(3)]
    for q in p32:
        q.set_deferred_order(True)
    p32[0].bind(chunks32[0][1])
    sf.pipeline(None, p32[0])
    for c in range(4):
        cur, nxt = p32[c % 3], p32[(c + 1) % 3]
        if c + 1 < 4:
            nxt.bind(chunks32[c + 1][1])
            sf.pipeline(cur, nxt)
        else:
            sf.pipeline(cur, None)
        cur.order()
        check(cur, chunks32[c][2], ("u32 chunk", c))
    for q in p32:
        q.close()
    for bufs, _, _ in chunks32:
        for b in bufs:
            b.close()


def test_deferred_ordering_refusals(sf):
    a, b = sf.Plan(2, [(64, 64)]), sf.Plan(2, [(64, 64)])
    buf = sf.DeviceBuffer(2 * 64 * 64)
    with pytest.raises(ValueError):
        a.order()  # nothing pending
    a.set_deferred_order(True)
    a.bind([buf.ptr])
    sf.pipeline(None, a)
    b.bind([buf.ptr])
    sf.capture_pipeline(a, b)
    sf.pipeline(a, b)  # a's ordering now pending
    with pytest.raises(ValueError):
        a.set_deferred_order(False)  # a setting change while one is pending
    with pytest.raises(ValueError):
        sf.pipeline(b, a)  # a cannot be scouted before it is ordered
    a.order()
    with pytest.raises(ValueError):
        a.order()  # done already
    a.set_deferred_order(False)
    a.bind([buf.ptr])
    sf.pipeline(None, a)
    b.bind([buf.ptr])
    with pytest.raises(ValueError):
        sf.replay_pipeline(a, b)  # captured with the other setting
    sf.pipeline(a, b)
    a.wait()
    assert a.counts()[0] == a.counts()[0]  # ordered by the step itself
    for q in (a, b):
        q.close()
    buf.close()
