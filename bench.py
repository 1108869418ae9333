"""Benchmark: the spot-finder hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chunk C]

Workload: a series of 1800 device-resident 4148 x 4362 (W x H) uint16
frames, Poisson background 0.3, ~1500 Gaussian spots a frame (sigma 0.8-2
px, 1-5 frames, peaks up to 20,000), masked module gaps at the 514 x 1030
pitch and 0.05% hot pixels -- generated on the device from a seed. A step is
one chunk of C frames through the hot path: the threshold with ordered
compaction (one graph replay), and the device 3-D labelling and centroid
pass for the chunk; two plans alternate, so a chunk's labelling overlaps the
next chunk's threshold. `value` is frames/s over the timed steps (all of
them thresholded and labelled) with the inputs already in HBM.

Also reported: the roofline of the dominant kernel, the sparse threshold's
scout (it reads every pixel of the chunk once: algorithmic bytes per launch
= chunk x frame bytes, over its HIP-event time on the plan's stream, against
8 TB/s), the same for the whole threshold sequence (frames read plus lists
written, over the scout, erosion and Poisson passes together), the plain
one-frame-per-launch form on the same frames, the reference's own CPU path
(oracle/_ref, compiled from the reference sources) timed on this host's
cores on a bounded sample, whose lists are also compared with the device's
(mask_diffs); then chunks of uneven frame sizes, a 32-bit series, and the
serving loop on four concurrent series in real time.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "dials-metal-spotfinder_amd"),
           os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HEIGHT, WIDTH = 4362, 4148
PEAK_HBM = 8.0e12
# the two launches of a pipelined step at 16 bits: half the next chunk's
# scout mixed with this chunk's erosion, then the other half mixed with its
# Poisson pass
MIXED_A = "scout_erode_kernel"
MIXED_B = "scout_poisson_kernel"
METRIC = ("frames/s + HBM GB/s (% peak), 4148×4362 uint16, 0 mask diffs; "
          "p99 latency")


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def host_description() -> dict:
    """The host the CPU baseline runs on: usable cores, CPU model, and the
    reference build's compiler and flags (oracle/Makefile)."""
    import platform
    import subprocess

    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    try:
        compiler = subprocess.run(["g++", "--version"], capture_output=True,
                                  text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        compiler = "g++"
    return {"nproc": nproc, "cpu_model": model,
            "flags": f"{compiler}: -O3 -DNDEBUG -ffp-contract=off (no -march)"}


def median_rate(fn, repeats, units):
    """Median of `repeats` timings of fn(), as units per second."""
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return units / float(np.median(times)), float(np.sum(times))


def cpu_baseline(sf, host_frames, device_lists, plan, chunks=None,
                 refl=None, repeats=5):
    """The reference's CPU path (oracle/_ref, compiled from its sources; the
    oracle's port where it was not built) on bounded samples of the same
    workloads, following the baseline plan:

    * the threshold alone at -j 1 and at -j nproc (one frame per thread,
      thread-local scratch), median of `repeats`, over the sample frames;
    * the grouping alone: the reference Labeller over the device lists of
      the measured chunk, on one thread; its Spot/Pixel/Counts tables are
      compared byte for byte with a fresh device labeller's over the same
      plan lists (table_diffs);
    * the whole path of the 12-frame 256 x 384 planted series at -j 4
      (threshold workers, grouping in frame order, the .refl table), beside
      the reference's published 516.4 images/s (README.md:95-107; that
      figure also reads the HDF5 file);
    * every sample frame's list against the device's (mask_diffs).

    `value` is the path's rate on the measured workload: the slower of the
    threshold at -j nproc and the grouping thread, as in the reference's
    pipeline."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    from oracle import cpu
    import frames as fr

    kind = "reference" if cpu.have_ref() else "port"
    host = host_description()
    nproc = host["nproc"]
    n = len(host_frames)
    if kind == "reference":
        def run(jobs, threads):
            return cpu.ref_dext_cycle(host_frames, jobs, threads)
    else:
        def run(jobs, threads):
            return cpu.oracle_dext_many(host_frames[:min(jobs, n)], threads)
    one_fps, spent1 = median_rate(lambda: run(2, 1), repeats, 2)
    jobs = max(nproc, n) if kind == "reference" else n
    many_fps, spent = median_rate(lambda: run(jobs, nproc), repeats, jobs)
    label_fps, table_diffs, tables = None, None, None
    if kind == "reference":
        labeller = cpu.RefLabeller(HEIGHT, WIDTH)
        t0 = time.perf_counter()
        for f, signal in enumerate(device_lists):
            labeller.add(f, signal)
        labeller.finish()
        label_fps = len(device_lists) / (time.perf_counter() - t0)
        want = labeller.results()
        dev = sf.Labeller(HEIGHT, WIDTH)
        dev.add_plan(plan, 0)
        dev.finish()
        got = dev.results()
        dev.close()
        table_diffs = int(got[0].tobytes() != want[0].tobytes()) + \
            int(got[1].tobytes() != want[1].tobytes()) + int(got[2] != want[2])
        tables = {"frames": len(device_lists), "spots": len(want[0]),
                  "pixels": len(want[1]), "accepted": want[2]["accepted"]}
    # parity on the sample: every SignalPixel field of every frame
    check = cpu.ref_dext if kind == "reference" else cpu.dext
    with ThreadPoolExecutor(min(nproc, 16)) as pool:
        expected = list(pool.map(
            lambda fr_: check(fr_, capacity=fr_.size // 8)[1], host_frames))
    diffs = sum(int(device_lists[f].tobytes() != expected[f].tobytes())
                for f in range(n))
    # the whole path on the 12-frame planted series, -j 4
    pipeline = None
    if kind == "reference":
        stack, _ = fr.planted_series()
        path = os.path.join(tempfile.mkdtemp(), "planted.refl")
        status, n_sig, n_spots, _ = cpu.ref_pipeline(stack, 4, path)
        ips, _ = median_rate(lambda: cpu.ref_pipeline(stack, 4, path), 25,
                             len(stack))
        fixture = fr.nxmx_fixture()
        kat = cpu.ref_pipeline(fixture, 4, path)
        pipeline = {
            "workload": "12 x 256x384 u16, Poisson 1, 48 planted "
                        "reflections; threshold on 4 threads, grouping, "
                        ".refl with shoeboxes (frames in memory)",
            "images_per_s": round(ips, 1),
            "published_images_per_s": 516.4,
            "signal_pixels": n_sig, "spots": n_spots, "status": status,
            "fixture_kat": {"signal_pixels": kat[1], "spots": kat[2],
                            "published": [8928, 48]},
        }
    value = many_fps if label_fps is None else min(many_fps, label_fps)
    # the oracle's decompress::image on up to two of the sample's
    # bitshuffle+LZ4 chunks, one thread (a plain restatement: bitshuffle's
    # SIMD library is not here), checked against the frames
    decode_ms, decode_diffs = None, 0
    sample = chunks[:min(2, n)] if chunks else []
    if sample:
        t0 = time.perf_counter()
        for f, chunk in enumerate(sample):
            out, message = cpu.decompress_image(chunk, "bslz4", 16, HEIGHT,
                                                WIDTH)
            if message or out != host_frames[f].tobytes():
                decode_diffs += 1
        decode_ms = (time.perf_counter() - t0) / len(sample) * 1e3
    # the reference's own refl::write on the driver's tables: the same bytes
    # as the library's writer, and its time beside it
    ref_write_ms, refl_same = None, None
    if refl is not None and kind == "reference":
        other = refl["path"] + ".reference"
        t0 = time.perf_counter()
        cpu.ref_refl_write(other, refl["spots"], refl["pixels"], WIDTH)
        ref_write_ms = (time.perf_counter() - t0) * 1e3
        with open(other, "rb") as a, open(refl["path"], "rb") as b:
            refl_same = a.read() == b.read()
    px = HEIGHT * WIDTH
    return {
        "value": round(value, 3),
        "unit": "frames/s",
        "cores": nproc,
        "kind": kind,
        "sample": f"{n} frames of the measured series: dext at -j 1 (2 "
                  f"frames a repeat) and at -j {nproc} ({jobs} frames a "
                  f"repeat, one per thread), median of {repeats}; the "
                  f"reference Labeller over the {len(device_lists)}-frame "
                  f"chunk's lists on one thread",
        "host": host,
        "threshold_j1_fps": round(one_fps, 3),
        "threshold_j1_gpx_per_s": round(one_fps * px / 1e9, 4),
        "threshold_jn_fps": round(many_fps, 3),
        "threshold_jn_gpx_per_s": round(many_fps * px / 1e9, 4),
        "threshold_seconds": round(spent1 + spent, 1),
        "labeller_fps": None if label_fps is None else round(label_fps, 1),
        "labeller_signal_px_per_s": None if label_fps is None else round(
            label_fps * sum(len(x) for x in device_lists) /
            len(device_lists), 1),
        "table_diffs": table_diffs,
        "tables": tables,
        "parity_frames": n,
        "whole_path_planted_j4": pipeline,
        "decode_port_ms_per_frame": None if decode_ms is None
        else round(decode_ms, 1),
        "decode_port_diffs": decode_diffs,
        "refl_reference_write_ms": None if ref_write_ms is None
        else round(ref_write_ms, 1),
        "refl_identical": refl_same,
    }, diffs


def cpu_u32_threshold(host_frames, threads, repeats=5):
    """The reference's dext on the 32-bit series' frames at -j nproc (one
    frame per thread), median of `repeats`: frames/s."""
    from oracle import cpu

    if not cpu.have_ref():
        return None
    jobs = max(threads, len(host_frames))
    fps, spent = median_rate(
        lambda: cpu.ref_dext_cycle(host_frames, jobs, threads), repeats, jobs)
    return {"fps": round(fps, 3), "cores": threads,
            "sample": f"{len(host_frames)} frames of 4148x4362 u32, {jobs} "
                      f"a repeat, median of {repeats} ({spent:.1f} s)"}


BATCH_SHAPES = [(2162, 2068)] * 32 + [(3269, 3110)] * 16 + [(4362, 4148)] * 8


def reserve_for(labeller, chunks_done, chunks_ahead):
    """Sizes the labeller's result tables for `chunks_ahead` more chunks
    like the `chunks_done` it has labelled (with a quarter to spare), so that
    timed steps allocate nothing: a large device allocation can stall a
    series for hundreds of milliseconds or more."""
    spots, pixels = labeller.sizes()
    per = 4 * max(chunks_done, 1)
    labeller.reserve(5 * spots * chunks_ahead // per + 1024,
                     5 * pixels * chunks_ahead // per + 1024)


def chunk_step_timing(sf, plan, chunks, steps, labeller=None, frames_per=0):
    """Graph replays of `plan` over `chunks` (lists of device pointers) in
    turn, with the labeller fed from each replay when given; returns
    (seconds per step, median threshold kernel ms, median ordering ms)."""
    plan.capture(chunks[0])
    first = 0

    def step(s):
        nonlocal first
        plan.bind(chunks[s % len(chunks)])
        plan.replay()
        if labeller is not None:
            labeller.add_plan(plan, first)
            first += frames_per

    step(0)
    plan.wait()
    if labeller is not None:
        reserve_for(labeller, 1, steps)
    sf.lib().sf_hip_synchronize()
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        step(s)
    plan.wait()
    if labeller is not None:
        labeller.wait()
    sf.lib().sf_hip_synchronize()
    per_step = (time.perf_counter() - t0) / steps
    kernel, order = [], []
    plan.set_timing(True)
    for s in range(3):
        plan.run(chunks[s % len(chunks)])
        plan.wait()
        a, b = plan.last_times()
        kernel.append(a)
        order.append(b)
    plan.set_timing(False)
    return per_step, float(np.median(kernel)), float(np.median(order))


def pipelined_step_timing(sf, plans, chunks, steps, labeller=None,
                          frames_per=0):
    """Pipelined steps over `chunks` (lists of device pointers) with two
    plans in turn: a step finishes one plan's chunk (erosion, Poisson,
    ordering) while it scouts the next chunk on the other, from two captured
    graphs, and hands the finished chunk to the labeller when given.
    Returns (seconds per step, median ms of an eager step's three launches:
    scout half + erosion, scout half + Poisson, ordering)."""
    n = len(chunks)
    first = [0]

    def bind(p, s):
        p.bind(chunks[s % n])

    def step(s):
        cur, nxt = plans[s % 2], plans[(s + 1) % 2]
        bind(nxt, s + 1)
        sf.replay_pipeline(cur, nxt)
        if labeller is not None:
            labeller.add_plan(cur, first[0])
        first[0] += frames_per

    def drain(last):
        sf.pipeline(plans[last % 2], None)
        if labeller is not None:
            labeller.add_plan(plans[last % 2], first[0])
        first[0] += frames_per
        for p in plans:
            p.wait()
        if labeller is not None:
            labeller.wait()

    sf.capture_pipeline(plans[0], plans[1])
    sf.capture_pipeline(plans[1], plans[0])
    bind(plans[0], 0)
    sf.pipeline(None, plans[0])
    step(0)
    drain(1)
    if labeller is not None:
        reserve_for(labeller, 2, steps + 1)
    sf.lib().sf_hip_synchronize()
    bind(plans[0], 0)
    sf.pipeline(None, plans[0])
    sf.lib().sf_hip_synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        step(s)
    if labeller is not None:
        labeller.wait()
    sf.lib().sf_hip_synchronize()
    per_step = (time.perf_counter() - t0) / steps
    drain(steps)
    times = []
    bind(plans[0], 0)
    sf.pipeline(None, plans[0])
    for s in range(3):
        cur = plans[s % 2]
        cur.set_timing(True)
        bind(plans[(s + 1) % 2], s + 1)
        sf.pipeline(cur, plans[(s + 1) % 2])
        cur.wait()
        times.append(cur.last_pass_times())
        cur.set_timing(False)
    sf.pipeline(plans[1], None)
    plans[1].wait()
    a, b, o = (float(np.median([t[k] for t in times])) for k in range(3))
    return per_step, a, b, o


SWEEP_DENSITIES = (1000, 6000, 30000, 130000)  # signal pixels a frame
SWEEP_SPANS = (1, 3, 10)  # frames a reflection spans
SWEEP_FRAMES = 40


def sweep_lists(signal_dtype, density, span):
    """The labeller sweep's frames (tests/bench_dials_spots.cc:62-111):
    blobs of the 21 pixels within sqrt(5) of a centre, each spanning `span`
    consecutive frames, scattered over a 4362 x 4148 detector and valued
    50 + int(1500 x rocking x radial); each frame ascending by index without
    duplicates. Generated here with numpy, seeded."""
    rng = np.random.default_rng(7)
    blobs = density * SWEEP_FRAMES // (19 * span)
    y = rng.integers(8, HEIGHT - 8, blobs)
    x = rng.integers(8, WIDTH - 8, blobs)
    first = np.arange(blobs) % (SWEEP_FRAMES - span)
    offs = [(dy, dx) for dy in range(-2, 3) for dx in range(-2, 3)
            if dy * dy + dx * dx <= 5]
    per = [[] for _ in range(SWEEP_FRAMES)]
    for step in range(span):
        middle = step - (span - 1) / 2.0
        rocking = np.exp(-middle * middle / 4.0)
        for dy, dx in offs:
            radial = np.exp(-(dy * dy + dx * dx) / 4.0)
            value = 50 + int(1500.0 * rocking * radial)
            index = (y + dy) * WIDTH + x + dx
            frame = first + step
            for f in range(SWEEP_FRAMES):
                sel = frame == f
                if sel.any():
                    per[f].append(index[sel])
                    per[f].append(np.full(int(sel.sum()), value, np.int64))
    lists = []
    for f in range(SWEEP_FRAMES):
        idx = np.concatenate(per[f][0::2]) if per[f] else np.zeros(0, np.int64)
        val = np.concatenate(per[f][1::2]) if per[f] else np.zeros(0, np.int64)
        order = np.argsort(idx, kind="stable")
        idx, val = idx[order], val[order]
        keep = np.ones(len(idx), bool)
        keep[1:] = idx[1:] != idx[:-1]
        sig = np.zeros(int(keep.sum()), signal_dtype)
        sig["index"] = idx[keep]
        sig["value"] = val[keep]
        lists.append(sig)
    return lists


def labeller_sweep(sf, repeats):
    """Labelling cost against density and span, as the reference's grouping
    benchmark sweeps it (tests/bench_dials_spots.cc:150-160): 40 frames of
    1k/6k/30k/130k signal pixels with reflections spanning 1/3/10 frames.
    The device labeller takes the lists from device memory (add_lists, one
    chunk) and is timed from the call to finish()'s return, best of
    `repeats`; the reference's Labeller (compiled from its sources) takes
    the host lists on one core, best of `repeats`; their tables are
    compared byte for byte."""
    from oracle import cpu

    rows = []
    for density in SWEEP_DENSITIES:
        for span in SWEEP_SPANS:
            lists = sweep_lists(sf.SIGNAL, density, span)
            counts = [len(x) for x in lists]
            total = sum(counts)
            buf = sf.DeviceBuffer(16 * max(total, 1))
            cbuf = sf.DeviceBuffer(4 * SWEEP_FRAMES)
            offsets = np.concatenate([[0], np.cumsum(counts)[:-1]])
            if total:
                buf.upload(np.concatenate(lists))
            cbuf.upload(np.array(counts, np.uint32))
            ptrs = [buf.ptr + 16 * int(o) for o in offsets]
            cptrs = [cbuf.ptr + 4 * k for k in range(SWEEP_FRAMES)]
            best, got = None, None
            for _ in range(repeats):
                lab = sf.Labeller(HEIGHT, WIDTH)
                sf.lib().sf_hip_synchronize()
                t0 = time.perf_counter()
                lab.add_lists(list(range(SWEEP_FRAMES)), ptrs, cptrs, counts,
                              counts)
                lab.finish()
                spent = time.perf_counter() - t0
                best = spent if best is None else min(best, spent)
                got = lab.results()
                lab.close()
            ref_best, want = None, None
            if cpu.have_ref():
                for _ in range(repeats):
                    ref = cpu.RefLabeller(HEIGHT, WIDTH)
                    t0 = time.perf_counter()
                    for f, x in enumerate(lists):
                        ref.add(f, x)
                    ref.finish()
                    spent = time.perf_counter() - t0
                    ref_best = spent if ref_best is None else min(ref_best,
                                                                  spent)
                    want = ref.results()
            buf.close()
            cbuf.close()
            per_frame = total / SWEEP_FRAMES
            row = {
                "pixels_per_frame": round(per_frame),
                "span": span,
                "device_ms_per_frame": round(1e3 * best / SWEEP_FRAMES, 4),
                "device_mpixel_per_s": round(total / best / 1e6, 1),
                "spots": len(got[0]),
            }
            if want is not None:
                row["reference_ms_per_frame"] = round(
                    1e3 * ref_best / SWEEP_FRAMES, 4)
                row["reference_mpixel_per_s"] = round(total / ref_best / 1e6,
                                                      1)
                row["table_diffs"] = (
                    int(got[0].tobytes() != want[0].tobytes()) +
                    int(got[1].tobytes() != want[1].tobytes()) +
                    int(got[2] != want[2]))
            rows.append(row)
    return {
        "workload": f"{SWEEP_FRAMES} frames of 4362x4148, 21-pixel blobs "
                    "spanning 1/3/10 frames at 1k/6k/30k/130k signal "
                    "pixels a frame (the reference's grouping benchmark)",
        "device": "lists in device memory, one chunk, add_lists to finish",
        "reference": "the reference's Labeller on one core, host lists",
        "repeats": repeats,
        "rows": rows,
    }


def batched_workload(sf, fr, steps):
    """Chunks of uneven frame sizes in one launch: 32 frames of 2162 x 2068,
    16 of 3269 x 3110 and 8 of 4362 x 4148 (H x W) u16, the 1800-frame
    series' value model scaled by area, two distinct chunks alternated."""
    bufs, chunks = [], [[], []]
    for k, (h, w) in enumerate(sorted(set(BATCH_SHAPES))):
        count = BATCH_SHAPES.count((h, w))
        buf = sf.DeviceBuffer(2 * h * w * count * 2)
        spots = fr.diffraction_spots(200 + k, h, w, 2 * count,
                                     1500.0 * h * w / (HEIGHT * WIDTH))
        sf.synth_frames(buf, 2, h, w, 2 * count, 200 + k, 0.3, True, 0.0005,
                        spots)
        bufs.append(buf)
        for c in range(2):
            chunks[c] += [buf.ptr + 2 * h * w * (c * count + f)
                          for f in range(count)]
    shapes = sorted(BATCH_SHAPES)
    plan = sf.Plan(2, shapes)
    tuned = plan.autotune(chunks[0], 3)
    per_step, kernel_ms, order_ms = chunk_step_timing(sf, plan, chunks, steps)
    plan.run(chunks[0])
    plan.wait()
    signal = float(np.sum(plan.counts()))
    pixel_bytes = float(sum(2 * h * w for h, w in shapes))
    algorithmic = pixel_bytes + 16.0 * signal
    plan.close()
    for buf in bufs:
        buf.close()
    return {
        "workload": "56-frame chunks: 32 x 2162x2068 + 16 x 3269x3110 + "
                    "8 x 4362x4148 uint16 (H x W), one graph replay a chunk",
        "frames_per_s": round(len(shapes) / per_step, 1),
        "chunks_per_s": round(1.0 / per_step, 1),
        "gb_per_s": round(algorithmic / per_step / 1e9, 1),
        "kernel_ms_per_chunk": round(kernel_ms, 4),
        "order_ms_per_chunk": round(order_ms, 4),
        "kernel_gb_per_s": round(algorithmic / (kernel_ms / 1e3) / 1e9, 1),
        "kernel_hbm_frac": round(algorithmic / (kernel_ms / 1e3) / PEAK_HBM,
                                 4),
        "algorithmic_bytes_per_chunk": int(algorithmic),
        "autotuned": {"band_rows": tuned[0], "blocks": tuned[1]},
    }


def u32_workload(sf, fr, steps, chunk=30, series=120, check=8):
    """The 32-bit path: a device-resident series of 4148 x 4362 uint32
    frames (Poisson 50, spots up to ~10^7 counts, masked gaps at
    0xFFFFFFFF, hot pixels), thresholded in chunks by the sparse passes in
    pipelined steps (a chunk's scout beside the previous chunk's erosion and
    Poisson pass), then labelled on the device. Returns the first `check`
    frames (host copies, compared with the reference's lists) and the
    record."""
    buf = sf.DeviceBuffer(4 * HEIGHT * WIDTH * series)
    spots = fr.diffraction_spots(300, HEIGHT, WIDTH, series, 1500.0,
                                 peak=1.0e6, min_peak=100.0)
    sf.synth_frames(buf, 4, HEIGHT, WIDTH, series, 300, 50.0, True, 0.0005,
                    spots)
    ptr = [buf.ptr + 4 * HEIGHT * WIDTH * f for f in range(series)]
    chunks = [ptr[c * chunk:(c + 1) * chunk] for c in range(series // chunk)]
    plan = sf.Plan(4, [(HEIGHT, WIDTH)] * chunk)
    other = sf.Plan(4, [(HEIGHT, WIDTH)] * chunk)
    labeller = sf.Labeller(HEIGHT, WIDTH)
    per_step, mixed_a, mixed_b, order_ms = pipelined_step_timing(
        sf, [plan, other], chunks, steps, labeller, chunk)
    kernel_ms = mixed_a + mixed_b + order_ms
    other.close()
    # one graph replay a chunk (clearing, scout, erosion, Poisson, ordering)
    # with a labeller of its own
    labeller.close()
    labeller = sf.Labeller(HEIGHT, WIDTH)
    plain_step, plain_ms, plain_order = chunk_step_timing(
        sf, plan, chunks, steps, labeller, chunk)
    plan.run(chunks[0])
    plan.wait()
    counts = plan.counts()
    signal = float(np.mean(counts))
    frame_bytes = 4.0 * HEIGHT * WIDTH
    algorithmic = chunk * (frame_bytes + 16.0 * signal)
    # full-size parity: the first `check` frames' lists against the
    # reference's own dext (the oracle's port where it was not built)
    from concurrent.futures import ThreadPoolExecutor
    from oracle import cpu

    host = np.zeros((check, HEIGHT, WIDTH), dtype=np.uint32)
    buf.download(host)
    ref = cpu.ref_dext if cpu.have_ref() else cpu.dext
    with ThreadPoolExecutor(check) as pool:
        expected = list(pool.map(lambda x: ref(x, capacity=x.size // 4)[1],
                                 host))
    diffs = sum(int(plan.output(f, int(counts[f])).tobytes() !=
                    expected[f].tobytes()) for f in range(check))
    plan.close()
    labeller.close()
    buf.close()
    return host, {
        "workload": f"{series}-frame series of 4148x4362 uint32, Poisson 50, "
                    f"~1500 spots/frame up to ~1e7 counts, gaps, 0.05% hot; "
                    f"{chunk} frames a step",
        "frames_per_s": round(chunk / per_step, 1),
        "path": "pipelined sparse threshold steps (two graphs replayed in "
                "turn) + 3-D labelling",
        "kernel_ms_per_chunk": round(kernel_ms, 4),
        "mixed_a_ms": round(mixed_a, 4),
        "mixed_b_ms": round(mixed_b, 4),
        "kernel_us_per_frame": round(kernel_ms / chunk * 1e3, 2),
        "kernel_gb_per_s": round(algorithmic / (kernel_ms / 1e3) / 1e9, 1),
        "kernel_hbm_frac": round(algorithmic / (kernel_ms / 1e3) / PEAK_HBM,
                                 4),
        "order_ms_per_chunk": round(order_ms, 4),
        "mean_signal_pixels_per_frame": round(signal, 1),
        "unpipelined": {
            "path": "one graph replay a chunk + 3-D labelling",
            "frames_per_s": round(chunk / plain_step, 1),
            "kernel_ms_per_chunk": round(plain_ms, 4),
            "kernel_us_per_frame": round(plain_ms / chunk * 1e3, 2),
        },
        "list_diffs": diffs,
        "parity_frames": check,
    }


SERVE_SERIES = (
    # (pixel_bytes, height, width, rate Hz, background, spot peak) -- the
    # four concurrent series of the serving workload
    (2, 2162, 2068, 750.0, 0.3, 20000.0),
    (2, 3269, 3110, 500.0, 0.3, 20000.0),
    (2, 4362, 4148, 250.0, 0.3, 20000.0),
    (4, 4362, 4148, 100.0, 50.0, 1.0e6),
)


def serve_latency(sf, fr, seconds, ring, budget_gb, max_batch, host=False,
                  which=None):
    """The serving loop on four concurrent series arriving in real time for
    `seconds` (3600 frames of the 750 Hz series at 4.8 s). Each series cycles
    through `ring` distinct frames, device-resident or (host) in pinned
    host memory, which the loop copies in on its batch streams; latencies
    run from each frame's scheduled arrival, so they include that copy."""
    server = sf.Server(int(budget_gb * (1 << 30)), max_batch=max_batch)
    buffers, rings, ids, rates, counts = [], [], [], [], []
    chosen = [SERVE_SERIES[k] for k in (which or range(len(SERVE_SERIES)))]
    for k, (pb, h, w, rate, bg, peak) in enumerate(chosen):
        buf = sf.DeviceBuffer(pb * h * w * ring)
        spots = fr.diffraction_spots(100 + k, h, w, ring,
                                     1500.0 * h * w / (HEIGHT * WIDTH),
                                     peak=peak)
        sf.synth_frames(buf, pb, h, w, ring, 100 + k, bg, True, 0.0005, spots)
        if host:
            pinned = sf.HostBuffer(pb * h * w * ring)
            buf.download(pinned.array(np.uint16 if pb == 2 else np.uint32,
                                      (ring, h, w)))
            buf.close()
            buf = pinned
        buffers.append(buf)
        rings.append([buf.ptr + pb * h * w * f for f in range(ring)])
        ids.append(server.open(pb, h, w, host=host))
        rates.append(rate)
        counts.append(int(round(rate * seconds)))
    sf.lib().sf_hip_synchronize()
    t0 = time.perf_counter()
    server.play(ids, rates, counts, rings)
    wall = time.perf_counter() - t0
    series = []
    worst = 0.0
    for sid, (pb, h, w, rate, bg, peak), n in zip(ids, chosen, counts):
        server.close_series(sid)
        fm = server.frames(sid)
        assert len(fm["frame"]) == n
        mask, table = fm["mask_ms"], fm["table_ms"]
        p99 = float(np.percentile(mask, 99))
        worst = max(worst, p99)
        series.append({
            "shape": f"{h}x{w} u{8 * pb}", "rate_hz": rate, "frames": n,
            "mask_p50_ms": round(float(np.percentile(mask, 50)), 3),
            "mask_p99_ms": round(p99, 3),
            "table_p50_ms": round(float(np.percentile(table, 50)), 3),
            "table_p99_ms": round(float(np.percentile(table, 99)), 3),
            "mean_signal_pixels": round(float(np.mean(fm["count"])), 1),
        })
    stats = server.stats()
    server.close()
    for buf in buffers:
        buf.close()
    return {
        "seconds_of_arrivals": seconds,
        "wall_s": round(wall, 3),
        "offered_fps": round(sum(counts) / seconds, 1),
        "achieved_fps": round(sum(counts) / wall, 1),
        "series": series,
        "batches": stats["batches"],
        "mean_batch": round(stats["batched_frames"] / max(stats["batches"], 1),
                            2),
        "max_batches_in_flight": stats["max_batches_in_flight"],
        "admission_waits": stats["admission_waits"],
        "peak_slot_gb": round(stats["peak_bytes"] / 2 ** 30, 3),
        "budget_gb": budget_gb,
        "frames_from": "pinned host memory, copied in on the batch streams"
                       if host else "device memory",
        "ingest_gb_per_s": round(stats["ingested_bytes"] / wall / 1e9, 2)
                           if host else None,
    }, worst


def serving_figures(sf, fr, args):
    """Device-resident frames (the serving workload as specified), then the
    same arriving in pinned host memory, copied in by the loop (latency
    includes the copy; a host-fed slot holds a frame more, so twice the
    budget), and the 750 Hz series alone that way."""
    serving, worst = serve_latency(sf, fr, args.serve_seconds,
                                   args.serve_ring, args.serve_budget_gb,
                                   args.serve_batch)
    log(f"serving from device memory: {serving}")
    serving_host, worst_host = serve_latency(
        sf, fr, args.serve_seconds, min(args.serve_ring, 12),
        2 * args.serve_budget_gb, args.serve_batch, host=True)
    log(f"serving from host memory: {serving_host}")
    serving_host["750_hz_series_alone"], _ = serve_latency(
        sf, fr, args.serve_seconds, min(args.serve_ring, 12),
        2 * args.serve_budget_gb, args.serve_batch, host=True, which=[0])
    log(f"the 750 Hz series alone from host memory: "
        f"{serving_host['750_hz_series_alone']}")
    queues = os.environ.get("GPU_MAX_HW_QUEUES", "HIP's default")
    for record in (serving, serving_host):
        record["process"] = f"its own, {queues} hardware queues"
    return {"serving": serving, "serving_host": serving_host,
            "latency": round(worst, 3), "latency_host": round(worst_host, 3)}


def measure_decode(sf, series, chunk: int, frame_bytes: int,
                   steps: int) -> tuple[dict, list]:
    """One chunk of the series as the detector's HDF5 files hold frames
    (bitshuffle+LZ4, 8192-byte blocks), compressed on the host, resident in
    HBM, decoded into frames in one launch (sf_hip_decoder_*). Device time
    from HIP events on the decoder's stream; three decoded frames compared
    with the originals byte for byte."""
    from concurrent.futures import ThreadPoolExecutor

    host = np.empty((chunk, HEIGHT, WIDTH), np.uint16)
    series.download(host)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(16) as pool:
        chunks = list(pool.map(sf.bslz4_compress, host))
    compress_s = time.perf_counter() - t0
    sizes = [len(c) for c in chunks]
    offsets = np.cumsum([0] + [(n + 255) // 256 * 256 for n in sizes])
    stage = sf.DeviceBuffer(int(offsets[-1]))
    for c, off in zip(chunks, offsets):
        stage.upload(np.frombuffer(c, np.uint8), int(off))
    out = sf.DeviceBuffer(chunk * frame_bytes)
    dec = sf.Decoder(16, HEIGHT, WIDTH, max_frames=chunk)
    dec.index(chunks)
    src = [stage.ptr + int(off) for off in offsets[:chunk]]
    dst = [out.ptr + f * frame_bytes for f in range(chunk)]
    dec.decode(src, dst)
    dec.wait()
    dec.set_timing(True)
    ms = []
    for _ in range(steps):
        dec.decode(src, dst)
        dec.wait()
        ms.append(dec.last_ms())
    dec.set_timing(False)
    diffs = 0
    check = np.empty((HEIGHT, WIDTH), np.uint16)
    for f in (0, chunk // 2, chunk - 1):
        out.download(check, f * frame_bytes)
        diffs += int(np.count_nonzero(check != host[f]))
    dec.close()
    stage.close()
    out.close()
    t = float(np.median(ms)) / 1e3
    compressed = float(sum(sizes))
    decoded = float(chunk * frame_bytes)
    return chunks, {
        "workload": f"{chunk} frames of {HEIGHT} x {WIDTH} u16 as "
                    "bitshuffle+LZ4 chunks (8192-byte blocks) in HBM, "
                    "decoded into frames in one launch",
        "ms_per_chunk": round(t * 1e3, 4),
        "us_per_frame": round(t / chunk * 1e6, 3),
        "frames_per_s": round(chunk / t, 1),
        "compressed_bytes_per_frame": round(compressed / chunk),
        "ratio": round(decoded / compressed, 3),
        "decoded_GB_per_s": round(decoded / t / 1e9, 1),
        "hbm_frac": round((compressed + decoded) / t / 8e12, 4),
        "host_compress_s": round(compress_s, 2),
        "frame_diffs": diffs,
    }


def measure_driver(sf, chunks: list, repeats: int = 4):
    """The series driver (sf_hip_driver_*) over the compressed chunks from
    host memory: submit by number, decode, threshold and group in order,
    finish with the summary; host clock over the whole run, which includes
    the chunks' copies to the device (PCIe) and the driver's set-up. The
    fastest of `repeats` runs: a run's set-up allocates device and pinned
    memory, and one such allocation has taken hundreds of milliseconds on
    the MI355X box now and then."""
    import tempfile

    times, lines, spots, tables, write_s = [], [], 0, None, None
    path = os.path.join(tempfile.mkdtemp(), "strong.refl")
    for r in range(repeats):
        t0 = time.perf_counter()
        d = sf.Driver(16, HEIGHT, WIDTH, chunk=len(chunks))
        for z, c in enumerate(chunks):
            d.submit(z, c)
        d.finish()
        times.append(time.perf_counter() - t0)
        lines = d.log().splitlines()
        tables = d.results()
        spots = len(tables[0])
        if r == repeats - 1:  # the .refl writer on the chunk's tables
            t0 = time.perf_counter()
            d.write(path)
            write_s = time.perf_counter() - t0
        d.close()
    t = min(times)
    refl = {"path": path, "spots": tables[0], "pixels": tables[1],
            "write_ms": round(write_s * 1e3, 1),
            "bytes": os.path.getsize(path)}
    return refl, {
        "workload": f"{len(chunks)} bitshuffle+LZ4 chunks of {HEIGHT} x "
                    f"{WIDTH} u16 in host memory through the driver (copy "
                    "to the device, decode, threshold, group, summary)",
        "s_per_run": round(t, 4),
        "s_per_run_median": round(float(np.median(times)), 4),
        "frames_per_s": round(len(chunks) / t, 1),
        "spots": spots,
        "summary": lines[1:3],
        "refl_write_ms": refl["write_ms"],
        "refl_bytes": refl["bytes"],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="HIP hardware queues of the process "
                         "(GPU_MAX_HW_QUEUES; 0 keeps the environment's)")
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chunk", type=int, default=60)
    ap.add_argument("--frames", type=int, default=1800)
    ap.add_argument("--band", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=32)
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-label", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.8,
                    help="seconds of real-time arrivals in the serving "
                         "measurement (0: skip it)")
    ap.add_argument("--serve-ring", type=int, default=48)
    ap.add_argument("--serve-budget-gb", type=float, default=4.0)
    ap.add_argument("--serve-batch", type=int, default=16)
    ap.add_argument("--serve-hw-queues", type=int, default=4,
                    help="hardware queues of the serving loop's process")
    ap.add_argument("--serve-only", action="store_true",
                    help="measure only the serving loop and print its "
                         "figures as one JSON line (the bench runs this "
                         "in a process of its own)")
    ap.add_argument("--other-steps", type=int, default=10,
                    help="steps of the batched and 32-bit workloads (0: skip)")
    ap.add_argument("--decode-steps", type=int, default=5,
                    help="timed decodes of a chunk of bitshuffle+LZ4 chunks "
                         "(0: skip)")
    ap.add_argument("--sweep-repeats", type=int, default=3,
                    help="the labeller density x span sweep, best of this "
                         "many (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy, three times (the "
                         "known-size kernel that scales FETCH_SIZE and "
                         "WRITE_SIZE under rocprofv3 --pmc)")
    args = ap.parse_args()
    # HIP maps streams onto hardware queues round robin and reads this once,
    # when it starts. A pipelined series keeps four streams of its own busy
    # (two plans', the labeller's and its snapshots'); with four queues, the
    # default of HIP and of the environment these runs start in, they share
    # queues with the process's other streams, and work on a shared queue
    # waits for work it does not depend on. Eight: 1.24 against 1.30 ms a
    # pipelined step.
    if args.hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.calibrate_copy:
        import torch

        x = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the hot path has no multi-GPU form "
              "here (frames shard across GPUs with no exchange; only the "
              "3-D labelling would pass a frontier between shards)",
              file=sys.stderr)
        return 2

    import spotfinder as sf
    import frames as fr

    if args.serve_only:
        print(json.dumps(serving_figures(sf, fr, args)))
        return 0

    if not sf.available():
        print("no HIP device", file=sys.stderr)
        return 1
    sf.report_windows(False)
    frame_bytes = HEIGHT * WIDTH * 2
    n_frames = args.frames
    chunk = args.chunk
    series = sf.DeviceBuffer(frame_bytes * n_frames)
    spots = fr.diffraction_spots(1, HEIGHT, WIDTH, n_frames, 1500.0)
    t0 = time.perf_counter()
    sf.synth_frames(series, 2, HEIGHT, WIDTH, n_frames, 1, 0.3, True,
                    0.0005, spots)
    log(f"generated {n_frames} frames in {time.perf_counter() - t0:.2f} s")
    ptr = [series.ptr + frame_bytes * f for f in range(n_frames)]
    n_chunks = n_frames // chunk

    plan = sf.Plan(2, [(HEIGHT, WIDTH)] * chunk, band_rows=args.band)
    tuned = plan.autotune(ptr[:chunk], 3)
    log(f"autotuned (band rows, blocks): {tuned}")
    labeller = None
    if not args.no_label and hasattr(sf, "Labeller"):
        labeller = sf.Labeller(HEIGHT, WIDTH)
    path = ("pipelined threshold steps (two graphs replayed in turn: "
            "scout of chunk k+1 mixed with erosion, then with the Poisson "
            "pass, of chunk k; ordering of chunk k)"
            + (" + 3-D labelling and centroids" if labeller else ""))

    # Two plans in turn. A step finishes the chunk bound to one plan (its
    # erosion, Poisson pass and ordering) while it scouts the next chunk,
    # bound to the other: two captured graphs, replayed alternately, with
    # the next chunk's frame pointers bound before each. The finished chunk
    # then goes to the labeller (which gathers its lists on a stream of its
    # own; the plan's next write of them waits for that), whose kernels
    # overlap the next step's. Frame numbers keep ascending as the series
    # is replayed.
    plans = [plan, sf.Plan(2, [(HEIGHT, WIDTH)] * chunk, band_rows=args.band)]
    next_frame = [0]  # the first frame number of the next chunk to label

    def bind(p, s):
        c = s % n_chunks
        p.bind(ptr[c * chunk:(c + 1) * chunk])

    def step(s, graphs=True):
        cur, nxt = plans[s % 2], plans[(s + 1) % 2]
        bind(nxt, s + 1)
        (sf.replay_pipeline if graphs else sf.pipeline)(cur, nxt)
        if labeller is not None:
            labeller.add_plan(cur, first_frame=next_frame[0])
        next_frame[0] += chunk

    def drain(last):
        sf.pipeline(plans[last % 2], None)
        if labeller is not None:
            labeller.add_plan(plans[last % 2], first_frame=next_frame[0])
        next_frame[0] += chunk
        for p in plans:
            p.wait()
        if labeller is not None:
            labeller.wait()

    sf.capture_pipeline(plans[0], plans[1])
    sf.capture_pipeline(plans[1], plans[0])
    bind(plans[0], 0)
    sf.pipeline(None, plans[0])
    for s in range(args.warmup):
        step(s)
    drain(args.warmup)
    if labeller is not None:
        # the timed steps, their drain, the unpipelined steps and the
        # labeller alone below
        reserve_for(labeller, args.warmup + 1, 2 * args.steps + 6)
    sf.lib().sf_hip_synchronize()
    bind(plans[0], 0)
    sf.pipeline(None, plans[0])
    sf.lib().sf_hip_synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(s)
    # every chunk of the timed steps labelled, not just enqueued
    if labeller is not None:
        labeller.wait()
    sf.lib().sf_hip_synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps * chunk / elapsed
    drain(args.steps)
    sf.lib().sf_hip_synchronize()

    # Each launch of a pipelined step by HIP events on the plan's stream:
    # (scout half + erosion, scout half + Poisson, ordering), eager steps.
    mixed_a, mixed_b, mixed_o, counts = [], [], [], []
    bind(plans[0], 0)
    sf.pipeline(None, plans[0])
    for s in range(5):
        cur = plans[s % 2]
        cur.set_timing(True)
        bind(plans[(s + 1) % 2], s + 1)
        sf.pipeline(cur, plans[(s + 1) % 2])
        cur.wait()
        a, b, o = cur.last_pass_times()
        cur.set_timing(False)
        mixed_a.append(a)
        mixed_b.append(b)
        mixed_o.append(o)
        counts.append(cur.counts())
    sf.pipeline(plans[1], None)
    plans[1].wait()

    # The same threshold unpipelined (one graph replay a chunk: clearing,
    # scout, erosion, Poisson, ordering) with the labeller, two plans in
    # turn, for comparison; and its passes by HIP events.
    pending = []

    def replay_step(s):
        c = s % n_chunks
        p = plans[s % 2]
        p.bind(ptr[c * chunk:(c + 1) * chunk])
        p.replay()
        if labeller is not None and pending:
            done, first = pending.pop()
            labeller.add_plan(done, first_frame=first)
        pending.append((p, next_frame[0]))
        next_frame[0] += chunk

    for p in plans:
        p.capture(ptr[:chunk])
    for s in range(2):
        replay_step(s)
    sf.lib().sf_hip_synchronize()
    t1 = time.perf_counter()
    for s in range(args.steps):
        replay_step(s)
    if labeller is not None and pending:
        done, first = pending.pop()
        labeller.add_plan(done, first_frame=first)
    pending.clear()
    if labeller is not None:
        labeller.wait()
    sf.lib().sf_hip_synchronize()
    replay_fps = args.steps * chunk / (time.perf_counter() - t1)
    plans[1].close()
    scout_ms, rest_ms, order_ms = [], [], []
    for s in range(5):
        c = s % n_chunks
        plan.set_timing(True)
        plan.run(ptr[c * chunk:(c + 1) * chunk])
        plan.wait()
        a, b, o = plan.last_pass_times()
        scout_ms.append(a)
        rest_ms.append(b)
        order_ms.append(o)
    plan.set_timing(False)
    # The labeller alone, by the host clock: a chunk thresholded and waited
    # for, then labelled and waited for, so that nothing shares the device.
    # Frame numbers carry on from the timed loop's.
    label_ms = []
    if labeller is not None:
        for s in range(4):
            c = s % n_chunks
            plan.run(ptr[c * chunk:(c + 1) * chunk])
            plan.wait()
            t_a = time.perf_counter()
            labeller.add_plan(plan, first_frame=next_frame[0])
            next_frame[0] += chunk
            labeller.wait()
            label_ms.append((time.perf_counter() - t_a) * 1e3)
    # The detector mask stamped into a chunk (sf_hip_stamp_mask): the module
    # gaps of the series' own geometry as device bits. Those pixels already
    # hold the sentinel, so stamping leaves the series as it is.
    frame0 = np.zeros((HEIGHT, WIDTH), dtype=np.uint16)
    series.download(frame0)
    gap_mask = sf.PixelMask(frame0 == 0xFFFF)
    chunk_ptrs = ptr[:chunk]
    gap_mask.stamp(2, chunk_ptrs)
    sf.lib().sf_hip_synchronize()
    stamp_s = []
    for _ in range(5):
        t_a = time.perf_counter()
        gap_mask.stamp(2, chunk_ptrs)
        sf.lib().sf_hip_synchronize()
        stamp_s.append(time.perf_counter() - t_a)
    masked = int(np.count_nonzero(frame0 == 0xFFFF))
    stamp = {
        "workload": f"the series' module-gap mask ({masked:,} of "
                    f"{HEIGHT * WIDTH:,} pixels) stamped into {chunk} frames",
        "us_per_frame": round(float(np.median(stamp_s)) / chunk * 1e6, 3),
        "mask_bytes": (HEIGHT * WIDTH + 31) // 32 * 4,
        "written_bytes_per_frame": 2 * masked,
    }
    gap_mask.close()
    decode = driver = decode_chunks = refl = None
    if args.decode_steps > 0:
        decode_chunks, decode = measure_decode(sf, series, chunk,
                                               frame_bytes, args.decode_steps)
        log(f"chunk decode: {decode}")
        refl, driver = measure_driver(sf, decode_chunks)
        log(f"driver: {driver}")
    mean_signal = float(np.mean(np.concatenate(counts)))
    # The dominant kernel: the slower of the two mixed launches. Its
    # algorithmic bytes are its half of the next chunk's frames (the scout
    # reads every pixel once) and, for the Poisson launch, the survivors it
    # writes (16 bytes each); the erosion's and Poisson pass's reads around
    # active cells are not algorithmic.
    a_s = float(np.median(mixed_a)) / 1e3
    b_s = float(np.median(mixed_b)) / 1e3
    o_s = float(np.median(mixed_o)) / 1e3
    half_frames = chunk // 2
    kernel_bytes = {
        MIXED_A: half_frames * float(frame_bytes),
        MIXED_B: (chunk - half_frames) * float(frame_bytes) +
        16.0 * chunk * mean_signal,
    }
    kernel = MIXED_B if b_s >= a_s else MIXED_A
    kernel_s = max(a_s, b_s)
    algorithmic = kernel_bytes[kernel]
    achieved = algorithmic / kernel_s
    # the whole threshold of a chunk: one chunk of frames read, its ordered
    # lists written, over the step's three parts
    threshold_s = a_s + b_s + o_s
    threshold_bytes = chunk * (frame_bytes + 16.0 * mean_signal)
    threshold_achieved = threshold_bytes / threshold_s
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "threshold_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            t = json.load(fh)
        if t.get("chunk") == chunk and t.get("kernel") == kernel:
            traffic = t.get("bytes_per_launch")

    # The plain one-frame-per-launch form on the same frames.
    plain = sf.Plan(2, [(HEIGHT, WIDTH)] * chunk, band_rows=args.band,
                    form=1)
    plain.run(ptr[:chunk])
    plain.wait()
    t1 = time.perf_counter()
    for s in range(args.plain_steps):
        c = s % n_chunks
        plain.run(ptr[c * chunk:(c + 1) * chunk])
    plain.wait()
    plain_fps = args.plain_steps * chunk / (time.perf_counter() - t1)
    plain.close()

    # The measured chunk's device lists (and a fresh device labelling of
    # them, inside cpu_baseline) for the CPU path, which thresholds a
    # sample of the same frames and groups the whole chunk.
    sample = min(args.cpu_sample, chunk)
    plan.run(ptr[:chunk])
    plan.wait()
    counts0 = plan.counts()
    device_lists = [plan.output(f, int(counts0[f])) for f in range(chunk)]
    host = np.zeros((sample, HEIGHT, WIDTH), dtype=np.uint16)
    series.download(host)
    baseline, diffs = cpu_baseline(sf, host, device_lists, plan,
                                   decode_chunks, refl, args.cpu_repeats)
    del host

    sweep = None
    if args.sweep_repeats > 0:
        sweep = labeller_sweep(sf, args.sweep_repeats)
        log(f"labeller sweep: {sweep['rows']}")

    others = {}
    if args.other_steps > 0:
        plan.close()
        labeller = None
        series.close()
        others["batched_uneven"] = batched_workload(sf, fr, args.other_steps)
        log(f"batched: {others['batched_uneven']}")
        u32_host, others["u32"] = u32_workload(sf, fr, args.other_steps)
        log(f"u32: {others['u32']}")
        baseline["u32_threshold"] = cpu_u32_threshold(
            u32_host, baseline["host"]["nproc"], args.cpu_repeats)
        del u32_host

    latency, latency_host, serving, serving_host = None, None, None, None
    if args.serve_seconds > 0:
        plan.close()
        labeller = None
        series.close()
        # The serving loop in a process of its own, with HIP's default of
        # four hardware queues. Its streams (a few batch streams, and each
        # series' labeller and snapshot streams) then share queues; with
        # eight, as the pipelined steps above use, the chunk-to-table p99
        # roughly doubled (0.61-0.79 against 1.2-1.56 ms) at the same
        # throughput.
        import subprocess

        child = subprocess.run(
            [sys.executable, os.path.abspath(__file__), "--serve-only",
             "--hw-queues", str(args.serve_hw_queues),
             "--serve-seconds", str(args.serve_seconds),
             "--serve-ring", str(args.serve_ring),
             "--serve-budget-gb", str(args.serve_budget_gb),
             "--serve-batch", str(args.serve_batch)],
            capture_output=True, text=True)
        sys.stderr.write(child.stderr)
        if child.returncode != 0:
            raise SystemExit(f"the serving loop's process failed "
                             f"({child.returncode})")
        lines = [x for x in child.stdout.splitlines() if x.startswith("{")]
        figures = json.loads(lines[-1])
        serving, serving_host = figures["serving"], figures["serving_host"]
        latency, latency_host = figures["latency"], figures["latency_host"]

    record = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic (device-generated, seeded)",
        "config": {
            "workload": "1800-frame series of 4148x4362 uint16, "
                        "Poisson 0.3, ~1500 spots/frame, gaps, 0.05% hot",
            "frames_per_step": chunk,
            "autotuned": {"band_rows": tuned[0], "blocks": tuned[1]},
            "series_frames": n_frames,
            "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "HIP's default"),
            "path": path,
        },
        "gb_per_s": round(fps * (frame_bytes + 16 * mean_signal) / 1e9, 2),
        "hbm_frac_whole_path": round(
            fps * (frame_bytes + 16 * mean_signal) / PEAK_HBM, 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved / 1e9, 2),
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM, 4),
            "traffic": traffic,
            "kernel": kernel,
            "kernel_ms_per_launch": round(kernel_s * 1e3, 4),
            "algorithmic_bytes_per_launch": int(algorithmic),
        },
        "threshold_roofline": {
            "kernels": f"{MIXED_A} + {MIXED_B} + ordering (a pipelined "
                       "step: one chunk's frames read, one chunk's lists "
                       "written)",
            "achieved": round(threshold_achieved / 1e9, 2),
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": round(threshold_achieved / PEAK_HBM, 4),
            "ms_per_step": round(threshold_s * 1e3, 4),
            "mixed_a_ms": round(a_s * 1e3, 4),
            "mixed_b_ms": round(b_s * 1e3, 4),
            "order_ms": round(o_s * 1e3, 4),
            "algorithmic_bytes_per_step": int(threshold_bytes),
        },
        "unpipelined": {
            "path": "one graph replay a chunk (clear, scout, erosion, "
                    "Poisson, ordering) + 3-D labelling",
            "frames_per_s": round(replay_fps, 2),
            "scout_ms": round(float(np.median(scout_ms)), 4),
            "erode_poisson_ms": round(float(np.median(rest_ms)), 4),
            "order_ms": round(float(np.median(order_ms)), 4),
            "threshold_frac": round(
                threshold_bytes / (float(np.median(np.add(scout_ms, rest_ms)))
                                   / 1e3) / PEAK_HBM, 4),
        },
        "labeller_alone": None if not label_ms else {
            "ms_per_chunk": round(float(np.median(label_ms)), 4),
            "chunk_frames": chunk,
            "frames_per_s": round(chunk / float(np.median(label_ms)) * 1e3,
                                  1),
        },
        "pixel_mask": stamp,
        "chunk_decode": decode,
        "driver": driver,
        "plain_form_fps": round(plain_fps, 2),
        "speedup_vs_plain": round(fps / plain_fps, 3),
        "mean_signal_pixels_per_frame": round(mean_signal, 1),
        "mask_diffs": diffs,
        "table_diffs": baseline["table_diffs"],
        "cpu_baseline": baseline,
        "p99_latency_ms": latency,
        "p99_latency_host_ms": latency_host,
        "serving": serving,
        "serving_host": serving_host,
        "other_workloads": others,
        "labeller_sweep": sweep,
    }
    print(json.dumps(record))
    plan.close()
    series.close()
    decoded_ok = (decode is None or decode["frame_diffs"] == 0) and \
        baseline["decode_port_diffs"] == 0 and \
        baseline["refl_identical"] is not False
    tables_ok = baseline["table_diffs"] in (None, 0) and \
        others.get("u32", {}).get("list_diffs", 0) == 0 and \
        all(r.get("table_diffs", 0) == 0 for r in (sweep or {}).get("rows", []))
    return 0 if diffs == 0 and decoded_ok and tables_ok else 3


if __name__ == "__main__":
    sys.exit(main())
