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
# the dominant kernel of a form-0 plan at 16 bits: the sparse threshold's
# scout (its erosion and Poisson passes follow as erode_kernel and
# poisson_kernel)
KERNEL = "scout_kernel"
METRIC = ("frames/s + HBM GB/s (% peak), 4148×4362 uint16, 0 mask diffs; "
          "p99 latency")


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def cpu_baseline(sf, host_frames, device_lists, threads, seconds,
                 chunks=None, refl=None):
    """The reference CPU path (or the oracle port) on a bounded sample: the
    threshold over the sample frames on `threads` host threads, repeated
    until about `seconds` have passed, and the reference's streaming
    Labeller over the same frames' lists on one thread. The path's rate is
    the slower of the two stages, as in the reference's pipeline (threshold
    workers feeding one grouping thread)."""
    from oracle import cpu

    kind = "reference" if cpu.have_ref() else "port"
    many = cpu.ref_dext_many if kind == "reference" else cpu.oracle_dext_many
    n = len(host_frames)
    passes, total, elapsed = 0, 0, 0.0
    while passes == 0 or elapsed < seconds:
        t0 = time.perf_counter()
        total += many(host_frames, threads)
        elapsed += time.perf_counter() - t0
        passes += 1
    threshold_fps = passes * n / elapsed
    label_fps = None
    if kind == "reference":
        labeller = cpu.RefLabeller(HEIGHT, WIDTH)
        t0 = time.perf_counter()
        for f, signal in enumerate(device_lists):
            labeller.add(f, signal)
        labeller.finish()
        label_fps = len(device_lists) / (time.perf_counter() - t0)
    # parity on part of the same sample: every SignalPixel field
    diffs = 0
    for f, frame in enumerate(host_frames[:8]):
        expected = (cpu.ref_dext if kind == "reference" else cpu.dext)(
            frame, capacity=frame.size // 8)[1]
        got = device_lists[f]
        if len(got) != len(expected) or got.tobytes() != expected.tobytes():
            diffs += 1
    value = threshold_fps if label_fps is None else min(threshold_fps,
                                                        label_fps)
    # the oracle's decompress::image on up to two of the sample's
    # bitshuffle+LZ4 chunks, one thread (a plain restatement: bitshuffle's SIMD library is
    # not here), checked against the frames
    decode_ms, decode_diffs = None, 0
    sample = chunks[:min(2, len(host_frames))] if chunks else []
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
    return {
        "value": round(value, 3),
        "unit": "frames/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{n} frames of the same series: dext on {threads} host "
                  f"threads (one frame per thread), {passes} passes, "
                  f"{elapsed:.1f} s; then the reference Labeller over their "
                  f"lists on one thread",
        "threshold_fps": round(threshold_fps, 3),
        "labeller_fps": None if label_fps is None else round(label_fps, 1),
        "signal_pixels_per_pass": int(total // passes),
        "parity_frames": min(n, 8),
        "decode_port_ms_per_frame": None if decode_ms is None
        else round(decode_ms, 1),
        "decode_port_diffs": decode_diffs,
        "refl_reference_write_ms": None if ref_write_ms is None
        else round(ref_write_ms, 1),
        "refl_identical": refl_same,
    }, diffs


BATCH_SHAPES = [(2162, 2068)] * 32 + [(3269, 3110)] * 16 + [(4362, 4148)] * 8


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


def u32_workload(sf, fr, steps, chunk=30, series=120):
    """The 32-bit path: a device-resident series of 4148 x 4362 uint32
    frames (Poisson 50, spots up to ~10^7 counts, masked gaps at
    0xFFFFFFFF, hot pixels), thresholded in chunks by the sweep kernel with
    autotuned bands, then labelled on the device."""
    buf = sf.DeviceBuffer(4 * HEIGHT * WIDTH * series)
    spots = fr.diffraction_spots(300, HEIGHT, WIDTH, series, 1500.0,
                                 peak=1.0e6, min_peak=100.0)
    sf.synth_frames(buf, 4, HEIGHT, WIDTH, series, 300, 50.0, True, 0.0005,
                    spots)
    ptr = [buf.ptr + 4 * HEIGHT * WIDTH * f for f in range(series)]
    chunks = [ptr[c * chunk:(c + 1) * chunk] for c in range(series // chunk)]
    plan = sf.Plan(4, [(HEIGHT, WIDTH)] * chunk)
    tuned = plan.autotune(chunks[0], 3)
    labeller = sf.Labeller(HEIGHT, WIDTH)
    per_step, kernel_ms, order_ms = chunk_step_timing(
        sf, plan, chunks, steps, labeller, chunk)
    plan.run(chunks[0])
    plan.wait()
    signal = float(np.mean(plan.counts()))
    frame_bytes = 4.0 * HEIGHT * WIDTH
    algorithmic = chunk * (frame_bytes + 16.0 * signal)
    plan.close()
    labeller.close()
    buf.close()
    return {
        "workload": f"{series}-frame series of 4148x4362 uint32, Poisson 50, "
                    f"~1500 spots/frame up to ~1e7 counts, gaps, 0.05% hot; "
                    f"{chunk} frames a step",
        "frames_per_s": round(chunk / per_step, 1),
        "path": "sweep threshold + ordered compaction + 3-D labelling",
        "kernel_ms_per_chunk": round(kernel_ms, 4),
        "kernel_us_per_frame": round(kernel_ms / chunk * 1e3, 2),
        "kernel_gb_per_s": round(algorithmic / (kernel_ms / 1e3) / 1e9, 1),
        "kernel_hbm_frac": round(algorithmic / (kernel_ms / 1e3) / PEAK_HBM,
                                 4),
        "order_ms_per_chunk": round(order_ms, 4),
        "mean_signal_pixels_per_frame": round(signal, 1),
        "autotuned": {"band_rows": tuned[0], "blocks": tuned[1]},
    }


SERVE_SERIES = (
    # (pixel_bytes, height, width, rate Hz, background, spot peak) -- the
    # four concurrent series of the serving workload
    (2, 2162, 2068, 750.0, 0.3, 20000.0),
    (2, 3269, 3110, 500.0, 0.3, 20000.0),
    (2, 4362, 4148, 250.0, 0.3, 20000.0),
    (4, 4362, 4148, 100.0, 50.0, 1.0e6),
)


def serve_latency(sf, fr, seconds, ring, budget_gb, max_batch):
    """The serving loop on four concurrent series arriving in real time for
    `seconds` (3600 frames of the 750 Hz series at 4.8 s). Each series cycles
    through `ring` distinct device-resident frames; latencies run from each
    frame's scheduled arrival."""
    server = sf.Server(int(budget_gb * (1 << 30)), max_batch=max_batch)
    buffers, rings, ids, rates, counts = [], [], [], [], []
    for k, (pb, h, w, rate, bg, peak) in enumerate(SERVE_SERIES):
        buf = sf.DeviceBuffer(pb * h * w * ring)
        spots = fr.diffraction_spots(100 + k, h, w, ring,
                                     1500.0 * h * w / (HEIGHT * WIDTH),
                                     peak=peak)
        sf.synth_frames(buf, pb, h, w, ring, 100 + k, bg, True, 0.0005, spots)
        buffers.append(buf)
        rings.append([buf.ptr + pb * h * w * f for f in range(ring)])
        ids.append(server.open(pb, h, w))
        rates.append(rate)
        counts.append(int(round(rate * seconds)))
    sf.lib().sf_hip_synchronize()
    t0 = time.perf_counter()
    server.play(ids, rates, counts, rings)
    wall = time.perf_counter() - t0
    series = []
    worst = 0.0
    for sid, (pb, h, w, rate, bg, peak), n in zip(ids, SERVE_SERIES, counts):
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
    }, worst


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


def measure_driver(sf, chunks: list, repeats: int = 2):
    """The series driver (sf_hip_driver_*) over the compressed chunks from
    host memory: submit by number, decode, threshold and group in order,
    finish with the summary; host clock over the whole run, which includes
    the chunks' copies to the device (PCIe)."""
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
        "frames_per_s": round(len(chunks) / t, 1),
        "spots": spots,
        "summary": lines[1:3],
        "refl_write_ms": refl["write_ms"],
        "refl_bytes": refl["bytes"],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chunk", type=int, default=60)
    ap.add_argument("--frames", type=int, default=1800)
    ap.add_argument("--band", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-label", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.8,
                    help="seconds of real-time arrivals in the serving "
                         "measurement (0: skip it)")
    ap.add_argument("--serve-ring", type=int, default=48)
    ap.add_argument("--serve-budget-gb", type=float, default=4.0)
    ap.add_argument("--serve-batch", type=int, default=16)
    ap.add_argument("--other-steps", type=int, default=10,
                    help="steps of the batched and 32-bit workloads (0: skip)")
    ap.add_argument("--decode-steps", type=int, default=5,
                    help="timed decodes of a chunk of bitshuffle+LZ4 chunks "
                         "(0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy, three times (the "
                         "known-size kernel that scales FETCH_SIZE and "
                         "WRITE_SIZE under rocprofv3 --pmc)")
    args = ap.parse_args()
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
    path = ("threshold + ordered compaction (one graph replay)"
            + (" + 3-D labelling and centroids" if labeller else ""))

    # Two plans in turn, each with one captured graph (threshold, ordering)
    # replayed per chunk; binding a chunk only rewrites the frame descriptors
    # it reads. Step s replays chunk s, then hands chunk s-1 (done or nearly)
    # to the labeller, so the host never waits on the threshold in flight
    # and chunk s-1's labelling overlaps chunk s's threshold. The labeller's
    # gather runs on the producing plan's stream, ahead of that plan's next
    # replay.
    plans = [plan, sf.Plan(2, [(HEIGHT, WIDTH)] * chunk, band_rows=args.band)]
    for p in plans:
        p.capture(ptr[:chunk])
    pending = []  # the replayed chunk not yet labelled: (plan, first frame)

    def step(s):
        c = s % n_chunks
        p = plans[s % 2]
        p.bind(ptr[c * chunk:(c + 1) * chunk])
        p.replay()
        if labeller is not None and pending:
            # frame numbers keep ascending as the series is replayed
            done, first = pending.pop()
            labeller.add_plan(done, first_frame=first)
        pending.append((p, s * chunk))

    def drain():
        if labeller is not None and pending:
            done, first = pending.pop()
            labeller.add_plan(done, first_frame=first)
        pending.clear()
        for p in plans:
            p.wait()
        if labeller is not None:
            labeller.wait()

    for s in range(args.warmup):
        step(s)
    drain()
    sf.lib().sf_hip_synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(args.warmup + s)
    drain()
    sf.lib().sf_hip_synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps * chunk / elapsed
    plans[1].close()

    # The scout alone, and the threshold's passes together, from HIP events
    # on the plan's stream.
    scout_ms, rest_ms, order_ms, counts = [], [], [], []
    for s in range(5):
        c = s % n_chunks
        plan.set_timing(True)
        plan.run(ptr[c * chunk:(c + 1) * chunk])
        plan.wait()
        a, b, o = plan.last_pass_times()
        scout_ms.append(a)
        rest_ms.append(b)
        order_ms.append(o)
        counts.append(plan.counts())
    plan.set_timing(False)
    # The labeller alone, by the host clock: a chunk thresholded and waited
    # for, then labelled and waited for, so that nothing shares the device.
    # Frame numbers carry on from the timed loop's.
    label_ms = []
    if labeller is not None:
        first = (args.warmup + args.steps) * chunk
        for s in range(4):
            c = s % n_chunks
            plan.run(ptr[c * chunk:(c + 1) * chunk])
            plan.wait()
            t_a = time.perf_counter()
            labeller.add_plan(plan, first_frame=first + s * chunk)
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
    kernel_s = float(np.median(scout_ms)) / 1e3
    threshold_s = float(np.median(np.add(scout_ms, rest_ms))) / 1e3
    mean_signal = float(np.mean(np.concatenate(counts)))
    algorithmic = chunk * float(frame_bytes)  # the scout reads every pixel
    achieved = algorithmic / kernel_s
    threshold_bytes = chunk * (frame_bytes + 16.0 * mean_signal)
    threshold_achieved = threshold_bytes / threshold_s
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "threshold_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            t = json.load(fh)
        if t.get("chunk") == chunk and t.get("kernel") == KERNEL:
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

    # Device lists for the CPU sample, and the CPU baseline itself.
    sample = min(args.cpu_sample, chunk)
    plan.run(ptr[:chunk])
    plan.wait()
    counts0 = plan.counts()
    device_lists = [plan.output(f, int(counts0[f])) for f in range(sample)]
    host = np.zeros((sample, HEIGHT, WIDTH), dtype=np.uint16)
    series.download(host)
    threads = min(os.cpu_count() or 1, 16)
    baseline, diffs = cpu_baseline(sf, host, device_lists, threads,
                                   args.cpu_seconds, decode_chunks, refl)

    others = {}
    if args.other_steps > 0:
        plan.close()
        labeller = None
        series.close()
        others["batched_uneven"] = batched_workload(sf, fr, args.other_steps)
        log(f"batched: {others['batched_uneven']}")
        others["u32"] = u32_workload(sf, fr, args.other_steps)
        log(f"u32: {others['u32']}")

    latency, serving = None, None
    if args.serve_seconds > 0:
        plan.close()
        labeller = None
        series.close()
        serving, worst = serve_latency(sf, fr, args.serve_seconds,
                                       args.serve_ring, args.serve_budget_gb,
                                       args.serve_batch)
        latency = round(worst, 3)
        log(f"serving: {serving}")

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
            "kernel": KERNEL,
            "kernel_ms_per_launch": round(kernel_s * 1e3, 4),
            "algorithmic_bytes_per_launch": int(algorithmic),
        },
        "threshold_roofline": {
            "kernels": "scout_kernel + erode_kernel + poisson_kernel",
            "achieved": round(threshold_achieved / 1e9, 2),
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": round(threshold_achieved / PEAK_HBM, 4),
            "ms_per_launch": round(threshold_s * 1e3, 4),
            "order_ms_per_launch": round(float(np.median(order_ms)), 4),
            "algorithmic_bytes_per_launch": int(threshold_bytes),
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
        "cpu_baseline": baseline,
        "p99_latency_ms": latency,
        "serving": serving,
        "other_workloads": others,
    }
    print(json.dumps(record))
    plan.close()
    series.close()
    decoded_ok = (decode is None or decode["frame_diffs"] == 0) and \
        baseline["decode_port_diffs"] == 0 and \
        baseline["refl_identical"] is not False
    return 0 if diffs == 0 and decoded_ok else 3


if __name__ == "__main__":
    sys.exit(main())
