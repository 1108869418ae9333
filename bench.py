"""Benchmark: the spot-finder hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chunk C]

Workload: a series of 1800 device-resident
4148 x 4362 (W x H) uint16 frames, Poisson background 0.3, ~1500 Gaussian
spots a frame (sigma 0.8-2 px, 1-5 frames, peaks up to 20,000), masked
module gaps at the 514 x 1030 pitch and 0.05% hot pixels -- generated on the
device from a seed. A step is one chunk of C frames through the hot path:
the fused threshold with ordered compaction (one graph replay), then the
device 3-D labelling and centroid pass for the chunk. `value` is frames/s
over the timed steps with the inputs already in HBM.

Also reported: the fused kernel's roofline (algorithmic bytes per launch /
its HIP-event time on the plan's stream, against 8 TB/s), the plain
one-frame-per-launch form on the same frames, and the reference's own CPU
path (oracle/_ref, compiled from the reference sources) timed on this
host's cores on a bounded sample, whose lists are also compared with the
device's (mask_diffs).
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
KERNEL = "probe_kernel"  # the u16 threshold kernel a plan of form 0 launches
METRIC = ("frames/s + HBM GB/s (% peak), 4148×4362 uint16, 0 mask diffs; "
          "p99 latency")


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def cpu_baseline(sf, host_frames, device_lists, threads, seconds):
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
    }, diffs


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
    labeller = None
    if not args.no_label and hasattr(sf, "Labeller"):
        labeller = sf.Labeller(HEIGHT, WIDTH)

    # One captured graph (zero counters, threshold, ordering) replayed per
    # chunk; binding a chunk only rewrites the frame descriptors it reads.
    plan.capture(ptr[:chunk])

    def step(s):
        c = s % n_chunks
        plan.bind(ptr[c * chunk:(c + 1) * chunk])
        plan.replay()
        if labeller is not None:
            # frame numbers keep ascending as the series is replayed
            labeller.add_plan(plan, first_frame=s * chunk)

    for s in range(args.warmup):
        step(s)
    plan.wait()
    sf.lib().sf_hip_synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(args.warmup + s)
    plan.wait()
    if labeller is not None:
        labeller.wait()
    sf.lib().sf_hip_synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps * chunk / elapsed

    # The fused kernel alone, from HIP events on the plan's stream.
    kernel_ms, order_ms, counts = [], [], []
    for s in range(5):
        c = s % n_chunks
        plan.set_timing(True)
        plan.run(ptr[c * chunk:(c + 1) * chunk])
        plan.wait()
        a, b = plan.last_times()
        kernel_ms.append(a)
        order_ms.append(b)
        counts.append(plan.counts())
    plan.set_timing(False)
    kernel_s = float(np.median(kernel_ms)) / 1e3
    mean_signal = float(np.mean(np.concatenate(counts)))
    algorithmic = chunk * (frame_bytes + 16.0 * mean_signal)
    achieved = algorithmic / kernel_s
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "fused_threshold_traffic.json")
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
                                   args.cpu_seconds)

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
            "workload": "config 1: 1800-frame series of 4148x4362 uint16, "
                        "Poisson 0.3, ~1500 spots/frame, gaps, 0.05% hot",
            "frames_per_step": chunk,
            "series_frames": n_frames,
            "path": "fused threshold + ordered compaction"
                    + (" + 3-D labelling and centroids" if labeller else ""),
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
            "order_ms_per_launch": round(float(np.median(order_ms)), 4),
            "algorithmic_bytes_per_launch": int(algorithmic),
        },
        "plain_form_fps": round(plain_fps, 2),
        "speedup_vs_plain": round(fps / plain_fps, 3),
        "mean_signal_pixels_per_frame": round(mean_signal, 1),
        "mask_diffs": diffs,
        "cpu_baseline": baseline,
        "p99_latency_ms": latency,
        "serving": serving,
    }
    print(json.dumps(record))
    plan.close()
    series.close()
    return 0 if diffs == 0 else 3


if __name__ == "__main__":
    sys.exit(main())
