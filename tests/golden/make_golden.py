"""Write the golden fixtures from the reference's own CPU path.

    python tests/golden/make_golden.py

Runs oracle/_ref/libdext_ref.so -- the reference's src/dext.cc,
src/dials_spots.cc and src/signal_order.cc compiled where they lie by
oracle/Makefile -- over deterministic inputs and stores inputs' SHA-256 and
the outputs in tests/golden/*.npz. The inputs are regenerated by the tests
from the same seeds (their hashes guard against generator drift), so only
small outputs are committed. Needs the reference sources to build _ref; the
fixtures then travel without them.
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

from oracle import cpu  # noqa: E402
import frames as fr  # noqa: E402


def sha(array: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def threshold_fixtures() -> dict:
    out = {}
    cases = list(fr.reference_test_frames())
    cases.append(("wrap", fr.wrap_frame()))
    cases.append(("u32", fr.u32_frame()))
    for k, (name, frame) in enumerate(cases):
        status, signal, found = cpu.ref_dext(frame)
        assert status == 0 and found == len(signal)
        out[f"t{k}_name"] = np.array(name)
        out[f"t{k}_sha"] = np.array(sha(frame))
        out[f"t{k}_signal"] = signal
    out["n_threshold"] = np.array(len(cases))
    return out


def label(stack_frames, numbers, **options):
    lab = cpu.RefLabeller(stack_frames[0].shape[0], stack_frames[0].shape[1],
                          **options)
    total = 0
    for z, frame in zip(numbers, stack_frames):
        status, signal, found = cpu.ref_dext(frame)
        assert status == 0
        total += found
        assert lab.add(z, signal) == 0
    lab.finish()
    return lab.results()


def labeller_fixtures() -> dict:
    out = {}
    nx = fr.nxmx_fixture()
    series, _ = fr.planted_series()
    cases = [
        ("nxmx", nx, list(range(12)), {}),
        ("nxmx_2d", nx, list(range(12)), {"two_d": True}),
        ("nxmx_holey", np.delete(nx, 5, axis=0),
         [z for z in range(12) if z != 5], {}),
        ("planted", series, list(range(12)), {}),
        ("planted_nosep", series, list(range(12)),
         {"max_separation": 0.0, "min_spot_size": 1}),
    ]
    for k, (name, stack, numbers, options) in enumerate(cases):
        spots, pixels, counts = label(list(stack), numbers, **options)
        out[f"l{k}_name"] = np.array(name)
        out[f"l{k}_sha"] = np.array(sha(stack))
        out[f"l{k}_numbers"] = np.array(numbers, dtype=np.int64)
        out[f"l{k}_options"] = np.array(repr(sorted(options.items())))
        out[f"l{k}_spots"] = spots
        out[f"l{k}_pixels"] = pixels
        out[f"l{k}_counts"] = np.array([counts[c] for c in cpu.COUNTS],
                                       dtype=np.uint64)
    out["n_labeller"] = np.array(len(cases))
    return out


def main() -> None:
    if not cpu.have_ref():
        raise SystemExit("oracle/_ref is not built (needs /root/reference)")
    np.savez_compressed(os.path.join(HERE, "threshold.npz"),
                        **threshold_fixtures())
    np.savez_compressed(os.path.join(HERE, "labeller.npz"),
                        **labeller_fixtures())
    for name in ("threshold.npz", "labeller.npz"):
        path = os.path.join(HERE, name)
        print(name, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
