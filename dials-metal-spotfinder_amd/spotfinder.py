"""Python host side of the MI355X spot-finder hot path (ctypes over the C ABI).

Mirrors the reference's device interface (src/dext_gpu.hh, namespace gpu):
`find` is gpu::find<T> -- same arguments, same ascending output, same -1 / -2
return codes -- plus `backend`, `available`, `memory_free`, the window and
profiling knobs, and device-resident batches (`Plan`) and the device
labeller (`Labeller`, mirroring dials_spots::Labeller) that the reference
does not have.

Nothing here computes: every call goes to libdext_hip.so, and importing this
module on a machine without the library raises rather than falling back.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libdext_hip.so")

# Hardware queues for the process's streams (read once, when HIP starts). A
# pipelined series uses four streams of its own (two plans', the labeller's
# and its snapshots'); with HIP's default of four queues they share queues
# with any other stream of the process, and work on a shared queue waits for
# work it does not depend on. Eight queues measured 2-3% faster on the
# pipelined bench. A caller's own setting is kept.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

SF_OK = 0
SF_BAD_ARGUMENT = -1
SF_OVERFLOW = -2
SF_DEVICE_ERROR = -3
SF_ORDER_ERROR = -4
SF_IO_ERROR = -5
SF_BAD_DATA = -6

DIRECT, TILE = 0, 1

SIGNAL = np.dtype(
    [("index", "<u4"), ("value", "<u4"), ("background", "<f4"),
     ("population", "<u2"), ("reserved", "<u2")]
)
PIXEL = np.dtype([("frame", "<i4"), ("index", "<u4"), ("value", "<u4")])
SPOT = np.dtype(
    [("bbox", "<i4", (6,)), ("position", "<f8", (3,)),
     ("variance", "<f8", (3,)), ("intensity", "<f8"),
     ("intensity_variance", "<f8"), ("n_signal", "<u4"), ("_pad", "<u4"),
     ("first", "<u8")]
)
COUNTS = ("signal_pixels", "groups", "too_small", "too_large", "separated",
          "accepted")

_sz = ctypes.c_size_t
_p = ctypes.c_void_p
_u32 = ctypes.c_uint32
_i = ctypes.c_int


class DeviceError(RuntimeError):
    """The device failed (gpu::find throws std::runtime_error here)."""


class Options(ctypes.Structure):
    """dials_spots::Options (src/dials_spots.hh:81-97)."""

    _fields_ = [("min_spot_size", ctypes.c_uint64),
                ("max_spot_size", ctypes.c_uint64),
                ("max_separation", ctypes.c_double),
                ("two_d", ctypes.c_int32)]


class ReflOptions(ctypes.Structure):
    """include/refl_hip.h sf_refl_options (src/refl.hh:52-63)."""
    _fields_ = [("panel", ctypes.c_uint64), ("id", ctypes.c_int32),
                ("shoeboxes", ctypes.c_int32), ("identifier", ctypes.c_char_p)]


class DriverOptions(ctypes.Structure):
    """include/driver_hip.h sf_driver_options (src/find_spots.cc:70-90)."""
    _fields_ = [("chunk", ctypes.c_int32), ("z_offset", ctypes.c_int64),
                ("grouping", Options)]


class DriverStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "dispatched", "read", "missing", "failed", "grouped",
        "out_of_order")]


class ServerStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "batches", "batched_frames", "admission_waits",
        "max_batches_in_flight", "bytes_in_use", "peak_bytes", "budget",
        "overflowed", "ingested_bytes")]


def _declare(lib: ctypes.CDLL) -> None:
    def fn(name, restype, *argtypes):
        f = getattr(lib, name)
        f.restype = restype
        f.argtypes = list(argtypes)

    fn("sf_hip_backend", ctypes.c_char_p)
    fn("sf_hip_available", _i)
    fn("sf_hip_memory_free", _sz)
    fn("sf_hip_host_alloc", _p, _sz)
    fn("sf_hip_host_free", None, _p)
    fn("sf_hip_last_error", ctypes.c_char_p)
    fn("sf_hip_device_alloc", _p, _sz)
    fn("sf_hip_device_free", None, _p)
    fn("sf_hip_copy_to_device", _i, _p, _p, _sz)
    fn("sf_hip_copy_to_host", _i, _p, _p, _sz)
    fn("sf_hip_synchronize", _i)
    fn("sf_hip_stage0_window", _i)
    fn("sf_hip_set_stage0_window", None, _i)
    fn("sf_hip_stage2_window", _i)
    fn("sf_hip_set_stage2_window", None, _i)
    fn("sf_hip_profile_stages", None, _i)
    fn("sf_hip_profiling", _i)
    fn("sf_hip_last_stage_times", None, ctypes.POINTER(ctypes.c_double))
    fn("sf_hip_report_windows", None, _i)
    for name in ("sf_hip_find_u16", "sf_hip_find_u32"):
        fn(name, _i, _p, _sz, _sz, _p, _sz, ctypes.POINTER(_sz))
    fn("sf_hip_take_found", _i, _p, _sz)
    fn("sf_hip_plan_create", _p, _i, _i, ctypes.POINTER(_u32),
       ctypes.POINTER(_u32), _i)
    fn("sf_hip_plan_destroy", None, _p)
    fn("sf_hip_plan_set_form", _i, _p, _i)
    fn("sf_hip_plan_set_grid", _i, _p, _i)
    fn("sf_hip_plan_autotune", _i, _p, ctypes.POINTER(_p), _i,
       ctypes.POINTER(_i), ctypes.POINTER(_i))
    fn("sf_hip_plan_run", _i, _p, ctypes.POINTER(_p))
    fn("sf_hip_plan_set_deferred_order", _i, _p, _i)
    fn("sf_hip_plan_order", _i, _p)
    fn("sf_hip_plan_bind", _i, _p, ctypes.POINTER(_p))
    fn("sf_hip_plan_capture", _i, _p, ctypes.POINTER(_p))
    fn("sf_hip_plan_replay", _i, _p)
    fn("sf_hip_plan_wait", _i, _p)
    fn("sf_hip_plan_counts", _i, _p, ctypes.POINTER(_u32))
    fn("sf_hip_plan_output", _p, _p, _i)
    fn("sf_hip_plan_capacity", _u32, _p, _i)
    fn("sf_hip_plan_set_timing", _i, _p, _i)
    fn("sf_hip_plan_last_times", _i, _p, ctypes.POINTER(ctypes.c_double),
       ctypes.POINTER(ctypes.c_double))
    fn("sf_hip_plan_last_pass_times", _i, _p,
       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
       ctypes.POINTER(ctypes.c_double))
    fn("sf_hip_plan_stream", _p, _p)
    for name in ("sf_hip_plan_pipeline", "sf_hip_plan_capture_pipeline",
                 "sf_hip_plan_replay_pipeline"):
        fn(name, _i, _p, _p)
    fn("sf_hip_synth_frames", _i, _p, _i, _u32, _u32, _u32, ctypes.c_uint64,
       ctypes.c_double, _i, ctypes.c_double, _p, _u32, ctypes.c_double)
    fn("sf_hip_labeller_create", _p, _sz, _sz, ctypes.POINTER(Options))
    fn("sf_hip_labeller_destroy", None, _p)
    fn("sf_hip_labeller_add", _i, _p, ctypes.c_int64, _p, _sz)
    fn("sf_hip_labeller_add_plan", _i, _p, _p, ctypes.c_int64)
    fn("sf_hip_labeller_add_lists", _i, _p, _i,
       ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_p),
       ctypes.POINTER(_p), ctypes.POINTER(_u32), ctypes.POINTER(_u32), _p)
    fn("sf_hip_labeller_flush", _i, _p)
    fn("sf_hip_labeller_finish", _i, _p)
    fn("sf_hip_labeller_wait", _i, _p)
    fn("sf_hip_labeller_live", _sz, _p)
    fn("sf_hip_labeller_sizes", _i, _p, ctypes.POINTER(_sz),
       ctypes.POINTER(_sz))
    fn("sf_hip_labeller_reserve", _i, _p, _sz, _sz)
    fn("sf_hip_labeller_spot_count", _sz, _p)
    fn("sf_hip_labeller_pixel_count", _sz, _p)
    fn("sf_hip_labeller_results", _i, _p, _p, _p, _p)
    fn("sf_hip_labeller_last_error", ctypes.c_char_p)
    fn("sf_hip_server_slot_bytes", _sz, _i, _sz, _sz)
    fn("sf_hip_server_create", _p, _sz, _i, _i)
    fn("sf_hip_server_destroy", None, _p)
    fn("sf_hip_server_open", _i, _p, _i, _sz, _sz, ctypes.POINTER(Options))
    fn("sf_hip_server_open_host", _i, _p, _i, _sz, _sz,
       ctypes.POINTER(Options))
    fn("sf_hip_server_submit", _i, _p, _i, ctypes.c_int64, _p)
    fn("sf_hip_server_submit_at", _i, _p, _i, ctypes.c_int64, _p,
       ctypes.c_double)
    fn("sf_hip_server_clock", ctypes.c_double)
    fn("sf_hip_server_drain", _i, _p)
    fn("sf_hip_server_close", _i, _p, _i)
    fn("sf_hip_server_labeller", _p, _p, _i)
    fn("sf_hip_server_frames", _i, _p, _i, _p, _p, _p, _p, _sz,
       ctypes.POINTER(_sz))
    fn("sf_hip_server_frame_list", _i, _p, _i, _sz, _p, _sz,
       ctypes.POINTER(_sz))
    fn("sf_hip_server_stats", _i, _p, ctypes.POINTER(ServerStats))
    fn("sf_hip_server_play", _i, _p, _i, ctypes.POINTER(_i),
       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
       ctypes.POINTER(_p), ctypes.POINTER(_i), ctypes.c_int64)
    fn("sf_hip_server_last_error", ctypes.c_char_p)
    fn("sf_hip_refl_write", _i, ctypes.c_char_p, _p, _sz, _p, _sz, _sz,
       ctypes.POINTER(ReflOptions))
    fn("sf_hip_stamp_mask", _i, _i, _p, _i, _sz, _sz, _p, _p)
    fn("sf_hip_pack_mask", _i, _p, _sz, _sz, _p, _p)
    fn("sf_hip_mask_last_error", ctypes.c_char_p)
    fn("sf_hip_refl_last_error", ctypes.c_char_p)
    fn("sf_hip_decoder_create", _i, ctypes.c_uint, _sz, _sz, _i,
       ctypes.POINTER(_p))
    fn("sf_hip_decoder_index", _i, _p, _p, _p, _i, ctypes.c_char_p)
    fn("sf_hip_decoder_decode", _i, _p, _p, _p)
    fn("sf_hip_decoder_wait", _i, _p)
    fn("sf_hip_decoder_image", _i, _p, _p, _sz, ctypes.c_char_p, _p)
    fn("sf_hip_decoder_stream", _p, _p)
    fn("sf_hip_decoder_set_timing", _i, _p, _i)
    fn("sf_hip_decoder_last_ms", ctypes.c_double, _p)
    fn("sf_hip_decoder_destroy", None, _p)
    fn("sf_hip_decode_last_error", ctypes.c_char_p)
    fn("sf_hip_driver_create", _p, ctypes.c_uint, _sz, _sz,
       ctypes.POINTER(DriverOptions))
    fn("sf_hip_driver_submit", _i, _p, ctypes.c_int64, _p, _sz,
       ctypes.c_char_p)
    fn("sf_hip_driver_finish", _i, _p)
    fn("sf_hip_driver_set_mask", _i, _p, _p, _sz, _sz)
    fn("sf_hip_driver_write", _i, _p, ctypes.c_char_p,
       ctypes.POINTER(ReflOptions))
    fn("sf_hip_driver_log", _sz, _p, _p, _sz)
    fn("sf_hip_driver_stats", _i, _p, ctypes.POINTER(DriverStats))
    fn("sf_hip_driver_labeller", _p, _p)
    fn("sf_hip_driver_destroy", None, _p)
    fn("sf_hip_driver_last_error", ctypes.c_char_p)
    fn("sf_bslz4_bound", _sz, _sz, _sz, _sz)
    fn("sf_bslz4_compress", _i, _p, _sz, _sz, _sz, _p, _sz,
       ctypes.POINTER(_sz))
    fn("sf_lz4_compress_block", _sz, _p, _sz, _p, _sz)


_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """The product library. Raises if it is not built: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(
                f"{LIBRARY} is missing; build it with "
                "dials-metal-spotfinder_amd/build.py")
        handle = ctypes.CDLL(LIBRARY)
        _declare(handle)
        _lib = handle
    return _lib


def last_error() -> str:
    return lib().sf_hip_last_error().decode()


def _check(status: int, what: str) -> None:
    if status == SF_DEVICE_ERROR:
        raise DeviceError(f"{what}: {last_error()}")
    if status != SF_OK:
        raise ValueError(f"{what}: status {status}: {last_error()}")


# --- src/dext_gpu.hh:28-139 -------------------------------------------------

def backend() -> str:
    return lib().sf_hip_backend().decode()


def available() -> bool:
    return bool(lib().sf_hip_available())


def memory_free() -> int:
    return int(lib().sf_hip_memory_free())


def stage0_window(window: int | None = None) -> int:
    if window is not None:
        lib().sf_hip_set_stage0_window(window)
    return int(lib().sf_hip_stage0_window())


def stage2_window(window: int | None = None) -> int:
    if window is not None:
        lib().sf_hip_set_stage2_window(window)
    return int(lib().sf_hip_stage2_window())


def profile_stages(on: bool | None = None) -> bool:
    if on is not None:
        lib().sf_hip_profile_stages(1 if on else 0)
    return bool(lib().sf_hip_profiling())


def last_stage_times() -> dict:
    times = (ctypes.c_double * 6)()
    lib().sf_hip_last_stage_times(times)
    return dict(zip(("upload", "stage0", "stage1", "stage2", "copy", "sort"),
                    list(times)))


def report_windows(on: bool) -> None:
    lib().sf_hip_report_windows(1 if on else 0)


def find(image: np.ndarray):
    """gpu::find<T> (src/dext_gpu.hh:156-158): (status, ascending list).

    status is 0, -1 (bad argument) or -2 (more signal pixels than
    height*width*sizeof(T)/16); a device failure raises DeviceError.
    """
    if image is None or image.ndim != 2 or image.dtype not in (np.uint16,
                                                               np.uint32):
        return SF_BAD_ARGUMENT, np.zeros(0, dtype=SIGNAL)
    image = np.ascontiguousarray(image)
    height, width = image.shape
    fn = lib().sf_hip_find_u16 if image.dtype == np.uint16 else lib().sf_hip_find_u32
    capacity = 4096
    out = np.zeros(capacity, dtype=SIGNAL)
    found = _sz(0)
    status = fn(image.ctypes.data if image.size else None, height, width,
                out.ctypes.data, capacity, ctypes.byref(found))
    if status == SF_DEVICE_ERROR:
        raise DeviceError(last_error())
    if status != SF_OK:
        return status, np.zeros(0, dtype=SIGNAL)
    n = found.value
    if n > capacity:
        out = np.zeros(n, dtype=SIGNAL)
        _check(lib().sf_hip_take_found(out.ctypes.data, n), "take_found")
        return SF_OK, out
    return SF_OK, out[:n].copy()


# --- device memory ----------------------------------------------------------

class DeviceBuffer:
    """A device allocation (hipMalloc), freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().sf_hip_device_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise DeviceError(f"device allocation of {nbytes} bytes: "
                              f"{last_error()}")

    def upload(self, array: np.ndarray, offset: int = 0) -> None:
        array = np.ascontiguousarray(array)
        assert offset + array.nbytes <= self.nbytes
        _check(lib().sf_hip_copy_to_device(self.ptr + offset,
                                           array.ctypes.data, array.nbytes),
               "copy to device")

    def download(self, array: np.ndarray, offset: int = 0) -> np.ndarray:
        assert array.flags.c_contiguous
        assert offset + array.nbytes <= self.nbytes
        _check(lib().sf_hip_copy_to_host(array.ctypes.data, self.ptr + offset,
                                         array.nbytes), "copy to host")
        return array

    def close(self) -> None:
        if self.ptr:
            lib().sf_hip_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory (hipHostMalloc), freed on close()/garbage
    collection; `array(dtype, shape)` views it."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().sf_hip_host_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise DeviceError(f"pinned allocation of {nbytes} bytes: "
                              f"{last_error()}")

    def array(self, dtype, shape, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = int(np.prod(shape))
        assert offset + count * dtype.itemsize <= self.nbytes
        raw = (ctypes.c_uint8 * (count * dtype.itemsize)).from_address(
            self.ptr + offset)
        return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape)

    def close(self) -> None:
        if self.ptr:
            lib().sf_hip_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def download_signal(device_ptr: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=SIGNAL)
    if n:
        _check(lib().sf_hip_copy_to_host(out.ctypes.data, device_ptr,
                                         out.nbytes), "copy to host")
    return out


def synth_frames(buffer: DeviceBuffer, pixel_bytes: int, height: int,
                 width: int, frames: int, seed: int, background: float,
                 gaps: bool, hot_fraction: float, spots: np.ndarray,
                 sigma_z: float = 0.7, offset: int = 0) -> None:
    """Fill device frames with the synthetic value model (see synth.hip)."""
    spots = np.ascontiguousarray(spots, dtype=np.float64).reshape(-1, 6)
    need = pixel_bytes * height * width * frames
    assert offset + need <= buffer.nbytes
    _check(lib().sf_hip_synth_frames(
        buffer.ptr + offset, pixel_bytes, height, width, frames, seed,
        background, 1 if gaps else 0, hot_fraction,
        spots.ctypes.data if len(spots) else None, len(spots), sigma_z),
        "synth_frames")


# --- device-resident batches -------------------------------------------------

class Plan:
    """Fixed frame shapes thresholded together on the device.

    form 0 is the fused default (the sparse passes; the register sweep for
    32-bit frames of 2^29 pixels or more), form 1 the plain three-kernel
    per-frame form, form 2 the register sweep. Frames are device pointers
    (ints).
    """

    def __init__(self, pixel_bytes: int, shapes, band_rows: int = 0,
                 form: int = 0):
        self.shapes = [(int(h), int(w)) for h, w in shapes]
        self.pixel_bytes = pixel_bytes
        n = len(self.shapes)
        heights = (_u32 * n)(*[h for h, _ in self.shapes])
        widths = (_u32 * n)(*[w for _, w in self.shapes])
        self.handle = lib().sf_hip_plan_create(pixel_bytes, n, heights,
                                               widths, band_rows)
        if not self.handle:
            raise DeviceError(f"plan: {last_error()}")
        if form:
            self.set_form(form)

    def _frames(self, frames):
        assert len(frames) == len(self.shapes)
        return (_p * len(frames))(*[int(f) for f in frames])

    def set_form(self, form: int) -> None:
        _check(lib().sf_hip_plan_set_form(self.handle, form), "set_form")

    def set_grid(self, blocks: int) -> None:
        _check(lib().sf_hip_plan_set_grid(self.handle, blocks), "set_grid")

    def autotune(self, frames, trials: int = 3):
        """Time the candidate band heights / grids on these frames and keep
        the fastest; returns (band_rows, blocks). Capture again afterwards."""
        band, blocks = _i(0), _i(0)
        _check(lib().sf_hip_plan_autotune(self.handle, self._frames(frames),
                                          trials, ctypes.byref(band),
                                          ctypes.byref(blocks)), "autotune")
        return band.value, blocks.value

    def set_timing(self, on: bool) -> None:
        _check(lib().sf_hip_plan_set_timing(self.handle, 1 if on else 0),
               "set_timing")

    def run(self, frames) -> None:
        _check(lib().sf_hip_plan_run(self.handle, self._frames(frames)), "run")

    def bind(self, frames) -> None:
        _check(lib().sf_hip_plan_bind(self.handle, self._frames(frames)),
               "bind")

    def capture(self, frames) -> None:
        _check(lib().sf_hip_plan_capture(self.handle, self._frames(frames)),
               "capture")

    def replay(self) -> None:
        _check(lib().sf_hip_plan_replay(self.handle), "replay")

    def set_deferred_order(self, deferred: bool) -> None:
        """Pipelined steps that finish this plan leave its ordering to
        order() (sf_hip_plan_set_deferred_order)."""
        _check(lib().sf_hip_plan_set_deferred_order(self.handle,
                                                     1 if deferred else 0),
               "set_deferred_order")

    def order(self) -> None:
        """The ordering a deferred pipelined step left pending, enqueued on
        the plan's stream (sf_hip_plan_order)."""
        _check(lib().sf_hip_plan_order(self.handle), "order")

    def wait(self) -> None:
        _check(lib().sf_hip_plan_wait(self.handle), "wait")

    def counts(self) -> np.ndarray:
        out = (_u32 * len(self.shapes))()
        _check(lib().sf_hip_plan_counts(self.handle, out), "counts")
        return np.array(list(out), dtype=np.int64)

    def capacity(self, frame: int) -> int:
        return int(lib().sf_hip_plan_capacity(self.handle, frame))

    def output_ptr(self, frame: int) -> int:
        return int(lib().sf_hip_plan_output(self.handle, frame))

    def output(self, frame: int, count: int | None = None) -> np.ndarray:
        if count is None:
            count = int(self.counts()[frame])
        count = min(count, self.capacity(frame))
        return download_signal(self.output_ptr(frame), count)

    def last_times(self):
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().sf_hip_plan_last_times(self.handle, ctypes.byref(a),
                                            ctypes.byref(b)), "last_times")
        return a.value, b.value

    def last_pass_times(self):
        """(first threshold kernel, the threshold's other kernels, ordering)
        of the last timed run, in ms."""
        a, b, c = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().sf_hip_plan_last_pass_times(
            self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
            "last_pass_times")
        return a.value, b.value, c.value

    def close(self) -> None:
        if self.handle:
            lib().sf_hip_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _handle(plan):
    return plan.handle if plan is not None else None


def pipeline(current: "Plan | None", next_plan: "Plan | None") -> None:
    """One pipelined step (sf_hip_plan_pipeline): finish `current` (scouted
    by the step before) while scouting `next_plan` (frames bound); either
    may be None."""
    _check(lib().sf_hip_plan_pipeline(_handle(current), _handle(next_plan)),
           "pipeline")


def capture_pipeline(current: "Plan | None", next_plan: "Plan | None") -> None:
    _check(lib().sf_hip_plan_capture_pipeline(_handle(current),
                                              _handle(next_plan)),
           "capture_pipeline")


def replay_pipeline(current: "Plan | None", next_plan: "Plan | None") -> None:
    _check(lib().sf_hip_plan_replay_pipeline(_handle(current),
                                             _handle(next_plan)),
           "replay_pipeline")


# --- 3-D grouping and centroids (dials_spots::Labeller) ----------------------

class Labeller:
    """dials_spots::Labeller on the device (src/dials_spots.hh:108-158).

    add() raises ValueError where the reference's add() throws; spots(),
    pixels() and counts() match the reference's row for row after finish().
    """

    def __init__(self, height: int, width: int, min_spot_size: int = 3,
                 max_spot_size: int = 1000, max_separation: float = 2.0,
                 two_d: bool = False):
        options = Options(min_spot_size, max_spot_size, float(max_separation),
                          1 if two_d else 0)
        self.handle = lib().sf_hip_labeller_create(height, width,
                                                   ctypes.byref(options))
        if not self.handle:
            raise ValueError(lib().sf_hip_labeller_last_error().decode())

    def _status(self, status: int, what: str) -> None:
        if status == SF_OK:
            return
        message = lib().sf_hip_labeller_last_error().decode()
        if status == SF_DEVICE_ERROR:
            raise DeviceError(f"{what}: {message}")
        raise ValueError(f"{what}: {message}")

    def add(self, frame: int, signal: np.ndarray) -> None:
        signal = np.ascontiguousarray(signal, dtype=SIGNAL)
        self._status(lib().sf_hip_labeller_add(
            self.handle, int(frame),
            signal.ctypes.data if len(signal) else None, len(signal)), "add")

    def add_plan(self, plan: "Plan", first_frame: int) -> None:
        """Add every frame of the plan's last run as consecutive frames from
        first_frame, straight from the device lists. Returns at once: the
        lists are gathered on the labeller's stream (by the ordering's
        places, where the plan's ordering is pending) and labelled by a
        worker thread; a failure there is raised by the next call."""
        self._status(lib().sf_hip_labeller_add_plan(
            self.handle, plan.handle, int(first_frame)), "add_plan")

    def add_lists(self, frames, lists, count_ptrs, capacities, counts,
                  stream=None) -> None:
        """Add frames whose ordered lists are in device memory: frame
        numbers, device pointers of the lists and of their counts, the
        lists' capacities and the counts as the host knows them."""
        n = len(frames)
        if not (len(lists) == len(count_ptrs) == len(capacities) ==
                len(counts) == n):
            raise ValueError("add_lists: one entry a frame in every argument")
        self._status(lib().sf_hip_labeller_add_lists(
            self.handle, n, (ctypes.c_int64 * n)(*frames),
            (_p * n)(*lists), (_p * n)(*count_ptrs),
            (_u32 * n)(*capacities), (_u32 * n)(*counts), stream),
            "add_lists")

    def flush(self) -> None:
        self._status(lib().sf_hip_labeller_flush(self.handle), "flush")

    def finish(self) -> None:
        self._status(lib().sf_hip_labeller_finish(self.handle), "finish")

    def wait(self) -> None:
        self._status(lib().sf_hip_labeller_wait(self.handle), "wait")

    def live(self) -> int:
        return int(lib().sf_hip_labeller_live(self.handle))

    def sizes(self):
        """(spot rows, pixel rows) so far."""
        n_spots, n_pixels = _sz(), _sz()
        self._status(lib().sf_hip_labeller_sizes(
            self.handle, ctypes.byref(n_spots), ctypes.byref(n_pixels)),
            "sizes")
        return int(n_spots.value), int(n_pixels.value)

    def reserve(self, spots: int, pixels: int) -> None:
        """Room for `spots` more spot rows and `pixels` more pixel rows,
        allocated now, so that the chunks that follow allocate nothing (an
        extension: the reference's tables are host vectors)."""
        self._status(lib().sf_hip_labeller_reserve(
            self.handle, int(spots), int(pixels)), "reserve")

    def results(self):
        """(spots, pixels, counts dict), as the reference's spots(),
        pixels() and counts()."""
        return _labeller_results(self.handle, self._status)

    def close(self) -> None:
        if self.handle:
            lib().sf_hip_labeller_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _labeller_results(handle, status):
    n_spots, n_pixels = _sz(0), _sz(0)
    status(lib().sf_hip_labeller_sizes(handle, ctypes.byref(n_spots),
                                       ctypes.byref(n_pixels)), "sizes")
    spots = np.zeros(n_spots.value, dtype=SPOT)
    pixels = np.zeros(n_pixels.value, dtype=PIXEL)
    counts = (ctypes.c_uint64 * 6)()
    status(lib().sf_hip_labeller_results(
        handle, spots.ctypes.data if len(spots) else None,
        pixels.ctypes.data if len(pixels) else None, counts), "results")
    return spots, pixels, dict(zip(COUNTS, [int(c) for c in counts]))


# --- the serving loop ---------------------------------------------------------

def server_slot_bytes(pixel_bytes: int, height: int, width: int) -> int:
    """Bytes of budget one frame of this shape takes in a Server."""
    n = int(lib().sf_hip_server_slot_bytes(pixel_bytes, height, width))
    if n == 0:
        raise ValueError("server_slot_bytes: bad shape or pixel type")
    return n


class Server:
    """Frames of several concurrent series, batched as they arrive, grouped
    per series on the device, admitted by an HBM budget (see serve.hip)."""

    def __init__(self, hbm_budget: int, max_batch: int = 16,
                 keep_lists: bool = False):
        self.handle = lib().sf_hip_server_create(int(hbm_budget),
                                                 int(max_batch),
                                                 1 if keep_lists else 0)
        if not self.handle:
            raise ValueError(self._error())

    @staticmethod
    def _error() -> str:
        return lib().sf_hip_server_last_error().decode()

    def _status(self, status: int, what: str) -> None:
        if status == SF_OK:
            return
        if status == SF_DEVICE_ERROR:
            raise DeviceError(f"{what}: {self._error()}")
        raise ValueError(f"{what}: status {status}: {self._error()}")

    def open(self, pixel_bytes: int, height: int, width: int,
             min_spot_size: int = 3, max_spot_size: int = 1000,
             max_separation: float = 2.0, two_d: bool = False,
             host: bool = False) -> int:
        """Open a series. host: its frames are submitted as pointers to
        (pinned) host memory and copied in by the serving loop."""
        options = Options(min_spot_size, max_spot_size, float(max_separation),
                          1 if two_d else 0)
        opener = (lib().sf_hip_server_open_host if host
                  else lib().sf_hip_server_open)
        sid = opener(self.handle, pixel_bytes, height, width,
                     ctypes.byref(options))
        if sid < 0:
            self._status(sid, "open")
        return sid

    def submit(self, series: int, frame: int, device_ptr: int,
               arrival: float | None = None) -> None:
        if arrival is None:
            status = lib().sf_hip_server_submit(self.handle, series,
                                                int(frame), int(device_ptr))
        else:
            status = lib().sf_hip_server_submit_at(
                self.handle, series, int(frame), int(device_ptr),
                float(arrival))
        self._status(status, "submit")

    @staticmethod
    def clock() -> float:
        return float(lib().sf_hip_server_clock())

    def drain(self) -> None:
        self._status(lib().sf_hip_server_drain(self.handle), "drain")

    def close_series(self, series: int) -> None:
        self._status(lib().sf_hip_server_close(self.handle, series), "close")

    def results(self, series: int):
        """The series' labeller tables (spots, pixels, counts dict)."""
        handle = lib().sf_hip_server_labeller(self.handle, series)
        if not handle:
            raise ValueError("no such series")

        def status(code, what):
            if code != SF_OK:
                raise ValueError(
                    f"{what}: {lib().sf_hip_labeller_last_error().decode()}")
        return _labeller_results(handle, status)

    def frames(self, series: int) -> dict:
        """Per completed frame: number, count, mask_ms and table_ms."""
        n = _sz(0)
        self._status(lib().sf_hip_server_frames(
            self.handle, series, None, None, None, None, 0, ctypes.byref(n)),
            "frames")
        k = n.value
        numbers = np.zeros(k, dtype=np.int64)
        counts = np.zeros(k, dtype=np.uint32)
        mask = np.zeros(k, dtype=np.float64)
        table = np.zeros(k, dtype=np.float64)
        if k:
            self._status(lib().sf_hip_server_frames(
                self.handle, series, numbers.ctypes.data, counts.ctypes.data,
                mask.ctypes.data, table.ctypes.data, k, ctypes.byref(n)),
                "frames")
        return {"frame": numbers, "count": counts, "mask_ms": mask,
                "table_ms": table}

    def frame_list(self, series: int, k: int) -> np.ndarray:
        n = _sz(0)
        self._status(lib().sf_hip_server_frame_list(
            self.handle, series, k, None, 0, ctypes.byref(n)), "frame_list")
        out = np.zeros(n.value, dtype=SIGNAL)
        if n.value:
            self._status(lib().sf_hip_server_frame_list(
                self.handle, series, k, out.ctypes.data, n.value,
                ctypes.byref(n)), "frame_list")
        return out

    def stats(self) -> dict:
        st = ServerStats()
        self._status(lib().sf_hip_server_stats(self.handle, ctypes.byref(st)),
                     "stats")
        return {name: int(getattr(st, name)) for name, _ in st._fields_}

    def play(self, series, rates_hz, n_frames, rings, first_frame: int = 0):
        """Arrival schedule: series[i] gets n_frames[i] frames at rates_hz[i],
        cycling through the device pointers rings[i]; then drains."""
        m = len(series)
        flat = [int(p) for ring in rings for p in ring]
        self._status(lib().sf_hip_server_play(
            self.handle, m, (_i * m)(*series),
            (ctypes.c_double * m)(*[float(r) for r in rates_hz]),
            (ctypes.c_int64 * m)(*[int(n) for n in n_frames]),
            (_p * len(flat))(*flat), (_i * m)(*[len(r) for r in rings]),
            int(first_frame)), "play")

    def close(self) -> None:
        if self.handle:
            lib().sf_hip_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- src/refl.hh:21-78 --------------------------------------------------------

def write_refl(path: str, spots: np.ndarray, pixels: np.ndarray, width: int,
               panel: int = 0, id: int = 0, identifier: str = "",
               shoeboxes: bool = True) -> None:
    """refl::write: the Labeller's spot and pixel tables (as results() returns
    them) to a DIALS reflection table. Raises ValueError on a bad table and
    OSError when the file cannot be written."""
    spots = np.ascontiguousarray(spots, dtype=SPOT)
    pixels = np.ascontiguousarray(pixels, dtype=PIXEL)
    options = ReflOptions(panel, id, int(bool(shoeboxes)),
                          identifier.encode() if identifier else None)
    status = lib().sf_hip_refl_write(
        os.fsencode(path), spots.ctypes.data, len(spots), pixels.ctypes.data,
        len(pixels), int(width), ctypes.byref(options))
    if status == SF_OK:
        return
    message = lib().sf_hip_refl_last_error().decode()
    if status == SF_IO_ERROR:
        raise OSError(message)
    raise ValueError(message)


# --- the detector's pixel mask (the reference's README.md:227-234) ------------

def _mask_status(status: int, what: str) -> None:
    if status != SF_OK:
        message = lib().sf_hip_mask_last_error().decode()
        if status == SF_DEVICE_ERROR:
            raise DeviceError(f"{what}: {message}")
        raise ValueError(f"{what}: {message}")


class PixelMask:
    """A detector pixel mask (nonzero = masked, as NXmx's pixel_mask) as the
    device bit array sf_hip_stamp_mask reads."""

    def __init__(self, mask: np.ndarray):
        mask = np.ascontiguousarray(mask)
        if mask.ndim != 2 or mask.size == 0:
            raise ValueError("PixelMask: a non-empty 2-D mask")
        self.height, self.width = mask.shape
        packed = np.packbits(mask.reshape(-1) != 0, bitorder="little")
        words = (mask.size + 31) // 32
        padded = np.zeros(4 * words, dtype=np.uint8)
        padded[:len(packed)] = packed
        self.bits = DeviceBuffer(4 * words)
        self.bits.upload(padded)

    def stamp(self, pixel_bytes: int, frames, stream: int = 0) -> None:
        """Writes the masked sentinel max(T) into each device frame
        (pointers), asynchronously on `stream`; then synchronise."""
        ptrs = [int(f) for f in frames]
        _mask_status(lib().sf_hip_stamp_mask(
            pixel_bytes, (_p * len(ptrs))(*ptrs), len(ptrs), self.height,
            self.width, self.bits.ptr, stream or None), "stamp")

    def close(self) -> None:
        self.bits.close()


def pack_mask(mask_device: int, height: int, width: int, bits_device: int,
              stream: int = 0) -> None:
    """sf_hip_pack_mask: a device byte mask to the device bit array."""
    _mask_status(lib().sf_hip_pack_mask(mask_device, height, width,
                                        bits_device, stream or None), "pack")


class DecodeError(RuntimeError):
    """A chunk that does not decode to its frame: decompress::image's
    std::runtime_error, with its message."""


def _decode_status(status: int, what: str) -> None:
    if status != SF_OK:
        message = lib().sf_hip_decode_last_error().decode()
        if status == SF_DEVICE_ERROR:
            raise DeviceError(f"{what}: {message}")
        if status == SF_BAD_DATA:
            raise DecodeError(message)
        raise ValueError(f"{what}: {message}")


def bslz4_compress(elements: np.ndarray, block_size: int = 0) -> bytes:
    """The HDF5 bitshuffle filter's chunk for these elements (host code):
    12-byte header, bitshuffled LZ4 blocks, the last n % 8 elements raw."""
    elements = np.ascontiguousarray(elements)
    n, es = elements.size, elements.itemsize
    cap = lib().sf_bslz4_bound(n, es, block_size)
    out = np.empty(cap, dtype=np.uint8)
    written = _sz(0)
    _decode_status(lib().sf_bslz4_compress(
        elements.ctypes.data, n, es, block_size, out.ctypes.data, cap,
        ctypes.byref(written)), "bslz4_compress")
    return out[:written.value].tobytes()


def lz4_compress_block(data: bytes) -> bytes:
    """One LZ4 block (the published block format), host code."""
    src = np.frombuffer(data, dtype=np.uint8)
    cap = len(data) + len(data) // 255 + 16
    out = np.empty(cap, dtype=np.uint8)
    got = lib().sf_lz4_compress_block(src.ctypes.data if len(data) else None,
                                      len(data), out.ctypes.data, cap)
    if got == 0:
        raise ValueError("lz4_compress_block: output too small")
    return out[:got].tobytes()


class Decoder:
    """HDF5 chunks ("bslz4", "lz4" or "" stored) decoded into device frames
    of height x width elements of bit_depth bits: decompress::image
    (src/decompress.cc:76-136) on the device, a batch a launch."""

    def __init__(self, bit_depth: int, height: int, width: int,
                 max_frames: int = 64):
        handle = _p()
        _decode_status(lib().sf_hip_decoder_create(
            bit_depth, height, width, max_frames, ctypes.byref(handle)),
            "Decoder")
        self.handle = handle.value
        self.frame_bytes = height * width * bit_depth // 8
        self._held = None

    def index(self, chunks, algorithm: str = "bslz4") -> None:
        """Walks the chunks (bytes in host memory) and uploads their block
        table; raises DecodeError with the reference's message."""
        chunks = [np.frombuffer(c if len(c) else b"\0", dtype=np.uint8)[
            :len(c)] for c in chunks]
        self._held = chunks  # the walk reads them; keep them alive
        ptrs = (_p * len(chunks))(*[c.ctypes.data for c in chunks])
        sizes = (_sz * len(chunks))(*[len(c) for c in chunks])
        _decode_status(lib().sf_hip_decoder_index(
            self.handle, ptrs, sizes, len(chunks), algorithm.encode()),
            "index")

    def decode(self, device_chunks, device_frames) -> None:
        """Decodes the indexed chunks, held at the device pointers
        device_chunks, into device_frames; asynchronous."""
        c = [int(x) for x in device_chunks]
        f = [int(x) for x in device_frames]
        _decode_status(lib().sf_hip_decoder_decode(
            self.handle, (_p * len(c))(*c), (_p * len(f))(*f)), "decode")

    def wait(self) -> None:
        _decode_status(lib().sf_hip_decoder_wait(self.handle), "wait")

    def image(self, data: bytes, algorithm: str, device_frame: int) -> None:
        """decompress::image into one device frame (synchronous)."""
        src = np.frombuffer(data if len(data) else b"\0", dtype=np.uint8)
        _decode_status(lib().sf_hip_decoder_image(
            self.handle, src.ctypes.data, len(data), algorithm.encode(),
            int(device_frame)), "image")

    def set_timing(self, on: bool) -> None:
        _decode_status(lib().sf_hip_decoder_set_timing(self.handle, int(on)),
                       "set_timing")

    def last_ms(self) -> float:
        """Device time of the last decode (HIP events on its stream)."""
        return float(lib().sf_hip_decoder_last_ms(self.handle))

    @property
    def stream(self) -> int:
        return lib().sf_hip_decoder_stream(self.handle) or 0

    def close(self) -> None:
        if self.handle:
            lib().sf_hip_decoder_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Driver:
    """The series driver (src/find_spots.cc:460-713) on the device: frames
    submitted by number as stored chunks (None: never written), decoded,
    thresholded and grouped in number order a chunk at a time; finish()
    logs dials.find_spots' summary lines."""

    def __init__(self, bit_depth: int, height: int, width: int,
                 chunk: int = 0, z_offset: int = 0, min_spot_size: int = 3,
                 max_spot_size: int = 1000, max_separation: float = 2.0,
                 two_d: bool = False):
        options = DriverOptions(chunk, z_offset, Options(
            min_spot_size, max_spot_size, max_separation, int(two_d)))
        self.height, self.width = height, width
        self.handle = lib().sf_hip_driver_create(bit_depth, height, width,
                                                 ctypes.byref(options))
        if not self.handle:
            raise ValueError(lib().sf_hip_driver_last_error().decode())

    def _status(self, status: int, what: str) -> None:
        if status != SF_OK:
            message = lib().sf_hip_driver_last_error().decode()
            if status == SF_DEVICE_ERROR:
                raise DeviceError(f"{what}: {message}")
            raise ValueError(f"{what}: {message}")

    def submit(self, number: int, data: bytes | None,
               algorithm: str = "bslz4") -> None:
        if data is None:
            status = lib().sf_hip_driver_submit(self.handle, number, None, 0,
                                                None)
        else:
            src = np.frombuffer(data if len(data) else b"\0", np.uint8)
            status = lib().sf_hip_driver_submit(
                self.handle, number, src.ctypes.data, len(data),
                algorithm.encode())
        self._status(status, "submit")

    def set_mask(self, mask: "PixelMask | None") -> None:
        """Stamp this detector mask into every frame before the threshold.
        The mask must have the driver's frame shape (ValueError otherwise,
        before anything is launched) and stay open until finish(): the
        driver holds a reference to it."""
        if mask is not None and (mask.height, mask.width) != (self.height,
                                                                self.width):
            raise ValueError(
                f"set_mask: a {mask.height} x {mask.width} mask for "
                f"{self.height} x {self.width} frames")
        if mask is None:
            status = lib().sf_hip_driver_set_mask(self.handle, None, 0, 0)
        else:
            status = lib().sf_hip_driver_set_mask(
                self.handle, mask.bits.ptr, mask.height, mask.width)
        self._status(status, "set_mask")
        self._mask = mask

    def finish(self) -> None:
        self._status(lib().sf_hip_driver_finish(self.handle), "finish")

    def write(self, path: str, identifier: str = "",
              shoeboxes: bool = True) -> None:
        options = ReflOptions(0, 0, int(shoeboxes), identifier.encode())
        self._status(lib().sf_hip_driver_write(self.handle, path.encode(),
                                               ctypes.byref(options)),
                     "write")

    def log(self) -> str:
        n = lib().sf_hip_driver_log(self.handle, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib().sf_hip_driver_log(self.handle, buf, n + 1)
        return buf.value.decode()

    def stats(self) -> dict:
        out = DriverStats()
        self._status(lib().sf_hip_driver_stats(self.handle,
                                               ctypes.byref(out)), "stats")
        return {name: int(getattr(out, name)) for name, _ in out._fields_}

    def results(self):
        """(spots, pixels, counts dict) of the driver's labeller."""
        handle = lib().sf_hip_driver_labeller(self.handle)
        return _labeller_results(handle, self._status)

    def close(self) -> None:
        if self.handle:
            lib().sf_hip_driver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
