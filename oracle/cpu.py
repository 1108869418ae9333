"""ctypes access to the CPU checkers -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time: nothing in the product imports it.

    oracle   liboracle.so, the plain-C restatement of src/dext.cc,
             src/signal_order.cc and src/dials_spots.cc (dext_oracle.c) and
             of src/decompress.cc (bslz4_oracle.c)
    ref      _ref/libdext_ref.so, the reference's own sources compiled by
             oracle/Makefile (absent when the reference was not there to
             build; every user checks `have_ref()`)
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdext_ref.so")

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
assert SIGNAL.itemsize == 16 and PIXEL.itemsize == 12 and SPOT.itemsize == 104

_c_size = ctypes.c_size_t
_ptr = ctypes.c_void_p


def _address(array: np.ndarray) -> int:
    return array.ctypes.data


_oracle = None
_ref = None


def oracle_lib() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built")
        lib = ctypes.CDLL(ORACLE_SO)
        for name in ("oracle_dext_u16", "oracle_dext_u32",
                     "oracle_dext_direct_u16", "oracle_dext_direct_u32"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [_ptr, _c_size, _c_size, _ptr, _c_size,
                           ctypes.POINTER(_c_size)]
        lib.oracle_dext_many_u16.restype = ctypes.c_longlong
        lib.oracle_dext_many_u16.argtypes = [_ptr, _c_size, _c_size, _c_size,
                                             ctypes.c_int]
        lib.oracle_by_index.restype = None
        lib.oracle_by_index.argtypes = [_ptr, _c_size]
        lib.oracle_labeller_new.restype = _ptr
        lib.oracle_labeller_new.argtypes = [_c_size, _c_size, _c_size, _c_size,
                                            ctypes.c_double, ctypes.c_int]
        lib.oracle_labeller_free.argtypes = [_ptr]
        lib.oracle_labeller_add.restype = ctypes.c_int
        lib.oracle_labeller_add.argtypes = [_ptr, ctypes.c_int64, _ptr, _c_size]
        lib.oracle_labeller_finish.argtypes = [_ptr]
        lib.oracle_labeller_spot_count.restype = _c_size
        lib.oracle_labeller_spot_count.argtypes = [_ptr]
        lib.oracle_labeller_pixel_count.restype = _c_size
        lib.oracle_labeller_pixel_count.argtypes = [_ptr]
        lib.oracle_labeller_results.argtypes = [_ptr, _ptr, _ptr, _ptr]
        lib.oracle_lz4_block.restype = ctypes.c_long
        lib.oracle_lz4_block.argtypes = [_ptr, _c_size, _ptr, _c_size]
        for name in ("oracle_bitshuffle", "oracle_bitunshuffle"):
            fn = getattr(lib, name)
            fn.restype = None
            fn.argtypes = [_ptr, _c_size, _c_size, _ptr]
        lib.oracle_bshuf_default_block.restype = _c_size
        lib.oracle_bshuf_default_block.argtypes = [_c_size]
        lib.oracle_bshuf_decompress_lz4.restype = ctypes.c_longlong
        lib.oracle_bshuf_decompress_lz4.argtypes = [_ptr, _c_size, _ptr,
                                                    _c_size, _c_size, _c_size]
        lib.oracle_decompress_image.restype = ctypes.c_int
        lib.oracle_decompress_image.argtypes = [
            _ptr, _c_size, ctypes.c_char_p, ctypes.c_uint, _c_size, _c_size,
            _ptr, _c_size, ctypes.c_char_p, _c_size]
        lib.oracle_synthetic_frame.restype = None
        lib.oracle_synthetic_frame.argtypes = [_ptr, _c_size, _c_size,
                                               ctypes.c_uint64, ctypes.c_double,
                                               ctypes.c_int]
        _oracle = lib
    return _oracle


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref_lib() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError(f"{REF_SO} is not built")
        lib = ctypes.CDLL(REF_SO)
        for name in ("ref_dext_u16", "ref_dext_u32"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [_ptr, _c_size, _c_size, _ptr, _c_size,
                           ctypes.POINTER(_c_size)]
        for name in ("ref_dext_many_u16", "ref_dext_many_u32"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_longlong
            fn.argtypes = [_ptr, _c_size, _c_size, _c_size, ctypes.c_int]
        for name in ("ref_dext_cycle_u16", "ref_dext_cycle_u32"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_longlong
            fn.argtypes = [_ptr, _c_size, _c_size, _c_size, _c_size,
                           ctypes.c_int]
        lib.ref_pipeline_u16.restype = ctypes.c_int
        lib.ref_pipeline_u16.argtypes = [
            _ptr, _c_size, _c_size, _c_size, ctypes.c_int, ctypes.c_char_p,
            ctypes.POINTER(_c_size), ctypes.POINTER(_c_size),
            ctypes.POINTER(_c_size)]
        lib.ref_by_index.argtypes = [_ptr, _c_size]
        lib.ref_labeller_new.restype = _ptr
        lib.ref_labeller_new.argtypes = [_c_size, _c_size, _c_size, _c_size,
                                         ctypes.c_double, ctypes.c_int]
        lib.ref_labeller_free.argtypes = [_ptr]
        lib.ref_labeller_add.restype = ctypes.c_int
        lib.ref_labeller_add.argtypes = [_ptr, ctypes.c_int64, _ptr, _c_size]
        lib.ref_labeller_finish.argtypes = [_ptr]
        for name in ("ref_labeller_live", "ref_labeller_spot_count",
                     "ref_labeller_pixel_count"):
            fn = getattr(lib, name)
            fn.restype = _c_size
            fn.argtypes = [_ptr]
        lib.ref_labeller_results.argtypes = [_ptr, _ptr, _ptr, _ptr]
        lib.ref_refl_write.restype = ctypes.c_int
        lib.ref_refl_write.argtypes = [ctypes.c_char_p, _ptr, _c_size, _ptr,
                                       _c_size, _c_size, _c_size, ctypes.c_int,
                                       ctypes.c_char_p, ctypes.c_int]
        _ref = lib
    return _ref


def ref_refl_write(path: str, spots: np.ndarray, pixels: np.ndarray,
                   width: int, panel: int = 0, id: int = 0,
                   identifier: str = "", shoeboxes: bool = True) -> int:
    """The reference's own refl::write (src/refl.cc) on these tables."""
    spots = np.ascontiguousarray(spots, dtype=SPOT)
    pixels = np.ascontiguousarray(pixels, dtype=PIXEL)
    return ref_lib().ref_refl_write(
        path.encode(), spots.ctypes.data, len(spots), pixels.ctypes.data,
        len(pixels), width, panel, id, identifier.encode(), int(shoeboxes))


def _threshold(fn, image: np.ndarray, capacity: int | None):
    image = np.ascontiguousarray(image)
    height, width = image.shape
    if capacity is None:
        capacity = height * width
    out = np.zeros(max(capacity, 1), dtype=SIGNAL)
    found = _c_size(0)
    status = fn(_address(image), height, width, _address(out), capacity,
                ctypes.byref(found))
    return status, out[: min(found.value, capacity)].copy(), found.value


def dext(image: np.ndarray, capacity: int | None = None):
    """(status, list, found) from the oracle's SAT form."""
    lib = oracle_lib()
    fn = lib.oracle_dext_u16 if image.dtype == np.uint16 else lib.oracle_dext_u32
    return _threshold(fn, image, capacity)


def dext_direct(image: np.ndarray, capacity: int | None = None):
    """(status, list, found) from the oracle's direct-window form."""
    lib = oracle_lib()
    fn = (lib.oracle_dext_direct_u16 if image.dtype == np.uint16
          else lib.oracle_dext_direct_u32)
    return _threshold(fn, image, capacity)


def ref_dext(image: np.ndarray, capacity: int | None = None):
    """(status, list, found) from the reference's own dext()."""
    lib = ref_lib()
    fn = lib.ref_dext_u16 if image.dtype == np.uint16 else lib.ref_dext_u32
    return _threshold(fn, image, capacity)


def ref_dext_many(frames: np.ndarray, threads: int) -> int:
    """Threshold a (n, h, w) stack on `threads` host threads; total found."""
    frames = np.ascontiguousarray(frames)
    n, height, width = frames.shape
    lib = ref_lib()
    fn = lib.ref_dext_many_u16 if frames.dtype == np.uint16 else lib.ref_dext_many_u32
    return int(fn(_address(frames), n, height, width, threads))


def ref_dext_cycle(frames: np.ndarray, jobs: int, threads: int) -> int:
    """`jobs` runs of the reference's dext over a (n, h, w) stack (frame
    j % n), on `threads` host threads; total found."""
    frames = np.ascontiguousarray(frames)
    n, height, width = frames.shape
    lib = ref_lib()
    fn = (lib.ref_dext_cycle_u16 if frames.dtype == np.uint16
          else lib.ref_dext_cycle_u32)
    return int(fn(_address(frames), n, height, width, jobs, threads))


def ref_pipeline(frames: np.ndarray, threads: int, path: str):
    """The reference's whole CPU path over a (n, h, w) u16 stack: dext on
    `threads` workers, the Labeller in frame order, refl::write to `path`.
    Returns (status, signal pixels, spots, pixels)."""
    frames = np.ascontiguousarray(frames, dtype=np.uint16)
    n, height, width = frames.shape
    signal, spots, pixels = _c_size(0), _c_size(0), _c_size(0)
    status = ref_lib().ref_pipeline_u16(
        _address(frames), n, height, width, threads, path.encode(),
        ctypes.byref(signal), ctypes.byref(spots), ctypes.byref(pixels))
    return status, signal.value, spots.value, pixels.value


def oracle_dext_many(frames: np.ndarray, threads: int) -> int:
    frames = np.ascontiguousarray(frames, dtype=np.uint16)
    n, height, width = frames.shape
    return int(oracle_lib().oracle_dext_many_u16(_address(frames), n, height,
                                                 width, threads))


def by_index(pixels: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(pixels, dtype=SIGNAL).copy()
    oracle_lib().oracle_by_index(_address(out), len(out))
    return out


def ref_by_index(pixels: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(pixels, dtype=SIGNAL).copy()
    ref_lib().ref_by_index(_address(out), len(out))
    return out


def synthetic_frame(height: int, width: int, seed: int, background: float,
                    with_mask: bool) -> np.ndarray:
    """The reference's test frame (tests/synthetic_frame.hh make_frame)."""
    out = np.zeros((height, width), dtype=np.uint16)
    oracle_lib().oracle_synthetic_frame(_address(out), height, width, seed,
                                        background, 1 if with_mask else 0)
    return out


class _LabellerBase:
    _prefix = ""

    def __init__(self, height, width, min_spot_size=3, max_spot_size=1000,
                 max_separation=2.0, two_d=False):
        self._lib = self._library()
        new = getattr(self._lib, self._prefix + "labeller_new")
        self._handle = new(height, width, min_spot_size, max_spot_size,
                           float(max_separation), 1 if two_d else 0)
        if not self._handle:
            raise ValueError("labeller options rejected")

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            getattr(self._lib, self._prefix + "labeller_free")(handle)
            self._handle = None

    def add(self, frame: int, signal: np.ndarray) -> int:
        signal = np.ascontiguousarray(signal, dtype=SIGNAL)
        return getattr(self._lib, self._prefix + "labeller_add")(
            self._handle, frame, _address(signal) if len(signal) else None,
            len(signal))

    def finish(self) -> None:
        getattr(self._lib, self._prefix + "labeller_finish")(self._handle)

    def results(self):
        """(spots, pixels, counts dict)."""
        n_spots = getattr(self._lib, self._prefix + "labeller_spot_count")(
            self._handle)
        n_pixels = getattr(self._lib, self._prefix + "labeller_pixel_count")(
            self._handle)
        spots = np.zeros(n_spots, dtype=SPOT)
        pixels = np.zeros(n_pixels, dtype=PIXEL)
        counts = (ctypes.c_uint64 * 6)()
        getattr(self._lib, self._prefix + "labeller_results")(
            self._handle, _address(spots) if n_spots else None,
            _address(pixels) if n_pixels else None, counts)
        return spots, pixels, dict(zip(COUNTS, [int(c) for c in counts]))


class Labeller(_LabellerBase):
    """The oracle's batch restatement of dials_spots::Labeller."""

    _prefix = "oracle_"

    @staticmethod
    def _library():
        return oracle_lib()


class RefLabeller(_LabellerBase):
    """The reference's own streaming dials_spots::Labeller."""

    _prefix = "ref_"

    @staticmethod
    def _library():
        return ref_lib()

    def live(self) -> int:
        return int(self._lib.ref_labeller_live(self._handle))


def _bytes_ptr(data: bytes) -> tuple[np.ndarray, int]:
    array = np.frombuffer(data if len(data) else b"\0", dtype=np.uint8)
    return array, array.ctypes.data


def lz4_block(data: bytes, cap: int) -> bytes | None:
    """The oracle's LZ4 block decoder: the bytes, or None when malformed."""
    src, at = _bytes_ptr(data)
    out = np.zeros(max(cap, 1), dtype=np.uint8)
    got = oracle_lib().oracle_lz4_block(at, len(data), out.ctypes.data, cap)
    return None if got < 0 else out[:got].tobytes()


def bitshuffle(elements: np.ndarray) -> np.ndarray:
    """bitshuffle's transpose of one block (size a multiple of 8) as bytes."""
    elements = np.ascontiguousarray(elements)
    out = np.zeros(elements.nbytes, dtype=np.uint8)
    oracle_lib().oracle_bitshuffle(elements.ctypes.data, elements.size,
                                   elements.itemsize, out.ctypes.data)
    return out


def bitunshuffle(planes: np.ndarray, dtype) -> np.ndarray:
    planes = np.ascontiguousarray(planes, dtype=np.uint8)
    dtype = np.dtype(dtype)
    out = np.zeros(planes.size // dtype.itemsize, dtype=dtype)
    oracle_lib().oracle_bitunshuffle(planes.ctypes.data, out.size,
                                     dtype.itemsize, out.ctypes.data)
    return out


def decompress_image(data: bytes, algorithm: str, bit_depth: int,
                     height: int, width: int) -> tuple[bytes | None, str]:
    """decompress::image (src/decompress.cc:76-136): (frame bytes, "") or
    (None, the reference's message)."""
    src, at = _bytes_ptr(data)
    expected = height * width * bit_depth // 8
    out = np.zeros(max(expected, 1), dtype=np.uint8)
    msg = ctypes.create_string_buffer(256)
    status = oracle_lib().oracle_decompress_image(
        at, len(data), algorithm.encode(), bit_depth, height, width,
        out.ctypes.data, expected, msg, len(msg))
    if status != 0:
        return None, msg.value.decode()
    return out[:expected].tobytes(), ""
