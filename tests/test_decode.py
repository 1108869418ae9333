"""Chunk decoding (src/decompress.cc), on the CPU: the oracle's restatement
of decompress::image and the product's host encoders.

What pins the oracle:
- LZ4 blocks: the system's liblz4 (an independent implementation, loaded
  with ctypes where installed): the oracle decodes every block liblz4's
  compressors write, liblz4 decodes every block the product's encoder
  writes, and on thousands of corrupted blocks the oracle and
  LZ4_decompress_safe agree on what decodes and to what.
- bitshuffle's bit planes: a numpy construction from the published layout
  (np.unpackbits / np.packbits) and hand-made known answers.
- decompress::image's checks: its messages, quoted from
  src/decompress.cc:76-136.
bitshuffle itself (kiyo-masui/bitshuffle) is absent here, so chunks come
from the product's encoder, each of whose LZ4 blocks liblz4 checks.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import cpu
import spotfinder as sf


def _liblz4():
    try:
        lib = ctypes.CDLL("liblz4.so.1")
    except OSError:
        return None
    lib.LZ4_compress_default.restype = ctypes.c_int
    lib.LZ4_compress_default.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_int]
    lib.LZ4_compress_HC.restype = ctypes.c_int
    lib.LZ4_compress_HC.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.LZ4_decompress_safe.restype = ctypes.c_int
    lib.LZ4_decompress_safe.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_int]
    return lib


LZ4 = _liblz4()
needs_lz4 = pytest.mark.skipif(LZ4 is None,
                               reason="the system liblz4 is not installed")


def lz4_compress(data: bytes, hc_level: int | None = None) -> bytes:
    src = np.frombuffer(data if data else b"\0", dtype=np.uint8)
    cap = len(data) + len(data) // 255 + 16
    out = np.zeros(cap, dtype=np.uint8)
    if hc_level is None:
        n = LZ4.LZ4_compress_default(src.ctypes.data, out.ctypes.data,
                                     len(data), cap)
    else:
        n = LZ4.LZ4_compress_HC(src.ctypes.data, out.ctypes.data, len(data),
                                cap, hc_level)
    assert n > 0
    return out[:n].tobytes()


def lz4_decompress(data: bytes, cap: int) -> bytes | None:
    src = np.frombuffer(data if data else b"\0", dtype=np.uint8)
    out = np.zeros(max(cap, 1) + 64, dtype=np.uint8)
    n = LZ4.LZ4_decompress_safe(src.ctypes.data, out.ctypes.data, len(data),
                                cap)
    return None if n < 0 else out[:n].tobytes()


def corpus() -> list[bytes]:
    rng = np.random.default_rng(5)
    items = [b"", b"a", bytes(7), bytes(12), bytes(13), bytes(8192),
             bytes(range(256)) * 32, rng.integers(0, 256, 8192, np.uint8)
             .tobytes(), (b"abcabcabd" * 1000)[:8000]]
    for n in (1, 5, 11, 12, 13, 17, 31, 64, 100, 1000):
        items.append(rng.integers(0, 3, n, np.uint8).tobytes())
    # the bit planes of diffraction blocks: what bslz4 feeds LZ4
    flat = cpu.synthetic_frame(64, 512, 3, 30.0, True).reshape(-1)
    for at in (0, 4096, 8192):
        items.append(cpu.bitshuffle(flat[at:at + 4096]).tobytes())
    # long runs with literals between: long lengths and overlapping matches
    runs = bytearray()
    for k in range(40):
        runs += bytes([k % 7]) * int(rng.integers(1, 700))
        runs += rng.integers(0, 256, int(rng.integers(0, 40)),
                             np.uint8).tobytes()
    items.append(bytes(runs))
    return items


@needs_lz4
@pytest.mark.parametrize("level", [None, 3, 9, 12])
def test_oracle_decodes_liblz4_blocks(level):
    for data in corpus():
        block = lz4_compress(data, level)
        assert cpu.lz4_block(block, len(data)) == data, len(data)


@needs_lz4
def test_encoder_blocks_decode_with_liblz4():
    for data in corpus():
        block = sf.lz4_compress_block(data)
        assert lz4_decompress(block, len(data)) == data, len(data)
        assert cpu.lz4_block(block, len(data)) == data, len(data)


@needs_lz4
def test_corrupt_blocks_agree_with_liblz4():
    """Bytes flipped, blocks cut short or lengthened, capacities too small:
    the oracle accepts exactly what LZ4_decompress_safe accepts, with the
    same bytes."""
    rng = np.random.default_rng(8)
    cases = 0
    for data in corpus():
        if not data:
            continue
        good = bytearray(lz4_compress(data))
        for trial in range(60):
            block = bytearray(good)
            cap = len(data)
            kind = trial % 4
            if kind == 0:
                for _ in range(int(rng.integers(1, 4))):
                    block[int(rng.integers(len(block)))] = int(
                        rng.integers(256))
            elif kind == 1:
                block = block[:int(rng.integers(len(block)))]
            elif kind == 2:
                block += rng.integers(0, 256, int(rng.integers(1, 9)),
                                      np.uint8).tobytes()
            else:
                cap = int(rng.integers(0, len(data) + 1))
            want = lz4_decompress(bytes(block), cap)
            got = cpu.lz4_block(bytes(block), cap)
            assert got == want, (len(data), trial)
            cases += 1
    assert cases > 1000


def test_bitshuffle_matches_bit_planes():
    """Plane 8 j + k holds bit k of byte j of every element, element i at
    bit i % 8 of byte i / 8 (numpy's little-endian bit order)."""
    rng = np.random.default_rng(3)
    for dtype, n in ((np.uint8, 64), (np.uint16, 4096), (np.uint32, 128),
                     (np.uint64, 8)):
        x = rng.integers(0, np.iinfo(dtype).max, n, dtype=dtype,
                         endpoint=True)
        bits = np.unpackbits(x.view(np.uint8).reshape(n, -1), axis=1,
                             bitorder="little")  # column 8 j + k
        planes = np.packbits(bits.T, axis=1, bitorder="little").reshape(-1)
        assert np.array_equal(cpu.bitshuffle(x), planes)
        assert np.array_equal(cpu.bitunshuffle(planes, dtype), x)
    one = np.zeros(8, np.uint16)
    one[0] = 1
    want = np.zeros(16, np.uint8)
    want[0] = 1  # plane 0, element 0
    assert np.array_equal(cpu.bitshuffle(one), want)
    top = np.full(16, 0x8000, np.uint16)
    want = np.zeros(32, np.uint8)
    want[30:32] = 0xFF  # plane 15: every element's top bit
    assert np.array_equal(cpu.bitshuffle(top), want)
    assert cpu.oracle_lib().oracle_bshuf_default_block(2) == 4096
    assert cpu.oracle_lib().oracle_bshuf_default_block(4) == 2048
    assert cpu.oracle_lib().oracle_bshuf_default_block(128) == 128


def split_chunk(chunk: bytes, n: int, es: int, block_elems: int):
    """The blocks of an HDF5 bitshuffle chunk: [(lz4 bytes, elements)] and
    the raw tail."""
    pos, blocks, done = 12, [], 0
    full, rest = divmod(n, block_elems)
    last = rest - rest % 8
    for k in range(full + (1 if last else 0)):
        count = block_elems if k < full else last
        comp = int.from_bytes(chunk[pos:pos + 4], "big")
        blocks.append((chunk[pos + 4:pos + 4 + comp], count))
        pos += 4 + comp
        done += count
    return blocks, chunk[pos:]


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32])
def test_bslz4_chunks_round_trip(dtype):
    rng = np.random.default_rng(np.dtype(dtype).itemsize)
    es = np.dtype(dtype).itemsize
    for n, block in ((1, 0), (7, 0), (8, 0), (100, 0), (4100, 0),
                     (9000, 0), (3000, 8), (5003, 64), (12345, 2048)):
        x = rng.poisson(0.5, n).astype(dtype)
        x[rng.random(n) < 0.01] = np.iinfo(dtype).max
        chunk = sf.bslz4_compress(x, block)
        assert int.from_bytes(chunk[:8], "big") == n * es
        block_elems = block or cpu.oracle_lib().oracle_bshuf_default_block(es)
        assert int.from_bytes(chunk[8:12], "big") == block_elems * es
        out, msg = cpu.decompress_image(chunk, "bslz4", 8 * es, 1, n)
        assert msg == "" and out == x.tobytes(), (n, block)
        # each block, decoded by liblz4, is the numpy bit planes of its
        # elements; the tail is the last n % 8 elements as they are
        blocks, tail = split_chunk(chunk, n, es, block_elems)
        done = 0
        for comp, count in blocks:
            if LZ4 is not None:
                planes = lz4_decompress(comp, count * es)
                part = x[done:done + count]
                bits = np.unpackbits(part.view(np.uint8).reshape(count, -1),
                                     axis=1, bitorder="little")
                want = np.packbits(bits.T, axis=1, bitorder="little")
                assert planes == want.tobytes(), (n, block, done)
            done += count
        assert tail == x[done:].tobytes() and n - done == n % 8


def lz4_chunk(data: bytes, block: int, header_block: int | None = None):
    """The HDF5 LZ4 filter's chunk (src/decompress.cc:40-44)."""
    out = bytearray(len(data).to_bytes(8, "big"))
    out += (block if header_block is None else header_block).to_bytes(4, "big")
    for at in range(0, len(data), block):
        comp = sf.lz4_compress_block(data[at:at + block])
        out += len(comp).to_bytes(4, "big") + comp
    return bytes(out)


def test_lz4_chunks_round_trip():
    rng = np.random.default_rng(4)
    x = rng.poisson(2.0, (30, 50)).astype(np.uint16)
    for block in (100, 1000, 3000, x.nbytes):
        chunk = lz4_chunk(x.tobytes(), block)
        out, msg = cpu.decompress_image(chunk, "lz4", 16, 30, 50)
        assert msg == "" and out == x.tobytes(), block
    # a zero block size in the header: the whole frame is one block
    chunk = lz4_chunk(x.tobytes(), x.nbytes, header_block=0)
    assert cpu.decompress_image(chunk, "lz4", 16, 30, 50)[0] == x.tobytes()


def lz4_pieces_chunk(data: bytes, header_block: int, pieces) -> bytes:
    """An LZ4 filter chunk whose blocks hold the given numbers of bytes of
    `data` in turn (each at most header_block, some fewer)."""
    out = bytearray(len(data).to_bytes(8, "big"))
    out += header_block.to_bytes(4, "big")
    at = 0
    for n in pieces:
        comp = sf.lz4_compress_block(data[at:at + n])
        out += len(comp).to_bytes(4, "big") + comp
        at += n
    assert at == len(data)
    return bytes(out)


SHORT_PIECES = [(1000, [600, 1000, 0, 1000, 400]),  # an empty block too
                (0, [700, 2300]),                   # 0: the whole frame
                (3000, [1, 2999])]


def test_lz4_chunks_with_short_blocks():
    """A block may decode to fewer bytes than the header's block size: the
    next block starts where it ended (src/decompress.cc:44-70 advances by
    what LZ4_decompress_safe produced), so the frame is the blocks' bytes in
    turn; a chunk whose blocks run out first is truncated."""
    rng = np.random.default_rng(5)
    x = rng.poisson(2.0, (30, 50)).astype(np.uint16)
    data = x.tobytes()
    for header_block, pieces in SHORT_PIECES:
        chunk = lz4_pieces_chunk(data, header_block, pieces)
        out, msg = cpu.decompress_image(chunk, "lz4", 16, 30, 50)
        assert msg == "" and out == data, (header_block, pieces)
    short = lz4_pieces_chunk(data[:2000], 1000, [600, 1000, 400])
    short = len(data).to_bytes(8, "big") + short[8:]
    assert cpu.decompress_image(short, "lz4", 16, 30, 50)[1] == \
        "lz4 data truncated"


def test_messages_follow_the_reference():
    """The messages of decompress::image (src/decompress.cc:83-135)."""
    x = np.arange(8, dtype=np.uint16)
    chunk = sf.bslz4_compress(x)
    dec = cpu.decompress_image
    assert dec(x.tobytes(), "", 16, 2, 4) == (x.tobytes(), "")
    assert dec(x.tobytes()[:3], "", 16, 2, 4)[1] == (
        "uncompressed frame is 3 bytes, expected 16")
    assert dec(chunk[:11], "bslz4", 16, 2, 4)[1] == (
        "too short for a filter header")
    assert dec(chunk, "bslz4", 16, 2, 5)[1] == (
        "filter header says 16 bytes uncompressed, expected 20")
    assert dec(chunk, "zstd", 16, 2, 4)[1] == (
        "unknown compression algorithm zstd")
    assert dec(chunk, "bslz4", 0, 2, 4)[1] == (
        "output buffer too small for the frame")
    odd = chunk[:8] + (6).to_bytes(4, "big") + chunk[12:]  # 3 elements
    assert dec(odd, "bslz4", 16, 2, 4)[1] == (
        "bitshuffle decompression failed with code -81")
    assert dec(chunk[:-3], "bslz4", 16, 2, 4)[1] == (
        "bitshuffle decompression failed with code -91")
    lz = lz4_chunk(x.tobytes(), 16)
    assert dec(lz[:14], "lz4", 16, 2, 4)[1] == "lz4 data truncated"
    zero = lz[:12] + bytes(4) + lz[16:]
    assert dec(zero, "lz4", 16, 2, 4)[1] == "lz4 block size out of range"
    assert dec(lz[:-1], "lz4", 16, 2, 4)[1] == "lz4 block size out of range"
    bad = bytearray(lz)
    assert bad[16] == 0xF0 and bad[17] == 1  # 16 literals: 15 + 1
    bad[17] = 200  # a literal run longer than the block
    assert dec(bytes(bad), "lz4", 16, 2, 4)[1].startswith(
        "lz4 decompression failed with code ")


def test_encoder_arguments():
    """The host encoder refuses a block size that is not a multiple of 8
    and writes the default block size into the header otherwise."""
    x = np.arange(100, dtype=np.uint16)
    with pytest.raises(ValueError, match="multiple of 8"):
        sf.bslz4_compress(x, 12)
    chunk = sf.bslz4_compress(x, 0)
    assert int.from_bytes(chunk[8:12], "big") == 8192
    assert sf.lib().sf_bslz4_bound(100, 2, 0) >= len(chunk)
    empty = sf.bslz4_compress(np.zeros(0, np.uint16))
    assert len(empty) == 12 and int.from_bytes(empty[:8], "big") == 0
    assert cpu.decompress_image(empty, "bslz4", 16, 0, 0)[1] == ""
