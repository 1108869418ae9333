"""Chunks decoded on the device (sf_hip_decoder_*): every frame equals the
oracle's decompress::image (src/decompress.cc:76-136) of its chunk, and a
decoded series thresholds and groups to the reference's published answer."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu
import frames as fr
from test_decode import (LZ4, SHORT_PIECES, lz4_chunk, lz4_compress,
                         lz4_pieces_chunk)

pytestmark = pytest.mark.gpu


def upload_chunks(sf, chunks, skew=0):
    """One device buffer holding the chunks, chunk f at a 256-byte boundary
    plus f * skew bytes; returns (buffer, device pointers)."""
    offsets, at = [], 0
    for f, c in enumerate(chunks):
        at = (at + 255) // 256 * 256 + f * skew
        offsets.append(at)
        at += len(c)
    buf = sf.DeviceBuffer(max(at, 1))
    for c, off in zip(chunks, offsets):
        if len(c):
            buf.upload(np.frombuffer(c, np.uint8), off)
    return buf, [buf.ptr + off for off in offsets]


def decode_batch(sf, chunks, algorithm, bit_depth, h, w, skew=0,
                 frame_skew=0):
    """Decodes the chunks into device frames; returns the frames' bytes."""
    n = len(chunks)
    frame_bytes = h * w * bit_depth // 8
    stride = (frame_bytes + 255) // 256 * 256 + 256
    frames = sf.DeviceBuffer(n * stride)
    frames.upload(np.full(n * stride, 0xA5, np.uint8))
    ptrs = [frames.ptr + f * stride + frame_skew for f in range(n)]
    buf, dev = upload_chunks(sf, chunks, skew)
    dec = sf.Decoder(bit_depth, h, w, max_frames=n)
    dec.index(chunks, algorithm)
    dec.decode(dev, ptrs)
    dec.wait()
    host = np.zeros(n * stride, np.uint8)
    frames.download(host)
    out = []
    for f in range(n):
        at = f * stride + frame_skew
        out.append(host[at:at + frame_bytes].tobytes())
        # nothing written outside the frame
        assert (host[at + frame_bytes:(f + 1) * stride] == 0xA5).all()
        assert (host[f * stride:at] == 0xA5).all()
    dec.close()
    buf.close()
    frames.close()
    return out


def check(sf, chunks, algorithm, bit_depth, h, w, **kw):
    got = decode_batch(sf, chunks, algorithm, bit_depth, h, w, **kw)
    for f, c in enumerate(chunks):
        want, msg = cpu.decompress_image(c, algorithm, bit_depth, h, w)
        assert msg == "", msg
        assert got[f] == want, f


def diffraction_frames(sf, h, w, n, seed):
    buf = sf.DeviceBuffer(2 * h * w * n)
    spots = fr.diffraction_spots(seed, h, w, n, 40.0)
    sf.synth_frames(buf, 2, h, w, n, seed, 0.3, True, 0.0005, spots)
    host = np.zeros((n, h, w), np.uint16)
    buf.download(host)
    buf.close()
    return host


def test_bslz4_u16_diffraction_frames(sf):
    h, w, n = 512, 1030, 4
    frames = diffraction_frames(sf, h, w, n, 21)
    chunks = [sf.bslz4_compress(frames[f]) for f in range(n)]
    check(sf, chunks, "bslz4", 16, h, w)
    got = decode_batch(sf, chunks, "bslz4", 16, h, w)
    assert all(got[f] == frames[f].tobytes() for f in range(n))


@pytest.mark.parametrize("bit_depth,h,w,block", [
    (8, 37, 211, 0), (16, 37, 211, 0), (32, 37, 211, 0), (64, 13, 29, 0),
    (16, 1, 7, 0), (16, 3, 5, 8), (16, 64, 100, 64), (32, 50, 77, 2048),
    (16, 100, 300, 8192)])
def test_bslz4_shapes_and_block_sizes(sf, bit_depth, h, w, block):
    """Ragged element counts (a raw tail), the last partial block, block
    sizes from 8 elements to 16 KiB, every element size."""
    rng = np.random.default_rng(bit_depth * h + w)
    dtype = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[
        bit_depth]
    chunks = []
    for f in range(3):
        x = rng.poisson(1.5 + 20 * f, (h, w)).astype(dtype)
        x[rng.random((h, w)) < 0.02] = np.iinfo(dtype).max
        chunks.append(sf.bslz4_compress(x, block))
    check(sf, chunks, "bslz4", bit_depth, h, w)


def test_unaligned_chunks_and_frames(sf):
    """Chunks at odd device addresses (the 16-byte staging loads start
    below them) and frames off 16-byte alignment (byte stores)."""
    h, w = 40, 333
    rng = np.random.default_rng(9)
    chunks = [sf.bslz4_compress(rng.poisson(3.0, (h, w)).astype(np.uint16))
              for _ in range(5)]
    check(sf, chunks, "bslz4", 16, h, w, skew=3)
    check(sf, chunks, "bslz4", 16, h, w, skew=7, frame_skew=2)


@pytest.mark.skipif(LZ4 is None, reason="the system liblz4 is not installed")
@pytest.mark.parametrize("level", [None, 12])
def test_liblz4_blocks(sf, level):
    """bslz4 chunks whose LZ4 blocks come from liblz4's own compressors
    (long overlapping matches at level 12)."""
    h, w = 64, 700
    frames = diffraction_frames(sf, h, w, 2, 5)
    chunks = []
    for x in frames:
        flat = x.reshape(-1)
        out = bytearray(flat.nbytes.to_bytes(8, "big"))
        out += (8192).to_bytes(4, "big")
        done = 0
        while flat.size - done >= 8:
            count = min(4096, (flat.size - done) // 8 * 8)
            comp = lz4_compress(cpu.bitshuffle(flat[done:done + count])
                                .tobytes(), level)
            out += len(comp).to_bytes(4, "big") + comp
            done += count
        out += flat[done:].tobytes()
        chunks.append(bytes(out))
    got = decode_batch(sf, chunks, "bslz4", 16, h, w)
    assert all(got[f] == frames[f].tobytes() for f in range(2))
    check(sf, chunks, "bslz4", 16, h, w)


def many_sequence_block(n_seq: int, cap: int, rng) -> bytes:
    """An LZ4 block of exactly n_seq sequences decoding to cap bytes:
    n_seq - 1 of one literal and a 4-byte match at offset 1, then the last
    literals (at least 12 bytes, as the format asks)."""
    out = bytearray()
    for _ in range(n_seq - 1):
        out += bytes([0x10, int(rng.integers(256)), 0x01, 0x00])
    rest = cap - 5 * (n_seq - 1)
    assert rest >= 12
    out.append(0xF0)
    v = rest - 15
    while v >= 255:
        out.append(255)
        v -= 255
    out.append(v)
    out += rng.integers(0, 256, rest, np.uint8).tobytes()
    return bytes(out)


@pytest.mark.parametrize("n_seq", [1, 255, 256, 257, 1500])
def test_blocks_with_many_sequences(sf, n_seq):
    """Blocks around the parse pass's 256 recorded sequences (more are
    parsed again in LDS), in a chunk of two blocks and a raw tail."""
    rng = np.random.default_rng(n_seq)
    h, w = 1, 8195  # two 4096-element blocks and 3 raw elements
    blocks = [many_sequence_block(n_seq, 8192, rng),
              many_sequence_block(max(1, n_seq // 2), 8192, rng)]
    assert cpu.lz4_block(blocks[0], 8192) is not None
    chunk = bytearray((h * w * 2).to_bytes(8, "big") + (8192).to_bytes(4, "big"))
    for b in blocks:
        chunk += len(b).to_bytes(4, "big") + b
    chunk += rng.integers(0, 256, 6, np.uint8).tobytes()
    check(sf, [bytes(chunk)] * 2, "bslz4", 16, h, w)


def test_lz4_and_stored_chunks(sf):
    rng = np.random.default_rng(6)
    h, w = 30, 70
    xs = [rng.poisson(2.0, (h, w)).astype(np.uint16) for _ in range(3)]
    for block in (64, 1000, h * w * 2):
        chunks = [lz4_chunk(x.tobytes(), block) for x in xs]
        check(sf, chunks, "lz4", 16, h, w)
    chunks = [lz4_chunk(x.tobytes(), h * w * 2, header_block=0) for x in xs]
    check(sf, chunks, "lz4", 16, h, w)
    chunks = [x.tobytes() for x in xs]
    check(sf, chunks, "", 16, h, w)


def test_lz4_chunks_with_short_blocks(sf):
    """Blocks that decode to fewer bytes than the header's block size (an
    empty one among them): each next block goes where the last one ended,
    as the reference's loop places them; equal to the oracle."""
    rng = np.random.default_rng(8)
    xs = [rng.poisson(2.0, (30, 50)).astype(np.uint16) for _ in range(3)]
    chunks = [lz4_pieces_chunk(x.tobytes(), hb, pieces)
              for x, (hb, pieces) in zip(xs, SHORT_PIECES)]
    check(sf, chunks, "lz4", 16, 30, 50)
    dec = sf.Decoder(16, 30, 50, max_frames=1)
    short = lz4_pieces_chunk(xs[0].tobytes()[:2000], 1000, [600, 1000, 400])
    short = (3000).to_bytes(8, "big") + short[8:]
    with pytest.raises(sf.DecodeError, match="^lz4 data truncated$"):
        dec.index([short], "lz4")
    dec.close()


def test_errors_follow_the_reference(sf):
    h, w = 16, 32
    x = np.arange(h * w, dtype=np.uint16)
    good = sf.bslz4_compress(x)
    dec = sf.Decoder(16, h, w, max_frames=4)
    frame = sf.DeviceBuffer(2 * h * w)
    # header checks on the host, the reference's messages
    with pytest.raises(sf.DecodeError,
                       match="^filter header says 1024 bytes uncompressed, "
                             "expected 1026$"):
        sf.Decoder(16, 1, 513).index([good])
    with pytest.raises(sf.DecodeError, match="^too short for a filter header$"):
        dec.index([good[:5]])
    with pytest.raises(sf.DecodeError,
                       match="^unknown compression algorithm zstd$"):
        dec.index([good], "zstd")
    with pytest.raises(sf.DecodeError, match="^chunk 1: .* code -91$"):
        dec.index([good, good[:-9]])
    with pytest.raises(ValueError, match="exceed the device decoder"):
        dec.index([sf.bslz4_compress(x, 16384)])
    with pytest.raises(sf.DecodeError, match="^uncompressed frame is 5 bytes"):
        dec.image(b"12345", "", frame.ptr)
    # a block that is well framed but does not decode: found on the device
    bad = bytearray(good)
    bad[16] ^= 0xFF  # the first block's token
    want = cpu.decompress_image(bytes(bad), "bslz4", 16, h, w)[1]
    assert want == "bitshuffle decompression failed with code -91"
    with pytest.raises(sf.DecodeError, match="^" + want + "$"):
        dec.image(bytes(bad), "bslz4", frame.ptr)
    buf, dev = upload_chunks(sf, [good, bytes(bad)])
    dec.index([good, bytes(bad)])
    dec.decode(dev, [frame.ptr, frame.ptr])
    with pytest.raises(sf.DecodeError, match="^chunk 1: bitshuffle .* -91$"):
        dec.wait()
    lz = bytearray(lz4_chunk(x.tobytes(), 256))
    lz[16:19] = b"\xf0\xff\xff"  # a literal run past the block
    with pytest.raises(sf.DecodeError,
                       match="^lz4 decompression failed with code -1$"):
        dec.image(bytes(lz), "lz4", frame.ptr)
    # and the decoder still works
    dec.image(good, "bslz4", frame.ptr)
    out = np.zeros(h * w, np.uint16)
    frame.download(out)
    assert np.array_equal(out, x)
    buf.close()
    frame.close()
    dec.close()


def test_decoded_fixture_gives_the_published_answer(sf):
    """The 12-frame make_test_nxmx.py series, bslz4-compressed as its HDF5
    file holds it, decoded on the device and thresholded there: 8,928
    signal pixels and 48 spots (README.md:95-107)."""
    stack = fr.nxmx_fixture()
    n, h, w = stack.shape
    chunks = [sf.bslz4_compress(stack[z]) for z in range(n)]
    frames = sf.DeviceBuffer(stack.nbytes)
    ptrs = [frames.ptr + 2 * h * w * z for z in range(n)]
    buf, dev = upload_chunks(sf, chunks)
    dec = sf.Decoder(16, h, w, max_frames=n)
    dec.index(chunks)
    dec.decode(dev, ptrs)
    dec.wait()
    plan = sf.Plan(2, [(h, w)] * n)
    plan.run(ptrs)
    plan.wait()
    counts = plan.counts()
    assert int(np.sum(counts)) == 8928
    lab = sf.Labeller(h, w)
    lab.add_plan(plan, first_frame=0)
    lab.finish()
    spots, _, _ = lab.results()
    assert len(spots) == 48
    # the one-chunk drop-in gives the same frames
    one = sf.DeviceBuffer(2 * h * w)
    host = np.zeros((h, w), np.uint16)
    for z in (0, 5, 11):
        dec.image(chunks[z], "bslz4", one.ptr)
        one.download(host)
        assert np.array_equal(host, stack[z])
    one.close()
    lab.close()
    plan.close()
    dec.close()
    buf.close()
    frames.close()


def test_lz4_whole_frame_blocks(sf):
    """The HDF5 LZ4 filter's default: the whole frame one LZ4 block, which
    one wave decodes in global memory (the fixture's 256 x 384 frames,
    196 KB a block, thousands of sequences)."""
    stack = fr.nxmx_fixture()[:3]
    n, h, w = stack.shape
    chunks = [lz4_chunk(stack[z].tobytes(), h * w * 2, header_block=0)
              for z in range(n)]
    got = decode_batch(sf, chunks, "lz4", 16, h, w)
    assert all(got[z] == stack[z].tobytes() for z in range(n))
    check(sf, chunks, "lz4", 16, h, w)
