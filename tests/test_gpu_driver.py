"""The series driver on the device (sf_hip_driver_*), held to the reference's
find_spots (src/find_spots.cc:460-713) and its regression invariants
(tests/regression.sh:141-195): the published summary of the 12-frame
fixture (README.md:93-106), --2d giving three times the spots, a frame
never written splitting the spots that spanned it, and a table that does
not depend on the chunk size. Spot tables are compared with the oracle
labeller fed the same frames; the .refl bytes with the reference's own
writer where it is built."""

from __future__ import annotations

import re

import numpy as np
import pytest

from oracle import cpu
import frames as fr

pytestmark = pytest.mark.gpu

H, W = 256, 384


@pytest.fixture(scope="module")
def fixture_chunks(sf):
    stack = fr.nxmx_fixture()
    assert stack.shape == (12, H, W)
    return stack, [sf.bslz4_compress(stack[z]) for z in range(12)]


def run_driver(sf, chunks, order=None, missing=(), chunk=4, z_offset=0,
               two_d=False, numbers=None):
    d = sf.Driver(16, H, W, chunk=chunk, z_offset=z_offset, two_d=two_d)
    numbers = list(range(len(chunks))) if numbers is None else numbers
    for k in (order if order is not None else range(len(chunks))):
        d.submit(numbers[k], None if k in missing else chunks[k])
    d.finish()
    return d


def oracle_tables(stack, frames, z_offset=0, two_d=False):
    lab = cpu.Labeller(H, W, two_d=two_d)
    for z in frames:
        assert lab.add(z + z_offset, cpu.dext(stack[z])[1]) == 0
    lab.finish()
    return lab.results()


def same_tables(a, b):
    return (a[0].tobytes() == b[0].tobytes() and
            a[1].tobytes() == b[1].tobytes() and a[2] == b[2])


def test_published_summary(sf, fixture_chunks, tmp_path, monkeypatch):
    """The summary lines of README.md:93-106, word for word (the timing in
    the first aside), and the table equal to the reference writer's."""
    stack, chunks = fixture_chunks
    monkeypatch.chdir(tmp_path)
    d = run_driver(sf, chunks)
    d.write("strong.refl", identifier="1c0ffee0")
    lines = d.log().splitlines()
    assert re.fullmatch(r"Thresholded 12 of 12 images in \d+\.\d s "
                        r"\(\d+\.\d images/s\), 0 never written, 0 failures",
                        lines[0]), lines[0]
    assert lines[1:] == [
        "Found 8928 signal pixels on 12 frames",
        "Extracted 48 spots",
        "Removed 0 spots with size < 3 pixels",
        "Removed 0 spots with size > 1000 pixels",
        "Calculated 48 spot centroids",
        "Calculated 48 spot intensities",
        "Filtered 48 of 48 spots by peak-centroid distance",
        "Wrote 48 reflections to strong.refl (0.1 MB, most of it shoeboxes)",
    ]
    want = oracle_tables(stack, range(12))
    assert same_tables(d.results(), want)
    if cpu.have_ref():
        cpu.ref_refl_write(str(tmp_path / "ref.refl"), want[0], want[1], W,
                           identifier="1c0ffee0")
        assert (tmp_path / "strong.refl").read_bytes() == (
            tmp_path / "ref.refl").read_bytes()
    d.close()


def test_two_d_gives_three_times_the_spots(sf, fixture_chunks):
    """tests/regression.sh:141-161: every reflection spans three frames."""
    stack, chunks = fixture_chunks
    three = run_driver(sf, chunks)
    two = run_driver(sf, chunks, two_d=True)
    n3, n2 = len(three.results()[0]), len(two.results()[0])
    assert n3 == 48 and n2 == 3 * n3
    assert same_tables(two.results(), oracle_tables(stack, range(12),
                                                    two_d=True))
    three.close()
    two.close()


def test_table_does_not_depend_on_the_chunk(sf, fixture_chunks, tmp_path):
    """tests/regression.sh:183-195: byte-identical tables whatever the chunk
    size and the order frames arrive in within the series."""
    _, chunks = fixture_chunks
    rng = np.random.default_rng(4)
    tables = []
    for chunk in (1, 3, 5, 12, 64):
        order = rng.permutation(12) if chunk >= 12 else None
        d = run_driver(sf, chunks, order=order, chunk=chunk)
        path = tmp_path / f"c{chunk}.refl"
        d.write(str(path))
        tables.append(path.read_bytes())
        assert d.stats()["out_of_order"] == 0
        d.close()
    assert all(t == tables[0] for t in tables)


@pytest.mark.parametrize("chunk", [4, 5])
def test_frame_never_written(sf, fixture_chunks, chunk):
    """tests/regression.sh:163-181: a frame never written is reported and
    nothing connects across it."""
    stack, chunks = fixture_chunks
    d = run_driver(sf, chunks, missing={5}, chunk=chunk)
    assert ", 1 never written, 0 failures" in d.log().splitlines()[0]
    assert "Found" in d.log() and "on 11 frames" in d.log()
    want = oracle_tables(stack, [z for z in range(12) if z != 5])
    assert same_tables(d.results(), want)
    assert len(want[0]) > 48  # spots across frame 5 split in two
    d.close()


def test_z_offset(sf, fixture_chunks):
    stack, chunks = fixture_chunks
    d = run_driver(sf, chunks, z_offset=100)
    assert same_tables(d.results(), oracle_tables(stack, range(12), 100))
    d.close()


def test_late_frame_is_dropped(sf, fixture_chunks):
    """src/find_spots.cc:571-578: a frame that arrives after a later one
    was grouped is dropped and counted, with the warning line."""
    stack, chunks = fixture_chunks
    d = sf.Driver(16, H, W, chunk=4)
    for z in range(8):
        d.submit(z, chunks[z])
    d.submit(2, chunks[2])  # late: frames 0-7 are grouped
    for z in range(8, 12):
        d.submit(z, chunks[z])
    d.finish()
    log = d.log()
    assert "warning: 1 frames arrived after a later one and were dropped" \
        in log
    assert d.stats()["out_of_order"] == 1
    assert "Thresholded 13 of 13 images" in log
    assert same_tables(d.results(), oracle_tables(stack, range(12)))
    d.close()


def test_a_number_submitted_twice(sf, fixture_chunks):
    """A frame number that arrives twice in one chunk is grouped once; the
    second is dropped as late."""
    stack, chunks = fixture_chunks
    d = sf.Driver(16, H, W, chunk=6)
    for z in range(12):
        d.submit(z, chunks[z])
        if z == 4:
            d.submit(4, chunks[4])
    d.finish()
    assert d.stats()["out_of_order"] == 1 and d.stats()["failed"] == 0
    assert same_tables(d.results(), oracle_tables(stack, range(12)))
    d.close()


def test_failures_are_logged_and_leave_a_gap(sf, fixture_chunks):
    """src/find_spots.cc:528-534: a frame that does not decode is logged
    with the reference's message and counted; the rest are grouped."""
    stack, chunks = fixture_chunks
    bad = bytearray(chunks[3])
    bad[16] ^= 0xFF  # the first block's token
    short = chunks[7][:20]  # a truncated chunk
    data = list(chunks)
    data[3] = bytes(bad)
    data[7] = short
    d = run_driver(sf, data, chunk=6)
    log = d.log()
    want3 = cpu.decompress_image(bytes(bad), "bslz4", 16, H, W)[1]
    want7 = cpu.decompress_image(short, "bslz4", 16, H, W)[1]
    assert f"frame 3: {want3}" in log.splitlines()
    assert f"frame 7: {want7}" in log.splitlines()
    assert ", 0 never written, 2 failures" in log
    assert d.stats()["failed"] == 2
    want = oracle_tables(stack, [z for z in range(12) if z not in (3, 7)])
    assert same_tables(d.results(), want)
    d.close()


def test_u32_frames_in_every_algorithm(sf):
    """32-bit frames whose chunks come in all three forms in one batch
    (bitshuffle+LZ4, plain LZ4, stored: nxmx.cc gives a chunk whose filter
    was skipped the empty algorithm), against the oracle labeller."""
    from test_decode import lz4_chunk

    base = fr.u32_frame()
    h, w = base.shape
    frames = [np.roll(base, 9 * z, axis=1) for z in range(6)]
    forms = ["bslz4", "lz4", "", "bslz4", "", "lz4"]
    d = sf.Driver(32, h, w, chunk=4)
    for z, (x, form) in enumerate(zip(frames, forms)):
        if form == "bslz4":
            data = sf.bslz4_compress(x)
        elif form == "lz4":
            data = lz4_chunk(x.tobytes(), 4096)
        else:
            data = x.tobytes()
        d.submit(z, data, form)
    d.finish()
    assert d.stats()["failed"] == 0 and d.stats()["grouped"] == 6
    lab = cpu.Labeller(h, w)
    for z, x in enumerate(frames):
        assert lab.add(z, cpu.dext(x)[1]) == 0
    lab.finish()
    want = lab.results()
    assert len(want[0]) > 0
    assert same_tables(d.results(), want)
    d.close()


def test_detector_mask_before_the_threshold(sf, fixture_chunks):
    """sf_hip_driver_set_mask: the mask stamped into each decoded frame
    before the threshold (the step README.md:227-234 says the reference
    leaves out), against the oracle labeller over the host-stamped
    frames."""
    stack, chunks = fixture_chunks
    rng = np.random.default_rng(21)
    mask = rng.random((H, W)) < 0.02
    mask[60:63, :] = True  # rows through planted spots
    mask[:, 100:104] = True
    pm = sf.PixelMask(mask)
    d = sf.Driver(16, H, W, chunk=5)
    d.set_mask(pm)
    for z in range(12):
        d.submit(z, chunks[z])
    d.finish()
    stamped = np.where(mask[None], np.uint16(0xFFFF), stack)
    want = oracle_tables(stamped, range(12))
    got = d.results()
    assert same_tables(got, want)
    assert got[2]["signal_pixels"] < 8928  # the mask took pixels away
    d.close()
    pm.close()


def test_detector_mask_of_another_shape_is_refused(sf):
    """A mask whose shape is not the driver's frame shape is refused before
    anything is launched: by the binding (ValueError) and by the C entry
    point itself (SF_BAD_ARGUMENT), for a smaller mask (whose bit array
    the stamp would read past) and for one with the same pixel count but
    another width (which would mask the wrong pixels)."""
    d = sf.Driver(16, H, W, chunk=5)
    for shape in ((H // 2, W), (W, H)):
        pm = sf.PixelMask(np.zeros(shape, dtype=bool))
        with pytest.raises(ValueError):
            d.set_mask(pm)
        status = sf.lib().sf_hip_driver_set_mask(d.handle, pm.bits.ptr,
                                                 shape[0], shape[1])
        assert status == sf.SF_BAD_ARGUMENT, shape
        assert b"mask for" in sf.lib().sf_hip_driver_last_error()
        pm.close()
    # a mask of the right shape, then none, are accepted
    pm = sf.PixelMask(np.zeros((H, W), dtype=bool))
    d.set_mask(pm)
    d.set_mask(None)
    d.close()
    pm.close()
