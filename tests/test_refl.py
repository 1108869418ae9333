"""The .refl writer (sf_hip_refl_write, the reference's refl::write,
src/refl.hh:21-78): the same bytes as the reference's own writer
(oracle/_ref, where built) on the same tables, and a structure and value
check through an independent msgpack decoder. Host code only: no GPU."""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import cpu
from test_oracle import random_sequence

import spotfinder as sf

msgpack = pytest.importorskip("msgpack")

COLUMNS = ["bbox", "flags", "id", "intensity.sum.value",
           "intensity.sum.variance", "n_signal", "panel", "shoebox",
           "xyzobs.px.value", "xyzobs.px.variance"]
TYPES = {"bbox": "int6", "flags": "std::size_t", "id": "int",
         "intensity.sum.value": "double", "intensity.sum.variance": "double",
         "n_signal": "int", "panel": "std::size_t", "shoebox": "Shoebox<>",
         "xyzobs.px.value": "vec3<double>", "xyzobs.px.variance": "vec3<double>"}


def labelled(seed, h, w, frames, density, **options):
    lab = cpu.Labeller(h, w, **options)
    for z, p in random_sequence(np.random.default_rng(seed), h, w, frames,
                                density, gap_every=3):
        lab.add(z, p)
    lab.finish()
    spots, pixels, _ = lab.results()
    return spots, pixels


def many_small_spots(n, w):
    """n one-pixel spots on distinct pixels: more rows than any 16-bit
    msgpack length holds."""
    spots = np.zeros(n, dtype=sf.SPOT)
    pixels = np.zeros(n, dtype=sf.PIXEL)
    idx = np.arange(n, dtype=np.uint32) * 3
    pixels["index"] = idx
    pixels["frame"] = np.arange(n) % 7
    pixels["value"] = np.arange(n) % 1000 + 1
    spots["bbox"][:, 0] = idx % w
    spots["bbox"][:, 1] = idx % w + 1
    spots["bbox"][:, 2] = idx // w
    spots["bbox"][:, 3] = idx // w + 1
    spots["bbox"][:, 4] = pixels["frame"]
    spots["bbox"][:, 5] = pixels["frame"] + 1
    spots["n_signal"] = 1
    spots["first"] = np.arange(n)
    spots["intensity"] = pixels["value"]
    spots["intensity_variance"] = pixels["value"]
    spots["position"] = np.random.default_rng(2).random((n, 3))
    spots["variance"] = 1.0 / 12.0
    return spots, pixels


CASES = [
    ("labelled", lambda: labelled(3, 40, 50, 8, 0.2), 50,
     dict(panel=0, id=0, identifier="", shoeboxes=True)),
    ("labelled, no boxes", lambda: labelled(4, 40, 50, 8, 0.2), 50,
     dict(panel=3, id=7, identifier="abc", shoeboxes=False)),
    ("long identifier, negative id", lambda: labelled(5, 30, 70, 6, 0.3), 70,
     dict(panel=1, id=-2, identifier="x" * 300, shoeboxes=True)),
    ("no spots", lambda: (np.zeros(0, sf.SPOT), np.zeros(0, sf.PIXEL)), 9,
     dict(panel=0, id=0, identifier="i" * 40, shoeboxes=True)),
    ("70,000 rows", lambda: many_small_spots(70000, 1000), 1000,
     dict(panel=0, id=1, identifier="e", shoeboxes=True)),
]


@pytest.mark.parametrize("name,make,width,options", CASES,
                         ids=[c[0] for c in CASES])
def test_bytes_equal_the_reference_writer(tmp_path, name, make, width,
                                          options):
    if not cpu.have_ref():
        pytest.skip("the reference's own writer (oracle/_ref) is not built")
    spots, pixels = make()
    ours, theirs = tmp_path / "ours.refl", tmp_path / "theirs.refl"
    sf.write_refl(str(ours), spots, pixels, width, **options)
    assert cpu.ref_refl_write(str(theirs), spots, pixels, width,
                              **options) == 0
    assert ours.read_bytes() == theirs.read_bytes()


@pytest.mark.parametrize("name,make,width,options", CASES,
                         ids=[c[0] for c in CASES])
def test_structure_and_values_decode(tmp_path, name, make, width, options):
    spots, pixels = make()
    path = tmp_path / "t.refl"
    sf.write_refl(str(path), spots, pixels, width, **options)
    doc = msgpack.unpackb(path.read_bytes(), raw=False, strict_map_key=False)
    assert doc[0] == "dials::af::reflection_table" and doc[1] == 2
    table = doc[2]
    assert list(table) == ["identifiers", "nrows", "data"]
    ident = options["identifier"]
    assert table["identifiers"] == (
        {options["id"] % 2 ** 64: ident} if ident else {})
    n = len(spots)
    assert table["nrows"] == n
    names = [c for c in COLUMNS if c != "shoebox" or options["shoeboxes"]]
    assert list(table["data"]) == names
    for c in names:
        kind, (rows, blob) = table["data"][c]
        assert kind == TYPES[c] and rows == n, c
    col = {c: table["data"][c][1][1] for c in names}
    assert np.array_equal(np.frombuffer(col["bbox"], "<i4").reshape(-1, 6),
                          spots["bbox"])
    assert np.all(np.frombuffer(col["flags"], "<u8") == 32)  # Strong
    assert np.all(np.frombuffer(col["id"], "<i4") == options["id"])
    assert np.array_equal(np.frombuffer(col["intensity.sum.value"], "<f8"),
                          spots["intensity"])
    assert np.array_equal(np.frombuffer(col["intensity.sum.variance"], "<f8"),
                          spots["intensity_variance"])
    assert np.array_equal(np.frombuffer(col["n_signal"], "<i4"),
                          spots["n_signal"])
    assert np.all(np.frombuffer(col["panel"], "<u8") == options["panel"])
    assert np.array_equal(
        np.frombuffer(col["xyzobs.px.value"], "<f8").reshape(-1, 3),
        spots["position"])
    assert np.array_equal(
        np.frombuffer(col["xyzobs.px.variance"], "<f8").reshape(-1, 3),
        spots["variance"])
    if options["shoeboxes"]:
        check_shoeboxes(col["shoebox"], spots, pixels, width,
                        options["panel"])


def check_shoeboxes(blob, spots, pixels, width, panel):
    """Each box: panel, bbox, version 2, then float data, byte mask and
    float background over (z, y, x), with the spot's pixels placed."""
    at = 0
    for s in spots[:200]:
        x0, x1, y0, y1, z0, z1 = (int(v) for v in s["bbox"])
        nx, ny, nz = x1 - x0, y1 - y0, z1 - z0
        m = nx * ny * nz
        assert np.frombuffer(blob, "<u4", 1, at)[0] == panel
        assert list(np.frombuffer(blob, "<i4", 6, at + 4)) == list(s["bbox"])
        assert blob[at + 28] == 2
        data = np.frombuffer(blob, "<f4", m, at + 29).reshape(nz, ny, nx)
        mask = np.frombuffer(blob, "u1", m, at + 29 + 4 * m).reshape(nz, ny, nx)
        back = np.frombuffer(blob, "<f4", m, at + 29 + 5 * m)
        want = np.zeros((nz, ny, nx), np.float32)
        code = np.zeros((nz, ny, nx), np.uint8)
        for p in pixels[s["first"]:s["first"] + s["n_signal"]]:
            z, y, x = p["frame"] - z0, p["index"] // width - y0, \
                p["index"] % width - x0
            want[z, y, x] = p["value"]
            code[z, y, x] = 5  # valid | foreground
        assert np.array_equal(data, want) and np.array_equal(mask, code)
        assert not back.any()
        at += 29 + 9 * m


def test_header_widths_are_the_shortest():
    """msgpack's shortest forms at every width boundary, against the
    independent decoder: str 31/32/255/256, uint 127/128/255/256/65535/65536,
    bin lengths through the shoebox column of one growing spot."""
    for length in (0, 31, 32, 255, 256, 65536):
        ident = "q" * length
        spots, pixels = many_small_spots(3, 10)
        out = os.path.join(os.environ.get("TMPDIR", "/tmp"),
                           f"sf_refl_width_{os.getpid()}.refl")
        try:
            for id_ in (0, 127, 128, 255, 256, 65535, 65536, 2 ** 31 - 1):
                sf.write_refl(out, spots, pixels, 10, id=id_,
                              identifier=ident)
                raw = open(out, "rb").read()
                doc = msgpack.unpackb(raw, raw=False, strict_map_key=False)
                assert doc[2]["identifiers"] == ({id_: ident} if ident
                                                 else {})
                key = msgpack.packb(id_) + msgpack.packb(ident)
                assert (key in raw) == bool(ident)
        finally:
            if os.path.exists(out):
                os.remove(out)


def test_errors(tmp_path):
    spots, pixels = labelled(6, 20, 30, 4, 0.3)
    with pytest.raises(ValueError):
        sf.write_refl(str(tmp_path / "a.refl"), spots, pixels, 0)
    bad = spots.copy()
    bad["first"][0] = len(pixels)  # pixels past the table
    with pytest.raises(ValueError):
        sf.write_refl(str(tmp_path / "b.refl"), bad, pixels, 30)
    assert not (tmp_path / "b.refl").exists()
    # without shoeboxes the pixel table is not read, as in the reference
    sf.write_refl(str(tmp_path / "c.refl"), bad, pixels, 30, shoeboxes=False)
    with pytest.raises(OSError):
        sf.write_refl(str(tmp_path / "no" / "such" / "dir.refl"), spots,
                      pixels, 30)
