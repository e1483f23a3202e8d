"""The reflection-table writer (sf_refl_*), against the reference's own writer,
without a GPU.

- The reference's golden test (tests/spots/test_refl_golden.cc:33-118): its
  fixed spots, written with shoeboxes and an identifier, are 2026 bytes with
  FNV-1a checksum 11545107909170840019.
- Byte equality with the reference's refl::write (src/spots/refl.cc, compiled
  into oracle/_ref) on the reference Labeller's own 2D and 3D spot lists of
  the golden sweep, with and without shoeboxes and identifiers, and at every
  msgpack size class the headers can take (rows, identifiers and columns
  across 128, 256, 65536 and the 32-character string limit).
- An independent decode with the msgpack library: the document's structure
  and every column read back as the spots' own values.
"""
import os

import msgpack
import numpy as np
import pytest

import spotfinder_hip as sf
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_spots(width=64):
    """test_refl_golden.cc:33-70's fixed spots and pixels."""
    spots, pixels = [], []
    at = 0
    for s in range(5):
        bbox = [12 + s * 5, 12 + s * 5 + 3, 10 + s * 7, 10 + s * 7 + 3, s, s + 2 + s % 2]
        first, n, intensity = at, 0, 0.0
        for z in range(2 + s % 2):
            for y in range(3):
                for x in range(3):
                    v = 3 + (x + y + z + s) % 7
                    pixels.append((z + s, (10 + s * 7 + y) * width + 12 + s * 5 + x, v))
                    at += 1
                    n += 1
                    intensity += v
        spots.append((bbox, (bbox[0] + 1.5, bbox[2] + 1.5, s + 0.5), (0.4, 0.5, 0.6),
                      intensity, intensity, n, first))
    return np.array(spots, sf.SPOT), np.array(pixels, sf.PIXEL), width


def fnv1a64(data):
    h = 1469598103934665603
    for c in data:
        h ^= c
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.fixture(scope="module")
def ref():
    r = O.reference()
    if r is None:
        pytest.fail("oracle/_ref is not built: run __graft_entry__.build()")
    return r


@pytest.fixture(scope="module")
def sweep():
    return dict(np.load(os.path.join(GOLDEN, "labeller_sweep.npz")))


def test_the_reference_golden_table(ref):
    spots, pixels, width = golden_spots()
    ident = "0123abcd-0000-1111-2222-333344445555"
    ours = sf.refl_encode(spots, pixels, width, identifier=ident)
    assert len(ours) == 2026
    assert fnv1a64(ours) == 11545107909170840019
    assert ref.refl_bytes(spots, pixels, width, identifier=ident) == ours


@pytest.mark.parametrize("kind", ["2d", "3d"])
@pytest.mark.parametrize("shoeboxes", [True, False])
@pytest.mark.parametrize("ident,id,panel", [("", 0, 0), ("expt-7", 7, 0),
                                            ("x" * 40, -3, 5), ("y" * 300, 200, 1 << 33)])
def test_equals_the_reference_writer(ref, sweep, kind, shoeboxes, ident, id, panel):
    spots, pixels = sweep[f"spots_{kind}"], sweep[f"pixels_{kind}"]
    width = int(sweep["desc_width"])
    kw = dict(panel=panel, id=id, identifier=ident, shoeboxes=shoeboxes)
    assert sf.refl_encode(spots, pixels, width, **kw) == \
        ref.refl_bytes(spots, pixels, width, **kw)


@pytest.mark.parametrize("rows", [0, 1, 5, 127, 128, 255, 256, 2731, 65536])
def test_size_classes(ref, sweep, rows):
    """Rows and column lengths across each msgpack form (fixint, uint8/16/32,
    bin8/16/32): the golden sweep's spots repeated to `rows`."""
    base, pixels = sweep["spots_2d"], sweep["pixels_2d"]
    width = int(sweep["desc_width"])
    spots = np.resize(base, rows)
    for shoeboxes in (False, True) if rows <= 2731 else (False,):
        assert sf.refl_encode(spots, pixels, width, shoeboxes=shoeboxes) == \
            ref.refl_bytes(spots, pixels, width, shoeboxes=shoeboxes)


def test_decodes_as_the_spots(sweep, tmp_path):
    spots, pixels = sweep["spots_3d"], sweep["pixels_3d"]
    width = int(sweep["desc_width"])
    path = tmp_path / "spots.refl"
    sf.refl_write(path, spots, pixels, width, id=2, identifier="abc")
    doc = msgpack.unpackb(path.read_bytes(), raw=False, strict_map_key=False)
    assert doc[0] == "dials::af::reflection_table" and doc[1] == 2
    table = doc[2]
    assert table["identifiers"] == {2: "abc"} and table["nrows"] == len(spots)
    data = table["data"]
    assert list(data) == sorted(data) and len(data) == 10
    col = {k: (t, n, blob) for k, (t, (n, blob)) in data.items()}
    assert all(n == len(spots) for _, n, _ in col.values())
    np.testing.assert_array_equal(
        np.frombuffer(col["bbox"][2], np.int32).reshape(-1, 6), spots["bbox"])
    np.testing.assert_array_equal(np.frombuffer(col["flags"][2], np.uint64), 32)
    np.testing.assert_array_equal(np.frombuffer(col["id"][2], np.int32), 2)
    np.testing.assert_array_equal(np.frombuffer(col["intensity.sum.value"][2]),
                                  spots["intensity"])
    np.testing.assert_array_equal(np.frombuffer(col["n_signal"][2], np.int32),
                                  spots["n_signal"])
    np.testing.assert_array_equal(
        np.frombuffer(col["xyzobs.px.value"][2]).reshape(-1, 3), spots["position"])
    # The shoeboxes: each spot's pixels are where its box says, valid and
    # foreground, and nothing else is set.
    blob, at = col["shoebox"][2], 0
    for s in spots:
        nx, ny, nz = (s["bbox"][1] - s["bbox"][0], s["bbox"][3] - s["bbox"][2],
                      s["bbox"][5] - s["bbox"][4])
        n = nx * ny * nz
        assert np.frombuffer(blob, np.int32, 6, at + 4).tolist() == s["bbox"].tolist()
        assert blob[at + 28] == 2
        dat = np.frombuffer(blob, np.float32, n, at + 29).reshape(nz, ny, nx)
        msk = np.frombuffer(blob, np.uint8, n, at + 29 + 4 * n).reshape(nz, ny, nx)
        bg = np.frombuffer(blob, np.float32, n, at + 29 + 5 * n)
        px = pixels[s["first"]: s["first"] + s["n_signal"]]
        z = px["frame"] - s["bbox"][4]
        y = px["index"] // width - s["bbox"][2]
        x = px["index"] % width - s["bbox"][0]
        assert np.array_equal(dat[z, y, x], px["value"].astype(np.float32))
        assert (msk != 0).sum() == len(px) and np.all(msk[z, y, x] == 5)
        assert not bg.any()
        at += 29 + 9 * n
    assert at == len(blob)


def test_inconsistent_spots_fail(sweep):
    spots, pixels = sweep["spots_2d"].copy(), sweep["pixels_2d"]
    width = int(sweep["desc_width"])
    with pytest.raises(RuntimeError, match="width of zero"):
        sf.refl_encode(spots, pixels, 0)
    bad = spots.copy()
    bad[3]["bbox"][1] = bad[3]["bbox"][0]  # an empty box cannot hold its pixels
    with pytest.raises(RuntimeError, match="outside its box"):
        sf.refl_encode(bad, pixels, width)
    assert sf.refl_encode(bad, pixels, width, shoeboxes=False)
