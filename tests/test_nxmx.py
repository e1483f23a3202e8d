"""The NXmx chunk reader (libspotfinder_nxmx.so), without a device.

Series are written here the way a detector's writer writes them -- data files
holding one frame per chunk through H5Dwrite_chunk, and a master whose
/entry/data/data is a virtual dataset over them (tests/cpp/nxmx_writer.c) --
with chunks made by the package's bslz4 encoder.  The reader must give back
every chunk byte for byte with its shape, bit depth and algorithm, resolve
the mapping into the reference's blocks (src/spots/nxmx.cc:433-497: relative
names against the master's directory, "." as the master itself, source
offsets, unlimited mappings clipped to the extent), report a chunk never
written as absent (nxmx.cc:213-216), treat a skipped filter as a raw chunk
(nxmx.cc:225-229), and refuse what the reference refuses with its messages.
The chunks it reads decode, through the oracle's decompress::image, to the
frames they were made from."""
import ctypes
import os
import threading

import numpy as np
import pytest

from conftest import ROOT

WRITER = os.path.join(ROOT, "tests", "cpp", "libnxmx_writer.so")


@pytest.fixture(scope="module")
def writer():
    if not os.path.exists(WRITER):
        pytest.skip("the NXmx test writer is not built (no HDF5 headers)")
    W = ctypes.CDLL(WRITER)
    u64, vp = ctypes.c_uint64, ctypes.c_void_p
    W.nxw_write_data.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, u64,
                                 u64, u64, ctypes.c_uint, ctypes.c_uint, u64, vp, vp, vp]
    W.nxw_write_contiguous.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u64, u64, u64]
    W.nxw_write_master.argtypes = [ctypes.c_char_p, ctypes.c_int, u64, u64, u64,
                                   ctypes.c_uint, ctypes.c_size_t, vp, vp, vp, vp, vp]
    return W


@pytest.fixture(scope="module")
def sfn(sf):
    try:
        sf.nxmx_lib()
    except ImportError as e:
        pytest.skip(str(e))
    return sf


def write_data(W, path, dataset, chunks, h, w, bits, filt=32008, masks=None,
               append=False, frames=None, chunk_frames=1):
    n = len(chunks)
    frames = n if frames is None else frames
    bufs = [ctypes.create_string_buffer(c, len(c)) if c else None for c in chunks]
    ptrs = (ctypes.c_void_p * max(1, n))(*[ctypes.cast(b, ctypes.c_void_p) if b else None
                                            for b in bufs])
    sizes = (ctypes.c_uint64 * max(1, n))(*[len(c) if c else 0 for c in chunks])
    m = (ctypes.c_uint * max(1, n))(*(masks or [0] * n))
    assert W.nxw_write_data(str(path).encode(), dataset.encode(), int(append), frames,
                            h, w, bits, filt, chunk_frames, ptrs, sizes, m) == 0


def write_master(W, path, images, h, w, bits, maps, append=False):
    n = len(maps)
    files = (ctypes.c_char_p * n)(*[m[0].encode() for m in maps])
    dsets = (ctypes.c_char_p * n)(*[m[1].encode() for m in maps])
    firsts = (ctypes.c_uint64 * n)(*[m[2] for m in maps])
    counts = (ctypes.c_uint64 * n)(*[m[3] for m in maps])
    offsets = (ctypes.c_uint64 * n)(*[m[4] for m in maps])
    assert W.nxw_write_master(str(path).encode(), int(append), images, h, w, bits, n,
                              files, dsets, firsts, counts, offsets) == 0


def frames_of(sf, n, h, w, bits, seed):
    d = sf.synth_desc((w, h), bits=bits, seed=seed, spots_per_frame=30,
                      ring_radius=min(h, w) / 3)
    return [sf.synth_frame(d, z) for z in range(n)]


@pytest.fixture()
def series(tmp_path, writer, sfn):
    """Ten frames of 97 x 130 uint32 over three data files, the
    third read from source offset 2, frame 5 never written, frame 7 stored
    raw (filter skipped); the master in a subdirectory naming the data files
    relative to itself."""
    from spotfinder_hip import codec
    sf = sfn
    h, w, bits = 97, 130, 32
    frames = frames_of(sf, 10, h, w, bits, seed=8)
    chunks = [codec.bslz4_encode(f) for f in frames]
    chunks[7] = frames[7].tobytes()
    d = tmp_path / "run"
    d.mkdir()
    # data_000001: frames 0-3; data_000002: frames 4-6 (5 unwritten);
    # data_000003: two frames of padding, then frames 7-9
    write_data(writer, d / "data_000001.h5", "/entry/data/data", chunks[0:4], h, w, bits)
    write_data(writer, d / "data_000002.h5", "/entry/data/data",
               [chunks[4], None, chunks[6]], h, w, bits)
    write_data(writer, d / "data_000003.h5", "/entry/data/data",
               [chunks[0], chunks[1]] + chunks[7:10], h, w, bits,
               masks=[0, 0, 1, 0, 0])
    master = d / "master.h5"
    write_master(writer, master, 10, h, w, bits,
                 [("data_000001.h5", "/entry/data/data", 0, 4, 0),
                  ("data_000002.h5", "/entry/data/data", 4, 3, 0),
                  ("data_000003.h5", "/entry/data/data", 7, 3, 2)])
    return master, frames, chunks


def test_blocks_and_info(series, sfn):
    master, frames, _ = series
    s = sfn.NxmxSeries(master)
    assert (s.images, s.height, s.width) == (10, 97, 130)
    d = os.path.dirname(str(master))
    assert s.blocks() == [
        {"filename": f"{d}/data_000001.h5", "dataset": "/entry/data/data",
         "first": 0, "last": 3, "offset": 0},
        {"filename": f"{d}/data_000002.h5", "dataset": "/entry/data/data",
         "first": 4, "last": 6, "offset": 0},
        {"filename": f"{d}/data_000003.h5", "dataset": "/entry/data/data",
         "first": 7, "last": 9, "offset": 2}]


def test_every_chunk_byte_for_byte(series, sfn, oracle_lib):
    master, frames, chunks = series
    s = sfn.NxmxSeries(master)
    for z in range(10):
        got = s.read(z)
        if z == 5:
            assert got is None and s.frame_info(5) is None
            continue
        data, meta = got
        assert data == chunks[z], z
        assert meta["number"] == z and meta["bit_depth"] == 32
        assert (meta["height"], meta["width"]) == (97, 130)
        assert meta["algorithm"] == ("" if z == 7 else "bslz4"), z
        rc, out, msg = oracle_lib.decompress_image(data, meta["algorithm"], 32, 97, 130)
        assert rc == 0 and out == frames[z].tobytes(), (z, msg)


def test_read_chunks_for_the_decoder(series, sfn):
    master, frames, chunks = series
    s = sfn.NxmxSeries(master)
    arr, n, buf, present = s.read_chunks(3, [1000 + i for i in range(6)])
    assert present == [1, 1, 0, 1, 1, 1] and n == 5
    want = [3, 4, 6, 7, 8]
    base = ctypes.addressof(buf)
    at = 0
    for k, z in enumerate(want):
        c = arr[k]
        assert c.data == base + at and c.size == len(chunks[z])
        assert ctypes.string_at(c.data, c.size) == chunks[z]
        assert (c.bit_depth, c.height, c.width) == (32, 97, 130)
        assert c.algorithm == (b"" if z == 7 else b"bslz4")
        assert c.frame == 1000 + (z - 3)
        at += c.size


def test_frame_outside_every_mapping(series, sfn):
    master, _, _ = series
    s = sfn.NxmxSeries(master)
    with pytest.raises(ValueError, match="outside every virtual mapping"):
        s.read(10)


def test_dot_mapping_unlimited_and_plain_dataset(tmp_path, writer, sfn):
    """A mapping into the master itself ("."), named by a relative path with
    a directory in it (nxmx.cc:455-467); an unlimited mapping, clipped to the
    dataset's extent; and a master whose images are a plain chunked dataset."""
    from spotfinder_hip import codec
    sf = sfn
    h, w = 40, 64
    frames = frames_of(sf, 6, h, w, 16, seed=3)
    chunks = [codec.bslz4_encode(f) for f in frames]
    d = tmp_path / "sub"
    d.mkdir()
    master = d / "self.h5"
    write_data(writer, master, "/entry/data/data_000001", chunks[:3], h, w, 16)
    write_data(writer, d / "more.h5", "/entry/data/data", chunks[3:], h, w, 16)
    write_master(writer, master, 6, h, w, 16,
                 [(".", "/entry/data/data_000001", 0, 3, 0),
                  ("more.h5", "/entry/data/data", 3, 0, 0)], append=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        s = sf.NxmxSeries(os.path.join("sub", "self.h5"))
        b = s.blocks()
        assert b[0]["filename"] == os.path.join("sub", "self.h5")
        assert b[1]["filename"] == os.path.join("sub", "more.h5")
        assert (b[1]["first"], b[1]["last"]) == (3, 5)
        for z in range(6):
            data, meta = s.read(z)
            assert data == chunks[z] and meta["bit_depth"] == 16
        s.close()
    finally:
        os.chdir(cwd)
    plain = tmp_path / "plain.h5"
    write_data(writer, plain, "/entry/data/data", chunks, h, w, 16, filt=32004)
    s = sf.NxmxSeries(plain)
    assert s.blocks() == [{"filename": str(plain), "dataset": "/entry/data/data",
                           "first": 0, "last": 5, "offset": 0}]
    data, meta = s.read(4)
    assert data == chunks[4] and meta["algorithm"] == "lz4"


def test_refusals(tmp_path, writer, sfn):
    sf = sfn
    with pytest.raises(FileNotFoundError):
        sf.NxmxSeries(tmp_path / "not_yet.h5")
    text = tmp_path / "list.expt"
    text.write_text('  {"experiment": []}')
    with pytest.raises(sf.NxmxError, match="is not an HDF5 file: it looks like JSON"):
        sf.NxmxSeries(text)
    other = tmp_path / "notes.txt"
    other.write_text("frames")
    with pytest.raises(sf.NxmxError, match=r"is not an HDF5 file$"):
        sf.NxmxSeries(other)
    flat = tmp_path / "flat.h5"
    assert writer.nxw_write_contiguous(str(flat).encode(), b"/entry/data/data", 2, 8, 8) == 0
    s = sf.NxmxSeries(flat)
    with pytest.raises(sf.NxmxError, match="is not chunked, so it cannot be chunk-read"):
        s.read(0)
    multi = tmp_path / "multi.h5"
    write_data(writer, multi, "/entry/data/data", [], 8, 8, 32, filt=0, frames=4,
               chunk_frames=2)
    s = sf.NxmxSeries(multi)
    with pytest.raises(sf.NxmxError, match="span 2 frames; only one frame per chunk"):
        s.read(0)


def test_threads_share_a_series(series, sfn):
    """Readers on four threads at once (every HDF5 call is serialised inside
    the library) get every chunk intact."""
    master, _, chunks = series
    s = sfn.NxmxSeries(master)
    errors = []

    def work(t):
        for rep in range(20):
            for z in (0, 1, 2, 3, 4, 6, 7, 8, 9):
                data, _ = s.read(z)
                if data != chunks[z]:
                    errors.append((t, rep, z))

    ts = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors
