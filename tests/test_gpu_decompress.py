"""Device decompression against the known-answer vectors and the oracle,
through the C ABI.

The frames decoded on the device must equal, byte for byte, the frames the
chunks were made from: the bitshuffle 0.3.5 vectors (tests/golden/
bslz4_vectors.npz), chunks the package's encoders write (bitshuffle's default
8192-byte blocks, blocks of up to 65536 bytes, ragged frames with leftovers)
and a full EIGER2 16M uint32 frame, whose decoded threshold must also equal the
threshold of the original.  Malformed chunks fail as the oracle fails (the
reference's decompress::image throws), and only the chunk at fault.
"""
import os
import struct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VECTORS = os.path.join(ROOT, "tests", "golden", "bslz4_vectors.npz")
DTYPES = {8: np.uint8, 16: np.uint16, 32: np.uint32}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def vectors():
    return np.load(VECTORS)


def cases(v):
    for name in v["cases"]:
        bd, h, w, block = (int(x) for x in v[f"{name}_meta"])
        yield str(name), bd, h, w, block, v[f"{name}_frame"].tobytes()


def decode_one(sf, data, algo, bd, h, w):
    d = sf.DeviceArray(max(sf.frame_bytes(h, w, bd), 8))
    d.copy_from(np.full(d.nbytes, 0xAB, np.uint8))
    sf.decompress_image(data, algo, bd, h, w, d)
    out = d.to_numpy(np.uint8, (d.nbytes,))[: sf.frame_bytes(h, w, bd)].tobytes()
    d.free()
    return out


def test_vectors_decode_on_device(device, vectors):
    for name, bd, h, w, block, frame in cases(vectors):
        for algo in ("bslz4", "lz4"):
            say("vector", name, algo)
            got = decode_one(device, vectors[f"{name}_{algo}"].tobytes(), algo, bd, h, w)
            assert got == frame, (name, algo)
        assert decode_one(device, frame, "", bd, h, w) == frame


def test_one_load_of_every_kind_twice(device, vectors):
    """All vectors (bslz4, lz4, raw) in one load, decoded twice from the
    resident chunks into frames scribbled over in between."""
    sf = device
    chunks, frames, outs = [], [], []
    for name, bd, h, w, block, frame in cases(vectors):
        for algo in ("bslz4", "lz4", ""):
            data = vectors[f"{name}_{algo}"].tobytes() if algo else frame
            d = sf.DeviceArray(max(len(frame), 8))
            outs.append(d)
            frames.append(frame)
            chunks.append((data, algo, bd, h, w, d))
    dec = sf.Decoder()
    say("load", len(chunks), "chunks")
    dec.load(chunks)
    for rep in range(2):
        for d in outs:
            d.copy_from(np.full(d.nbytes, 0x5A + rep, np.uint8))
        dec.run()
        assert dec.sync() == [0] * len(chunks)
        for d, f in zip(outs, frames):
            assert d.to_numpy(np.uint8, (d.nbytes,))[: len(f)].tobytes() == f
    dec.close()
    for d in outs:
        d.free()


@pytest.mark.parametrize("dtype,block,shape", [
    (np.uint32, 0, (301, 277)),       # 8192-byte blocks, a short last, leftovers
    (np.uint16, 0, (257, 129)),
    (np.uint8, 0, (97, 333)),
    (np.uint32, 16384, (300, 300)),   # 65536-byte blocks: a wave of its own
    (np.uint16, 8192, (129, 200)),    # 16384-byte blocks
    (np.uint32, 104, (33, 35)),       # rows of other than whole words
])
def test_encoded_frames_round_trip(device, oracle_lib, dtype, block, shape):
    sf = device
    from spotfinder_hip import codec
    rng = np.random.default_rng(7)
    f = rng.poisson(2.0, size=shape).astype(dtype)
    f[::17, ::5] = np.iinfo(dtype).max
    f[5, :] = np.iinfo(dtype).max - 1
    f[rng.integers(0, shape[0], 40), rng.integers(0, shape[1], 40)] = 1000 % np.iinfo(dtype).max
    bd = 8 * f.dtype.itemsize
    chunk = codec.bslz4_encode(f, block)
    rc, ref, msg = oracle_lib.decompress_image(chunk, "bslz4", bd, *shape)
    assert rc == 0 and ref == f.tobytes(), msg
    say("round trip", f.dtype, block, shape)
    assert decode_one(sf, chunk, "bslz4", bd, *shape) == f.tobytes()
    lz = codec.lz4_encode(f, 0)
    assert decode_one(sf, lz, "lz4", bd, *shape) == f.tobytes()


def test_16m_frame_decodes_and_thresholds(device):
    """A full EIGER2 16M uint32 frame: decoded on the device, it equals the
    original, and its threshold equals the original's."""
    sf = device
    from spotfinder_hip import codec
    desc = sf.synth_desc("16m", spots_per_frame=2000, persist=1, seed=2024)
    h, w = desc.height, desc.width
    f = sf.synth_frame(desc, 3)
    chunk = codec.bslz4_encode(f)
    say("16M chunk", len(chunk), "bytes for", f.nbytes)
    assert len(chunk) < f.nbytes // 4
    d = sf.DeviceArray(f.nbytes)
    sf.decompress_image(chunk, "bslz4", 32, h, w, d)
    assert np.array_equal(d.to_numpy(np.uint32, (h, w)), f)
    orig = sf.DeviceArray.from_numpy(f)
    ws = sf.Workspace()
    lists = []
    for src in (d, orig):
        ws.threshold(src.ptr, 32, h, w, "fused")
        rc, n = ws.sync()
        assert rc == 0
        lists.append(ws.fetch(n))
    ws.close()
    assert len(lists[0]) > 1000 and lists[0].tobytes() == lists[1].tobytes()
    d.free()
    orig.free()


def test_malformed_chunks_fail_like_the_oracle(device, oracle_lib):
    """Header checks fail on the host with the reference's messages; a
    malformed block fails on the device; only the chunk at fault fails."""
    sf = device
    from spotfinder_hip import codec
    f = (np.arange(96 * 80, dtype=np.uint32) % 13).reshape(96, 80)
    f[::7] = 4000
    good = codec.bslz4_encode(f)
    d = sf.DeviceArray(f.nbytes)
    for data, algo, bd in [
        (good, "bslz4", 16),                         # header size mismatch
        (good[:11], "bslz4", 32),                    # too short for a header
        (good, "zstd", 32),                          # unknown algorithm
        (f.tobytes()[:-4], "", 32),                  # raw size mismatch
        (good[:12] + struct.pack(">I", 0x00FFFFFF) + good[16:], "bslz4", 32),
    ]:
        rc, _, msg = oracle_lib.decompress_image(data, algo, bd, 96, 80)
        assert rc == -1
        with pytest.raises(sf.DecompressError) as e:
            sf.decompress_image(data, algo, bd, 96, 80, d)
        assert str(e.value) == msg
    # Corrupt LZ4 bytes inside the first block: the device finds them.
    first_len = struct.unpack(">I", good[12:16])[0]
    rng = np.random.default_rng(3)
    bad_chunks = []
    for k in range(12):
        b = bytearray(good)
        pos = 16 + int(rng.integers(0, first_len))
        b[pos] ^= int(rng.integers(1, 256))
        bad_chunks.append(bytes(b))
    outs = [sf.DeviceArray(f.nbytes) for _ in range(len(bad_chunks) + 1)]
    dec = sf.Decoder()
    dec.load([(c, "bslz4", 32, 96, 80, o) for c, o in zip(bad_chunks + [good], outs)])
    dec.run()
    status = dec.sync()
    say("corrupt statuses", status)
    for c, st, o in zip(bad_chunks, status, outs):
        rc, ref, msg = oracle_lib.decompress_image(c, "bslz4", 32, 96, 80)
        assert (st == 0) == (rc == 0), (st, rc, msg)
        if rc == 0:  # a flip the format cannot see decodes as the oracle decodes
            assert o.to_numpy(np.uint8, (f.nbytes,)).tobytes() == ref
    assert status[-1] == 0
    assert outs[-1].to_numpy(np.uint32, f.shape).tobytes() == f.tobytes()
    dec.close()
    for o in outs + [d]:
        o.free()
