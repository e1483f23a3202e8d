"""The decompression oracle and the chunk encoders, pinned without a GPU.

What pins what:
- the bit transpose: tests/golden/bslz4_vectors.npz, written by an independent
  build of bitshuffle 0.3.5 (imagecodecs; tests/golden/make_bslz4_vectors.py),
  which is the library src/spots/decompress.cc:111-121 calls;
- LZ4 block decoding: the system's liblz4 1.9.3 (LZ4_decompress_safe, the
  routine decompress.cc:58-61 calls and bitshuffle's own decoder), on valid
  blocks and on truncated and corrupted ones;
- the chunk framing and its checks: decompress.cc:76-133 (header, messages), and
  round trips through the encoders, which reproduce the fixture chunks byte for
  byte.
"""
import ctypes
import ctypes.util
import os
import struct
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dials-metal_amd"))

from oracle import oracle as O  # noqa: E402
from spotfinder_hip import codec  # noqa: E402

VECTORS = os.path.join(ROOT, "tests", "golden", "bslz4_vectors.npz")
DTYPES = {1: np.uint8, 2: np.uint16, 4: np.uint32}


@pytest.fixture(scope="module")
def vectors():
    return np.load(VECTORS)


@pytest.fixture(scope="module")
def orc():
    return O.oracle()


@pytest.fixture(scope="module")
def liblz4():
    lib = ctypes.CDLL(ctypes.util.find_library("lz4") or "liblz4.so.1")
    lib.LZ4_decompress_safe.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.c_int, ctypes.c_int]
    lib.LZ4_decompress_safe.restype = ctypes.c_int

    def decode(block, cap):
        out = ctypes.create_string_buffer(max(cap, 1))
        r = lib.LZ4_decompress_safe(block, out, len(block), cap)
        return r, out.raw[:max(r, 0)]
    return decode


def cases(v):
    for name in v["cases"]:
        bd, h, w, block = (int(x) for x in v[f"{name}_meta"])
        frame = v[f"{name}_frame"].tobytes()
        arr = np.frombuffer(frame, DTYPES[bd // 8]).reshape(h, w)
        yield str(name), bd, h, w, block, arr


def test_bit_transpose_matches_bitshuffle(vectors, orc):
    """The encoder's and the oracle's transpose, block by block, against
    bitshuffle's own output for the whole frame (blocks of `block` elements,
    the last rounded down to a multiple of 8, the leftover raw)."""
    for name, bd, h, w, block, arr in cases(vectors):
        elem = bd // 8
        raw = arr.tobytes()
        n = arr.size
        b = block or codec.default_block(elem)
        assert b == orc.bslz4_default_block(elem) or block
        want = vectors[f"{name}_shuffled"].tobytes()
        sizes = [b] * (n // b) + ([n % b // 8 * 8] if n % b >= 8 else [])
        got_codec, got_oracle, at = [], [], 0
        for count in sizes:
            piece = raw[at * elem:(at + count) * elem]
            got_codec.append(codec.bitshuffle(piece, elem))
            got_oracle.append(orc.bitshuffle(piece, count, elem))
            assert orc.bitunshuffle(got_oracle[-1], count, elem) == piece
            at += count
        tail = raw[at * elem:]
        assert b"".join(got_codec) + tail == want, name
        assert b"".join(got_oracle) + tail == want, name


def test_known_bit_rows():
    """0..15 as uint32: bit-row k holds bit k of every element, element i at
    bit i % 8 of byte i // 8 (the first rows alternate 0xAA, 0xCC, 0xF0)."""
    rows = codec.bitshuffle(np.arange(16, dtype=np.uint32).tobytes(), 4)
    assert rows[:8] == bytes([0xAA, 0xAA, 0xCC, 0xCC, 0xF0, 0xF0, 0x00, 0xFF])
    assert rows[8:] == bytes(56)


def test_encoders_reproduce_the_vectors(vectors):
    for name, bd, h, w, block, arr in cases(vectors):
        assert codec.bslz4_encode(arr, block) == vectors[f"{name}_bslz4"].tobytes(), name
        assert codec.lz4_encode(arr, 4096) == vectors[f"{name}_lz4"].tobytes(), name


def test_oracle_decodes_the_vectors(vectors, orc):
    for name, bd, h, w, block, arr in cases(vectors):
        for algo in ("bslz4", "lz4"):
            rc, out, msg = orc.decompress_image(vectors[f"{name}_{algo}"].tobytes(),
                                                algo, bd, h, w)
            assert rc == 0 and out == arr.tobytes(), (name, algo, msg)
        rc, out, _ = orc.decompress_image(arr.tobytes(), "", bd, h, w)
        assert rc == 0 and out == arr.tobytes()


def _blocks(rng, trials):
    for trial in range(trials):
        n = int(rng.integers(1, 3000))
        kind = trial % 4
        if kind == 0:
            data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        elif kind == 1:
            data = (rng.poisson(0.3, n) % 256).astype(np.uint8).tobytes()
        elif kind == 2:
            data = bytes(n)
        else:
            data = codec.bitshuffle(
                rng.poisson(1.0, (n // 8 + 1) * 8).astype(np.uint32).tobytes(), 4)
        yield data, codec.lz4_block(data)


def test_lz4_blocks_decode_as_liblz4_does(orc, liblz4):
    rng = np.random.default_rng(11)
    for data, blk in _blocks(rng, 200):
        for cap in (len(data), len(data) + 9):
            assert orc.lz4_decode_block(blk, cap) == liblz4(blk, cap)
            assert orc.lz4_decode_block(blk, cap)[1] == data
        if len(data) > 1:  # too small a capacity is an error for both
            assert orc.lz4_decode_block(blk, len(data) - 1)[0] < 0
            assert liblz4(blk, len(data) - 1)[0] < 0


def test_lz4_corrupt_blocks_fail_as_liblz4_does(orc, liblz4):
    """Truncations and single-byte edits: wherever liblz4 rejects a block the
    oracle does, and wherever both accept, the bytes agree.  The one rule the
    oracle adds is the block format's: offset 0 is malformed (liblz4 1.9.3
    copies from the output's own end instead)."""
    rng = np.random.default_rng(12)
    checked = 0
    for data, blk in _blocks(rng, 200):
        for c in range(6):
            bb = bytearray(blk)
            if c < 2 and len(bb) > 1:
                bb = bb[:int(rng.integers(1, len(bb)))]
            else:
                bb[int(rng.integers(0, len(bb)))] = int(rng.integers(0, 256))
            bb = bytes(bb)
            a = orc.lz4_decode_block(bb, len(data))
            b = liblz4(bb, len(data))
            if b[0] < 0:
                assert a[0] < 0
            elif a[0] >= 0:
                assert a == b
            else:
                assert _has_zero_offset(bb), bb
            checked += 1
    assert checked == 1200


def _has_zero_offset(blk):
    ip = 0
    while ip < len(blk):
        t = blk[ip]
        ip += 1
        lit = t >> 4
        if lit == 15:
            while ip < len(blk):
                b = blk[ip]
                ip += 1
                lit += b
                if b != 255:
                    break
        ip += lit
        if ip + 2 > len(blk):
            return False
        if blk[ip] == 0 and blk[ip + 1] == 0:
            return True
        ip += 2
        if t & 15 == 15:
            while ip < len(blk) and blk[ip] == 255:
                ip += 1
            ip += 1
    return False


def test_image_checks_and_messages(orc):
    """decompress.cc:76-133: each check, with its message."""
    f = (np.arange(64 * 32, dtype=np.uint32) % 7).reshape(64, 32)
    chunk = codec.bslz4_encode(f)
    ok = orc.decompress_image(chunk, "bslz4", 32, 64, 32)
    assert ok[0] == 0 and ok[1] == f.tobytes()
    assert orc.decompress_image(chunk, "bslz4", 4, 64, 32)[2] == \
        "output buffer too small for the frame"
    assert orc.decompress_image(f.tobytes()[:-1], "", 32, 64, 32)[2] == \
        "uncompressed frame is 8191 bytes, expected 8192"
    assert orc.decompress_image(chunk[:11], "bslz4", 32, 64, 32)[2] == \
        "too short for a filter header"
    assert orc.decompress_image(chunk, "bslz4", 16, 64, 32)[2] == \
        "filter header says 8192 bytes uncompressed, expected 4096"
    assert orc.decompress_image(chunk, "zstd", 32, 64, 32)[2] == \
        "unknown compression algorithm zstd"
    # a block length past the chunk, and a block that decodes short
    bad = bytearray(chunk)
    bad[12:16] = struct.pack(">I", 0x00FFFFFF)
    assert orc.decompress_image(bytes(bad), "bslz4", 32, 64, 32)[2].startswith(
        "bitshuffle decompression failed with code")
    lz = codec.lz4_encode(f, 4096)
    assert orc.decompress_image(lz[:-3], "lz4", 32, 64, 32)[2] in (
        "lz4 block size out of range", "lz4 data truncated")
    assert orc.decompress_image(lz[:12 + 4 + struct.unpack(">I", lz[12:16])[0]],
                                "lz4", 32, 64, 32)[2] == "lz4 data truncated"
    zero = bytearray(lz)
    zero[12:16] = bytes(4)
    assert orc.decompress_image(bytes(zero), "lz4", 32, 64, 32)[2] == \
        "lz4 block size out of range"


def test_block_size_rules(orc):
    """A header block that is not a multiple of 8 elements is bitshuffle's
    error -81; a zero block means bitshuffle's default."""
    f = np.arange(3000, dtype=np.uint16).reshape(30, 100)
    chunk = bytearray(codec.bslz4_encode(f))
    assert struct.unpack(">I", chunk[8:12])[0] == 8192
    chunk[8:12] = struct.pack(">I", 0)
    rc, out, _ = orc.decompress_image(bytes(chunk), "bslz4", 16, 30, 100)
    assert rc == 0 and out == f.tobytes()
    chunk[8:12] = struct.pack(">I", 2 * 12)
    assert orc.decompress_image(bytes(chunk), "bslz4", 16, 30, 100)[2] == \
        "bitshuffle decompression failed with code -81"
