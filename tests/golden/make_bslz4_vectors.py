"""Writes tests/golden/bslz4_vectors.npz: known-answer vectors for the
decompression path, made with an independent build of the bitshuffle library.

Run with the image's conda interpreter, which has imagecodecs (built with
bitshuffle 0.3.5, the library the reference's src/spots/decompress.cc calls
through bshuf_decompress_lz4):

    /opt/conda/bin/python3.9 tests/golden/make_bslz4_vectors.py

For each case the file holds:
  <case>_frame     the frame's bytes (little-endian elements)
  <case>_meta      [bit_depth, height, width, block elements (0 = default)]
  <case>_shuffled  imagecodecs.bitshuffle_encode(frame, blocksize=block): the
                   blocked bit transpose, leftover elements raw
  <case>_bslz4     the HDF5 bitshuffle-LZ4 chunk: 12-byte header, then per block
                   a big-endian length and liblz4's LZ4_compress_default of the
                   block's slice of <case>_shuffled, then the leftover raw
  <case>_lz4       the HDF5 LZ4 chunk of the frame, blocks of 4096 bytes
"""
import ctypes
import ctypes.util
import os
import struct

import numpy as np
from imagecodecs import _bitshuffle

HERE = os.path.dirname(os.path.abspath(__file__))

lz4 = ctypes.CDLL(ctypes.util.find_library("lz4") or "liblz4.so.1")
lz4.LZ4_compressBound.argtypes = [ctypes.c_int]
lz4.LZ4_compress_default.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.c_int, ctypes.c_int]


def lz4_block(data):
    cap = lz4.LZ4_compressBound(len(data))
    out = ctypes.create_string_buffer(max(cap, 1))
    n = lz4.LZ4_compress_default(data, out, len(data), cap)
    assert n > 0
    return out.raw[:n]


def default_block(elem):
    return max((8192 // elem) // 8 * 8, 128)


def frame_of(dtype, h, w, seed):
    rng = np.random.default_rng(seed)
    f = rng.poisson(1.5, size=(h, w)).astype(dtype)
    top = np.iinfo(dtype).max
    # spots and hot pixels reach the high bits; a masked column and corner
    f[rng.integers(0, h, 12), rng.integers(0, w, 12)] = rng.integers(
        100, top // 2, 12).astype(dtype)
    f[:, w // 3] = top
    f[: h // 4, -2:] = top - 1
    return np.ascontiguousarray(f)


CASES = [
    # name, dtype, height, width, block (elements, 0 = default)
    ("u32_two_full", np.uint32, 64, 64, 0),
    ("u32_ragged", np.uint32, 37, 61, 0),
    ("u32_tiny", np.uint32, 5, 3, 0),
    ("u32_small_blocks", np.uint32, 40, 50, 256),
    ("u16_one_short", np.uint16, 45, 91, 0),
    ("u16_two_full", np.uint16, 64, 128, 0),
    ("u16_odd_blocks", np.uint16, 33, 77, 1000),
    ("u8_ragged", np.uint8, 100, 83, 0),
]


def main():
    out = {}
    for i, (name, dtype, h, w, block) in enumerate(CASES):
        f = frame_of(dtype, h, w, 100 + i)
        elem = f.dtype.itemsize
        raw = f.tobytes()
        n = f.size
        b = block or default_block(elem)
        shuffled = bytes(np.ascontiguousarray(
            _bitshuffle.bitshuffle_encode(f, itemsize=elem, blocksize=block)
        ).view(np.uint8).reshape(-1))
        assert len(shuffled) == len(raw)
        parts = [struct.pack(">QI", len(raw), b * elem)]
        sizes = [b] * (n // b)
        last = (n % b) // 8 * 8
        if last:
            sizes.append(last)
        at = 0
        for count in sizes:
            packed = lz4_block(shuffled[at * elem:(at + count) * elem])
            parts += [struct.pack(">I", len(packed)), packed]
            at += count
        parts.append(shuffled[at * elem:])  # leftover elements, raw
        chunk = b"".join(parts)
        lz = [struct.pack(">QI", len(raw), 4096)]
        for s in range(0, len(raw), 4096):
            packed = lz4_block(raw[s:s + 4096])
            lz += [struct.pack(">I", len(packed)), packed]
        out[f"{name}_frame"] = np.frombuffer(raw, np.uint8)
        out[f"{name}_meta"] = np.array([8 * elem, h, w, block], np.int64)
        out[f"{name}_shuffled"] = np.frombuffer(shuffled, np.uint8)
        out[f"{name}_bslz4"] = np.frombuffer(chunk, np.uint8)
        out[f"{name}_lz4"] = np.frombuffer(b"".join(lz), np.uint8)
    out["cases"] = np.array([c[0] for c in CASES])
    out["library"] = np.array([_bitshuffle.bitshuffle_version()])
    np.savez_compressed(os.path.join(HERE, "bslz4_vectors.npz"), **out)
    print("wrote", len(CASES), "cases with", _bitshuffle.bitshuffle_version())


if __name__ == "__main__":
    main()
