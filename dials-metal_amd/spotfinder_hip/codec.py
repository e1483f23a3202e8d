"""Host-side encoders for the chunk formats the device decoder reads.

These write frames the way an EIGER's HDF5 filters store them, so that tests and
the benchmark can hand the decoder realistic chunks:

- ``bslz4_encode``: the bitshuffle filter with LZ4 (HDF5 filter 32008;
  ``src/spots/nxmx.cc:38, 173-174``).  A 12-byte header (the frame's byte count as
  a big-endian u64, the block's byte count as a big-endian u32), then per block
  of ``block`` elements (the last rounded down to a multiple of 8) a big-endian
  u32 compressed length and an LZ4 block of the bit-transposed elements, then
  the ``n % 8`` leftover elements raw.  This is what ``decompress::image`` hands
  to ``bshuf_decompress_lz4`` (``src/spots/decompress.cc:111-121``).
- ``lz4_encode``: the HDF5 LZ4 filter: the same 12-byte header, then per block
  of ``block_bytes`` a big-endian u32 compressed length and an LZ4 block
  (``src/spots/decompress.cc:43-72``).

LZ4 blocks are compressed with the system's liblz4 (``LZ4_compress_default``).
The bit transpose follows bitshuffle: bit-row ``8j + k`` holds bit ``k`` of byte
``j`` of every element of the block, element ``i`` at bit ``i % 8`` of byte
``i // 8``.
"""
import ctypes
import ctypes.util
import struct

import numpy as np

_lz4 = None


def _lib():
    global _lz4
    if _lz4 is None:
        name = ctypes.util.find_library("lz4") or "liblz4.so.1"
        lib = ctypes.CDLL(name)
        lib.LZ4_compressBound.argtypes = [ctypes.c_int]
        lib.LZ4_compressBound.restype = ctypes.c_int
        lib.LZ4_compress_default.argtypes = [
            ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        lib.LZ4_compress_default.restype = ctypes.c_int
        _lz4 = lib
    return _lz4


def lz4_block(data):
    """One LZ4 block of ``data`` (bytes), as LZ4_compress_default writes it."""
    data = bytes(data)
    lib = _lib()
    cap = lib.LZ4_compressBound(len(data))
    out = ctypes.create_string_buffer(max(cap, 1))
    n = lib.LZ4_compress_default(data, out, len(data), cap)
    if n <= 0:
        raise RuntimeError("LZ4_compress_default failed")
    return out.raw[:n]


def default_block(elem):
    """bitshuffle's default block: 8192 bytes of elements, a multiple of 8,
    at least 128 elements."""
    b = (8192 // elem) // 8 * 8
    return max(b, 128)


def bitshuffle(block, elem):
    """The bit transpose of one block: ``block`` is uint8 of n * elem bytes,
    n a multiple of 8."""
    a = np.frombuffer(bytes(block), np.uint8)
    n = a.size // elem
    assert n * elem == a.size and n % 8 == 0
    bits = np.unpackbits(a.reshape(n, elem), axis=1, bitorder="little")  # (n, 8 elem)
    return np.packbits(bits.T, axis=1, bitorder="little").tobytes()


def _raw_bytes(frame):
    frame = np.ascontiguousarray(frame)
    return frame.view(np.uint8).reshape(-1), frame.dtype.itemsize


def bslz4_encode(frame, block=0):
    """A bslz4 chunk of ``frame`` (a uint8/16/32 array); ``block`` in elements,
    0 for bitshuffle's default."""
    raw, elem = _raw_bytes(frame)
    n = raw.size // elem
    block = block or default_block(elem)
    if block % 8:
        raise ValueError("a bitshuffle block is a multiple of 8 elements")
    parts = [struct.pack(">QI", raw.size, block * elem)]
    done = 0
    sizes = [block] * (n // block)
    last = (n % block) // 8 * 8
    if last:
        sizes.append(last)
    for count in sizes:
        packed = lz4_block(bitshuffle(raw[done * elem:(done + count) * elem], elem))
        parts.append(struct.pack(">I", len(packed)))
        parts.append(packed)
        done += count
    parts.append(raw[done * elem:].tobytes())
    return b"".join(parts)


def lz4_encode(frame, block_bytes=0):
    """An HDF5-LZ4 chunk of ``frame``; ``block_bytes`` 0 writes one block."""
    raw, _ = _raw_bytes(frame)
    total = raw.size
    bs = block_bytes or total
    parts = [struct.pack(">QI", total, block_bytes)]
    for at in range(0, total, bs):
        packed = lz4_block(raw[at:at + bs].tobytes())
        parts.append(struct.pack(">I", len(packed)))
        parts.append(packed)
    return b"".join(parts)
