"""ctypes bindings for the oracle (liboracle.so) and for the reference's own
CPU code compiled into oracle/_ref/libspots_ref.so.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU
baseline import this module; the product never does.

Reference interfaces these mirror (paths in the reference repository):
  dext<T>                     src/spots/dext.hh:96-107
  signal_order::by_index      src/spots/signal_order.hh:62-66
  dials_spots::Labeller       src/spots/dials_spots.hh:108-158
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SIGNAL_PIXEL = np.dtype(
    [("index", "<u4"), ("value", "<u4"), ("background", "<f4"),
     ("population", "<u2"), ("reserved", "<u2")])
PIXEL = np.dtype([("frame", "<i4"), ("index", "<u4"), ("value", "<u4")])
SPOT = np.dtype(
    [("bbox", "<i4", (6,)), ("position", "<f8", (3,)), ("variance", "<f8", (3,)),
     ("intensity", "<f8"), ("intensity_variance", "<f8"), ("n_signal", "<u4"),
     ("first", "<u8")], align=True)
COUNTS_FIELDS = ("signal_pixels", "groups", "too_small", "too_large",
                 "separated", "accepted")
assert SIGNAL_PIXEL.itemsize == 16 and PIXEL.itemsize == 12
assert SPOT.itemsize == 104


class _Counts(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in COUNTS_FIELDS]


class _Scan(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in ("largest", "bad", "joins", "over")]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Lib:
    """Shared loader for the two libraries, which expose the same surface
    under the prefixes oracle_ and ref_."""

    def __init__(self, path, prefix):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        L, P = self.lib, prefix
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        for t in ("u16", "u32"):
            f = getattr(L, f"{P}dext_{t}")
            f.argtypes = [vp, sz, sz, vp, sz, ctypes.POINTER(sz)]
            f.restype = ctypes.c_int
        getattr(L, f"{P}sort_by_index").argtypes = [vp, sz]
        getattr(L, f"{P}sort_by_index").restype = None
        getattr(L, f"{P}labeller_new").argtypes = [
            sz, sz, sz, sz, ctypes.c_double, ctypes.c_int, ctypes.c_char_p, sz]
        getattr(L, f"{P}labeller_new").restype = vp
        getattr(L, f"{P}labeller_free").argtypes = [vp]
        getattr(L, f"{P}labeller_add").argtypes = [
            vp, ctypes.c_int64, vp, sz, ctypes.c_char_p, sz]
        getattr(L, f"{P}labeller_add").restype = ctypes.c_int
        getattr(L, f"{P}labeller_finish").argtypes = [vp]
        getattr(L, f"{P}labeller_spot_count").argtypes = [vp]
        getattr(L, f"{P}labeller_spot_count").restype = sz
        getattr(L, f"{P}labeller_pixel_count").argtypes = [vp]
        getattr(L, f"{P}labeller_pixel_count").restype = sz
        getattr(L, f"{P}labeller_live").argtypes = [vp]
        getattr(L, f"{P}labeller_live").restype = sz
        f = getattr(L, f"{P}dext_u32_threads")
        f.argtypes = [vp, sz, sz, sz, sz, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        f.restype = ctypes.c_uint64
        if prefix == "oracle_":
            L.oracle_labeller_spots.argtypes = [vp]
            L.oracle_labeller_spots.restype = vp
            L.oracle_labeller_pixels.argtypes = [vp]
            L.oracle_labeller_pixels.restype = vp
            L.oracle_labeller_counts.argtypes = [vp]
            L.oracle_labeller_counts.restype = _Counts
            L.oracle_lz4_decode_block.argtypes = [vp, sz, vp, sz]
            L.oracle_lz4_decode_block.restype = ctypes.c_long
            L.oracle_bitshuffle.argtypes = [vp, vp, sz, sz]
            L.oracle_bitunshuffle.argtypes = [vp, vp, sz, sz]
            L.oracle_bslz4_default_block.argtypes = [sz]
            L.oracle_bslz4_default_block.restype = sz
            L.oracle_bslz4_decode.argtypes = [vp, sz, vp, sz, sz, sz]
            L.oracle_bslz4_decode.restype = ctypes.c_long
            L.oracle_decompress_image.argtypes = [
                vp, sz, ctypes.c_char_p, ctypes.c_uint, sz, sz, vp, sz,
                ctypes.c_char_p, sz]
            L.oracle_decompress_image.restype = ctypes.c_int
            L.oracle_narrow_to_16.argtypes = [vp, sz, vp, ctypes.POINTER(ctypes.c_uint32)]
            L.oracle_narrow_to_16.restype = ctypes.c_int
            L.oracle_scan_u16.argtypes = [vp, sz, ctypes.POINTER(_Scan)]
            L.oracle_scan_u32.argtypes = [vp, sz, ctypes.POINTER(_Scan)]
        else:
            L.ref_labeller_spots.argtypes = [vp, vp]
            L.ref_labeller_pixels.argtypes = [vp, vp]
            L.ref_labeller_counts.argtypes = [vp, ctypes.POINTER(_Counts)]
            L.ref_sizeof_spot.restype = sz
            L.ref_refl_write.argtypes = [ctypes.c_char_p, vp, sz, vp, sz, sz, sz,
                                         ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.c_char_p, sz]
            L.ref_refl_write.restype = ctypes.c_int

    def dext(self, image, capacity=None):
        """(status, SignalPixel array) for a 2-D uint16 or uint32 frame."""
        image = np.ascontiguousarray(image)
        if image.dtype == np.uint16:
            f = getattr(self.lib, f"{self.prefix}dext_u16")
        elif image.dtype == np.uint32:
            f = getattr(self.lib, f"{self.prefix}dext_u32")
        else:
            raise TypeError(f"no threshold for {image.dtype}")
        h, w = image.shape
        if capacity is None:
            capacity = h * w
        out = np.zeros(max(capacity, 1), SIGNAL_PIXEL)
        found = ctypes.c_size_t(0)
        rc = f(_p(image), h, w, _p(out), capacity, ctypes.byref(found))
        return rc, out[: min(found.value, capacity)].copy()

    def sort_by_index(self, pixels):
        pixels = np.ascontiguousarray(pixels, SIGNAL_PIXEL).copy()
        getattr(self.lib, f"{self.prefix}sort_by_index")(_p(pixels), len(pixels))
        return pixels

    def threads_u32(self, frames, threads, total=None):
        """`total` frames (default: each once) taken in turn from `frames`,
        on `threads` workers: (signal pixels, wall seconds)."""
        frames = np.ascontiguousarray(frames, np.uint32)
        n, h, w = frames.shape
        secs = ctypes.c_double(0)
        sig = getattr(self.lib, f"{self.prefix}dext_u32_threads")(
            _p(frames), n, n if total is None else int(total), h, w, threads,
            ctypes.byref(secs))
        return int(sig), secs.value

    # ---- decompression (oracle only: decompress.cc needs the absent
    # bitshuffle and lz4 sources, so it has no compiled counterpart) ----

    def lz4_decode_block(self, block, capacity):
        """(status, bytes): status is the byte count, or < 0 when malformed."""
        src = np.frombuffer(bytes(block), np.uint8)
        out = np.zeros(max(capacity, 1), np.uint8)
        r = self.lib.oracle_lz4_decode_block(_p(src), len(src), _p(out), capacity)
        return r, out[: max(r, 0)].tobytes()

    def bitshuffle(self, data, n, elem):
        src = np.frombuffer(bytes(data), np.uint8)
        out = np.zeros(n * elem, np.uint8)
        self.lib.oracle_bitshuffle(_p(src), _p(out), n, elem)
        return out.tobytes()

    def bitunshuffle(self, data, n, elem):
        src = np.frombuffer(bytes(data), np.uint8)
        out = np.zeros(n * elem, np.uint8)
        self.lib.oracle_bitunshuffle(_p(src), _p(out), n, elem)
        return out.tobytes()

    def bslz4_default_block(self, elem):
        return int(self.lib.oracle_bslz4_default_block(elem))

    def decompress_image(self, data, algorithm, bit_depth, height, width):
        """(0, frame bytes, "") or (-1, None, the reference's message)."""
        src = np.frombuffer(bytes(data), np.uint8)
        expected = height * width * bit_depth // 8
        out = np.zeros(max(expected, 1), np.uint8)
        err = ctypes.create_string_buffer(256)
        rc = self.lib.oracle_decompress_image(
            _p(src) if len(src) else None, len(src), (algorithm or "").encode(),
            bit_depth, height, width, _p(out), expected, err, len(err))
        if rc != 0:
            return rc, None, err.value.decode()
        return 0, out[:expected].tobytes(), ""

    def refl_bytes(self, spots, pixels, width, panel=0, id=0, identifier="",
                   shoeboxes=True):
        """The reference's refl::write of SPOT / PIXEL arrays, as bytes
        (the compiled reference only)."""
        import tempfile
        spots = np.ascontiguousarray(spots, SPOT)
        pixels = np.ascontiguousarray(pixels, PIXEL)
        err = ctypes.create_string_buffer(256)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "t.refl")
            rc = self.lib.ref_refl_write(
                path.encode(), _p(spots) if len(spots) else None, len(spots),
                _p(pixels) if len(pixels) else None, len(pixels), width, panel,
                id, identifier.encode(), int(shoeboxes), err, len(err))
            if rc != 0:
                raise RuntimeError(err.value.decode())
            with open(path, "rb") as fh:
                return fh.read()

    # ---- --gpu-force's narrowing and mxi_max's scan ----

    def narrow_to_16(self, frame):
        """(fits, offending, uint16 frame) for a uint32 frame."""
        f = np.ascontiguousarray(frame, np.uint32)
        out = np.zeros(f.shape, np.uint16)
        worst = ctypes.c_uint32(0)
        ok = self.lib.oracle_narrow_to_16(_p(f), f.size, _p(out), ctypes.byref(worst))
        return bool(ok), worst.value, out

    def scan(self, frame):
        """dict(largest, bad, joins, over) for a uint16 or uint32 frame."""
        f = np.ascontiguousarray(frame)
        s = _Scan()
        fn = {np.dtype(np.uint16): self.lib.oracle_scan_u16,
              np.dtype(np.uint32): self.lib.oracle_scan_u32}[f.dtype]
        fn(_p(f), f.size, ctypes.byref(s))
        return {k: int(getattr(s, k)) for k in ("largest", "bad", "joins", "over")}

    def labeller(self, height, width, min_spot_size=3, max_spot_size=1000,
                 max_separation=2.0, two_d=False):
        return Labeller(self, height, width, min_spot_size, max_spot_size,
                        max_separation, two_d)


class LabellerError(RuntimeError):
    pass


class Labeller:
    def __init__(self, lib, h, w, mn, mx, sep, two_d):
        self.L = lib
        err = ctypes.create_string_buffer(512)
        self.h = getattr(lib.lib, f"{lib.prefix}labeller_new")(
            h, w, mn, mx, sep, int(bool(two_d)), err, 512)
        if not self.h:
            raise LabellerError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.L.lib, f"{self.L.prefix}labeller_free")(self.h)
            self.h = None

    def add(self, frame, signal):
        signal = np.ascontiguousarray(signal, SIGNAL_PIXEL)
        err = ctypes.create_string_buffer(512)
        rc = getattr(self.L.lib, f"{self.L.prefix}labeller_add")(
            self.h, int(frame), _p(signal), len(signal), err, 512)
        if rc != 0:
            raise LabellerError(err.value.decode())

    def finish(self):
        getattr(self.L.lib, f"{self.L.prefix}labeller_finish")(self.h)

    def live(self):
        return getattr(self.L.lib, f"{self.L.prefix}labeller_live")(self.h)

    def spots(self):
        n = getattr(self.L.lib, f"{self.L.prefix}labeller_spot_count")(self.h)
        out = np.zeros(n, SPOT)
        if n:
            if self.L.prefix == "oracle_":
                ptr = self.L.lib.oracle_labeller_spots(self.h)
                ctypes.memmove(_p(out), ptr, n * SPOT.itemsize)
            else:
                self.L.lib.ref_labeller_spots(self.h, _p(out))
        return out

    def pixels(self):
        n = getattr(self.L.lib, f"{self.L.prefix}labeller_pixel_count")(self.h)
        out = np.zeros(n, PIXEL)
        if n:
            if self.L.prefix == "oracle_":
                ptr = self.L.lib.oracle_labeller_pixels(self.h)
                ctypes.memmove(_p(out), ptr, n * PIXEL.itemsize)
            else:
                self.L.lib.ref_labeller_pixels(self.h, _p(out))
        return out

    def counts(self):
        if self.L.prefix == "oracle_":
            c = self.L.lib.oracle_labeller_counts(self.h)
        else:
            c = _Counts()
            self.L.lib.ref_labeller_counts(self.h, ctypes.byref(c))
        return {f: int(getattr(c, f)) for f in COUNTS_FIELDS}


_oracle = None
_ref = None


def oracle():
    """The oracle's own library, built by `make -C oracle`."""
    global _oracle
    if _oracle is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run make -C oracle")
        _oracle = _Lib(path, "oracle_")
    return _oracle


def reference():
    """The reference's CPU code (oracle/_ref), or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libspots_ref.so")
        if not os.path.exists(path):
            return None
        _ref = _Lib(path, "ref_")
    return _ref
