"""Python binding of libspotfinder_hip.so, the gfx950 spot-finding backend.

A thin ctypes layer over the C ABI in include/spotfinder_hip.h: the tests and
the benchmark call the kernels through exactly the surface a C, C++ or
foreign-language caller would.  Loading fails loudly if the library was not
built (``make -C dials-metal_amd`` or ``__graft_entry__.build()``); there is no
CPU fallback.

Reference interfaces mirrored (paths in the reference repository):
  find                 gpu::find<T>            src/spots/dext_gpu.hh:171-180
  available, ...       gpu::available, ...     src/spots/dext_gpu.hh:137-154
  Labeller             dials_spots::Labeller   src/spots/dials_spots.hh:108-158
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPOTFINDER_HIP_LIBRARY names another build of the same library (the
# sanitizer build of `make -C dials-metal_amd check-sanitize`, for one).
LIBRARY = os.environ.get("SPOTFINDER_HIP_LIBRARY") or os.path.join(
    HERE, "libspotfinder_hip.so")

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

COMPONENT = np.dtype(
    [("first", "<u4"), ("n", "<u4"), ("x0", "<u4"), ("x1", "<u4"),
     ("y0", "<u4"), ("y1", "<u4"), ("peak", "<u8"), ("sum", "<u8"),
     ("sum_sq", "<u8", (2,)), ("sum_x", "<u8", (2,)), ("sum_y", "<u8", (2,)),
     ("sum_xx", "<u8", (2,)), ("sum_yy", "<u8", (2,))])
assert COMPONENT.itemsize == 120

KERNELS = {"fused": 0, "direct": 1}


class Grouping(ctypes.Structure):
    _fields_ = [("min_spot_size", ctypes.c_size_t),
                ("max_spot_size", ctypes.c_size_t),
                ("max_separation", ctypes.c_double), ("two_d", ctypes.c_int)]


class Counts(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in COUNTS_FIELDS]


class FrameDesc(ctypes.Structure):
    _fields_ = [("frame", ctypes.c_void_p), ("height", ctypes.c_uint32),
                ("width", ctypes.c_uint32), ("bits", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("seg_rows", ctypes.c_int), ("target_waves", ctypes.c_int),
                ("waves_per_block", ctypes.c_int),
                ("frames_per_launch", ctypes.c_int),
                ("lane_columns", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AutotuneReport(ctypes.Structure):
    _fields_ = [("best", Tuning), ("fused_ms_per_frame", ctypes.c_double),
                ("step_ms_per_frame", ctypes.c_double),
                ("shapes_tried", ctypes.c_uint32),
                ("launch_sizes_tried", ctypes.c_uint32)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "best"}
        d["best"] = self.best.as_dict()
        return d


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("ring", ctypes.c_uint32),
                ("label", ctypes.c_int32), ("reserve_bytes", ctypes.c_uint64)]


class ServerResult(ctypes.Structure):
    _fields_ = [("request", ctypes.c_uint64), ("sweep", ctypes.c_int32),
                ("status", ctypes.c_int32), ("signal_pixels", ctypes.c_uint64),
                ("components", ctypes.c_uint64), ("links", ctypes.c_uint64),
                ("latency_ms", ctypes.c_double),
                ("batch_frames", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class ServerStats(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("frames_per_s", ctypes.c_double), ("mean_batch", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SpotfinderError(RuntimeError):
    """A status other than success from the C ABI."""

    def __init__(self, status, message):
        super().__init__(f"status {status}: {message}")
        self.status = status


class StageTimes(ctypes.Structure):
    _fields_ = [(f, ctypes.c_double)
                for f in ("upload", "stage0", "stage1", "stage2", "fused",
                          "copy", "sort")]


class SynthDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("bits", ctypes.c_uint32),
                ("module_w", ctypes.c_uint32), ("gap_w", ctypes.c_uint32),
                ("module_h", ctypes.c_uint32), ("gap_h", ctypes.c_uint32),
                ("seed", ctypes.c_uint64),
                ("background", ctypes.c_double), ("ring_mean", ctypes.c_double),
                ("ring_radius", ctypes.c_double), ("ring_sigma", ctypes.c_double),
                ("spots_per_frame", ctypes.c_uint32),
                ("sigma_min", ctypes.c_double), ("sigma_max", ctypes.c_double),
                ("peak_min", ctypes.c_double), ("peak_max", ctypes.c_double),
                ("hot_fraction", ctypes.c_double),
                ("persist", ctypes.c_uint32), ("threads", ctypes.c_uint32)]


class NarrowJob(ctypes.Structure):
    _fields_ = [("frame", ctypes.c_void_p), ("out16", ctypes.c_void_p),
                ("pixels", ctypes.c_uint64), ("bits", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class FrameScan(ctypes.Structure):
    _fields_ = [("largest", ctypes.c_uint64), ("bad", ctypes.c_uint64),
                ("joins", ctypes.c_uint64), ("over", ctypes.c_uint64),
                ("fits", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ReflOptions(ctypes.Structure):
    _fields_ = [("panel", ctypes.c_uint64), ("id", ctypes.c_int32),
                ("shoeboxes", ctypes.c_int32), ("identifier", ctypes.c_char_p)]


class Chunk(ctypes.Structure):
    """sf_chunk: a chunk as HDF5 stored it, and its frame in device memory."""
    _fields_ = [("data", ctypes.c_void_p), ("size", ctypes.c_size_t),
                ("algorithm", ctypes.c_char_p), ("bit_depth", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("frame", ctypes.c_void_p)]


_lib = None


def lib():
    """The loaded library.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise ImportError(
            f"{LIBRARY} is not built: run `make -C dials-metal_amd` or "
            "__graft_entry__.build(); there is no CPU fallback")
    L = ctypes.CDLL(LIBRARY)
    vp, sz, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32
    sigs = {
        "sf_backend": ([], ctypes.c_char_p),
        "sf_available": ([], ctypes.c_int),
        "sf_unavailable_reason": ([], ctypes.c_char_p),
        "sf_memory_free": ([], sz),
        "sf_host_alloc": ([sz], vp),
        "sf_host_free": ([vp], None),
        "sf_last_error": ([], ctypes.c_char_p),
        "sf_device_alloc": ([sz], vp),
        "sf_device_free": ([vp], None),
        "sf_copy_to_device": ([vp, vp, sz], ctypes.c_int),
        "sf_copy_to_host": ([vp, vp, sz], ctypes.c_int),
        "sf_copy_on_device": ([vp, vp, sz], ctypes.c_int),
        "sf_device_synchronize": ([], ctypes.c_int),
        "sf_get_tuning": ([ctypes.POINTER(Tuning)], None),
        "sf_set_tuning": ([ctypes.POINTER(Tuning)], ctypes.c_int),
        "sf_set_kernel": ([ctypes.c_int], None),
        "sf_get_kernel": ([], ctypes.c_int),
        "sf_profile_stages": ([ctypes.c_int], None),
        "sf_profiling_stages": ([], ctypes.c_int),
        "sf_last_stage_times": ([], StageTimes),
        "sf_find_run": ([ctypes.c_int, vp, sz, sz, ctypes.POINTER(sz)], ctypes.c_int),
        "sf_find_fetch": ([vp, sz], ctypes.c_int),
        "sf_find_run_as": ([ctypes.c_int, vp, sz, sz, ctypes.c_int, ctypes.c_int,
                            ctypes.POINTER(sz)], ctypes.c_int),
        "sf_inject_find_failure": ([ctypes.c_uint64], None),
        "sf_find_idle": ([], ctypes.c_int),
        "sf_batch_device_bytes": ([vp], sz),
        "sf_threshold_noalloc": ([ctypes.c_int, vp, u32, u32, vp, sz,
                                  ctypes.POINTER(sz), vp], ctypes.c_int),
        "sf_noalloc_scratch_bytes": ([], sz),
        "sf_noalloc_output_records": ([u32, u32, ctypes.c_int], sz),
        "sf_ws_create": ([sz, vp, ctypes.POINTER(vp)], ctypes.c_int),
        "sf_ws_destroy": ([vp], None),
        "sf_ws_stream": ([vp], vp),
        "sf_ws_threshold": ([vp, ctypes.c_int, vp, u32, u32, ctypes.c_int], ctypes.c_int),
        "sf_ws_sync": ([vp, ctypes.POINTER(sz)], ctypes.c_int),
        "sf_ws_fetch": ([vp, vp, sz], ctypes.c_int),
        "sf_ws_device_signal": ([vp], vp),
        "sf_ws_set_timing": ([vp, ctypes.c_int], None),
        "sf_ws_kernel_ms": ([vp, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "sf_ws_fallbacks": ([vp], ctypes.c_uint64),
        "sf_ws_device_mask": ([vp, ctypes.POINTER(u32)], vp),
        "sf_ws_label": ([vp, ctypes.c_int], ctypes.c_int),
        "sf_ws_sync_label": ([vp, ctypes.POINTER(sz), ctypes.POINTER(sz),
                              ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_int)],
                             ctypes.c_int),
        "sf_ws_fetch_labels": ([vp, vp, sz, vp, sz, vp, sz], ctypes.c_int),
        "sf_ws_load_signal": ([vp, vp, sz, u32, u32], ctypes.c_int),
        "sf_batch_create": ([vp, ctypes.POINTER(vp)], ctypes.c_int),
        "sf_batch_destroy": ([vp], None),
        "sf_batch_stream": ([vp], vp),
        "sf_batch_threshold": ([vp, ctypes.POINTER(FrameDesc), sz, ctypes.c_int],
                               ctypes.c_int),
        "sf_batch_sync": ([vp, ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_int)],
                          ctypes.c_int),
        "sf_batch_fetch": ([vp, sz, vp, sz], ctypes.c_int),
        "sf_batch_label": ([vp, ctypes.c_int], ctypes.c_int),
        "sf_batch_label_counts": ([vp, sz, ctypes.POINTER(sz), ctypes.POINTER(sz),
                                   ctypes.POINTER(sz)], ctypes.c_int),
        "sf_batch_fetch_labels": ([vp, sz, vp, sz, vp, sz, vp, sz], ctypes.c_int),
        "sf_batch_set_timing": ([vp, ctypes.c_int], None),
        "sf_batch_kernel_ms": ([vp, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "sf_batch_fallbacks": ([vp], ctypes.c_uint64),
        "sf_labeller_add_batch": ([vp, ctypes.c_int64, vp], ctypes.c_int),
        "sf_grouping_default": ([ctypes.POINTER(Grouping)], None),
        "sf_labeller_create": ([sz, sz, ctypes.POINTER(Grouping),
                                ctypes.POINTER(vp)], ctypes.c_int),
        "sf_labeller_create_device": ([sz, sz, ctypes.POINTER(Grouping),
                                       ctypes.POINTER(vp)], ctypes.c_int),
        "sf_labeller_device_bytes": ([vp], sz),
        "sf_labeller_destroy": ([vp], None),
        "sf_labeller_add": ([vp, ctypes.c_int64, vp, sz], ctypes.c_int),
        "sf_labeller_add_ws": ([vp, ctypes.c_int64, vp], ctypes.c_int),
        "sf_labeller_add_labelled": ([vp, ctypes.c_int64, vp, sz, vp, vp, sz,
                                      vp, sz, ctypes.c_int], ctypes.c_int),
        "sf_labeller_finish": ([vp], ctypes.c_int),
        "sf_labeller_spot_count": ([vp], sz),
        "sf_labeller_spots": ([vp], vp),
        "sf_labeller_pixel_count": ([vp], sz),
        "sf_labeller_pixels": ([vp], vp),
        "sf_labeller_counts": ([vp], Counts),
        "sf_labeller_live": ([vp], sz),
        "sf_autotune": ([vp, sz, ctypes.c_int, ctypes.POINTER(AutotuneReport)],
                        ctypes.c_int),
        "sf_autotune_mix": ([vp, sz, ctypes.POINTER(AutotuneReport)], ctypes.c_int),
        "sf_tuning_for_mix": ([vp, sz, ctypes.POINTER(Tuning)], ctypes.c_int),
        "sf_clear_mix_tunings": ([], None),
        "sf_server_config_default": ([ctypes.POINTER(ServerConfig)], None),
        "sf_server_create": ([ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)],
                             ctypes.c_int),
        "sf_server_destroy": ([vp], None),
        "sf_server_open_sweep": ([vp, u32, u32, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "sf_server_submit": ([vp, ctypes.c_int, vp, vp, sz,
                              ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "sf_server_wait": ([vp, ctypes.c_uint64, ctypes.POINTER(ServerResult)],
                           ctypes.c_int),
        "sf_server_stats_get": ([vp, ctypes.POINTER(ServerStats)], ctypes.c_int),
        "sf_server_stats_reset": ([vp], None),
        "sf_server_replay": ([vp, vp, vp, vp, sz, vp], ctypes.c_int),
        "sf_server_submit_chunk": ([vp, ctypes.c_int, vp, sz, ctypes.c_char_p, vp, sz,
                                    ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "sf_server_replay_chunks": ([vp, vp, vp, vp, sz, vp], ctypes.c_int),
        "sf_synth_default": ([ctypes.POINTER(SynthDesc), u32, u32], None),
        "sf_synth_eiger": ([ctypes.POINTER(SynthDesc), ctypes.c_int, ctypes.c_int], None),
        "sf_synth_frame": ([ctypes.POINTER(SynthDesc), ctypes.c_int64, vp], ctypes.c_int),
        "sf_frame_bytes": ([sz, sz, ctypes.c_uint], sz),
        "sf_refl_options_default": ([ctypes.POINTER(ReflOptions)], None),
        "sf_refl_write": ([ctypes.c_char_p, vp, sz, vp, sz, sz,
                           ctypes.POINTER(ReflOptions)], ctypes.c_int),
        "sf_refl_encode": ([vp, sz, vp, sz, sz, ctypes.POINTER(ReflOptions), vp, sz,
                            ctypes.POINTER(sz)], ctypes.c_int),
        "sf_labeller_write_refl": ([vp, ctypes.c_char_p, ctypes.POINTER(ReflOptions)],
                                   ctypes.c_int),
        "sf_narrow_frames": ([ctypes.POINTER(NarrowJob), sz, ctypes.POINTER(FrameScan),
                              ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "sf_decompress_image": ([vp, sz, ctypes.c_char_p, ctypes.c_uint, u32, u32,
                                 vp, sz], ctypes.c_int),
        "sf_decoder_create": ([vp], vp),
        "sf_decoder_destroy": ([vp], None),
        "sf_decoder_stream": ([vp], vp),
        "sf_decoder_load": ([vp, ctypes.POINTER(Chunk), sz], ctypes.c_int),
        "sf_decoder_run": ([vp], ctypes.c_int),
        "sf_decoder_sync": ([vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "sf_decoder_set_timing": ([vp, ctypes.c_int], None),
        "sf_decoder_kernel_ms": ([vp, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def last_error():
    return lib().sf_last_error().decode()


def _check(status):
    if status != 0:
        raise SpotfinderError(status, last_error())


def backend():
    return lib().sf_backend().decode()


def available():
    return bool(lib().sf_available())


def unavailable_reason():
    return lib().sf_unavailable_reason().decode()


def memory_free():
    return int(lib().sf_memory_free())


def synchronize():
    _check(lib().sf_device_synchronize())


class DeviceArray:
    """A device allocation made through the library's own runtime (one GPU
    runtime per process: a second, such as a framework's bundled one, cannot
    open a device this one already holds)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().sf_device_alloc(self.nbytes)
        if not self.ptr:
            raise SpotfinderError(-4, last_error())

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _check(lib().sf_copy_to_device(d.ptr, _ptr(a), a.nbytes))
        return d

    def copy_from(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the allocation")
        _check(lib().sf_copy_to_device(self.ptr, _ptr(a), a.nbytes))

    def copy_device(self, src, nbytes=None):
        """Copies `nbytes` (all of `src` by default) from another DeviceArray,
        device to device."""
        n = src.nbytes if nbytes is None else int(nbytes)
        if n > self.nbytes or n > src.nbytes:
            raise ValueError("copy larger than an allocation")
        _check(lib().sf_copy_on_device(self.ptr, src.ptr, n))

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype)
        _check(lib().sf_copy_to_host(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().sf_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def get_tuning():
    t = Tuning()
    lib().sf_get_tuning(ctypes.byref(t))
    return t.as_dict()


def set_tuning(**kw):
    """Changes the named launch-shape settings (sf_tuning)."""
    t = Tuning()
    lib().sf_get_tuning(ctypes.byref(t))
    for k, v in kw.items():
        setattr(t, k, int(v))
    if lib().sf_set_tuning(ctypes.byref(t)) != 0:
        raise ValueError(last_error())


def inject_find_failure(nth):
    """The nth find from now (over all threads) fails as a device failure
    would, after its work is enqueued; 0 cancels."""
    lib().sf_inject_find_failure(int(nth))


def find_idle():
    """True when this thread's find stream has no work in flight."""
    return bool(lib().sf_find_idle())


def threshold_noalloc(dev_frame, bits, height, width, dev_out, out_records,
                      stream=None):
    """sf_threshold_noalloc: thresholds a device frame into the device list
    `dev_out` (out_records entries) without allocating; (status, count).
    -1 raises ValueError with the library's message; -3 raises."""
    found = ctypes.c_size_t(0)
    rc = lib().sf_threshold_noalloc(bits, dev_frame, height, width, dev_out,
                                    out_records, ctypes.byref(found), stream)
    if rc == -1:
        raise ValueError(last_error())
    if rc not in (0, -2):
        _check(rc)
    return rc, found.value


def noalloc_output_records(height, width, bits):
    return int(lib().sf_noalloc_output_records(height, width, bits))


def set_kernel(name):
    lib().sf_set_kernel(KERNELS[name])


def get_kernel():
    k = lib().sf_get_kernel()
    return {v: n for n, v in KERNELS.items()}[k]


def profile_stages(on):
    lib().sf_profile_stages(int(bool(on)))


def last_stage_times():
    t = lib().sf_last_stage_times()
    return {f: getattr(t, f) for f, _ in StageTimes._fields_}


def find(frame, kernel=None):
    """gpu::find<T>: (status, SignalPixel array ascending by index) for a
    host uint16 or uint32 frame.  Status as dext(): 0, -1 or -2; a device
    failure raises SpotfinderError.  `kernel` ("fused" or "direct") chooses
    for this call alone; None follows set_kernel."""
    frame = np.ascontiguousarray(frame)
    if frame.dtype == np.uint16:
        bits = 16
    elif frame.dtype == np.uint32:
        bits = 32
    else:
        raise TypeError(f"no threshold for {frame.dtype}")
    if frame.ndim != 2:
        raise ValueError("a frame is two-dimensional")
    h, w = frame.shape
    found = ctypes.c_size_t(0)
    if kernel is None:
        rc = lib().sf_find_run(bits, _ptr(frame) if frame.size else None, h, w,
                               ctypes.byref(found))
    else:
        rc = lib().sf_find_run_as(bits, _ptr(frame) if frame.size else None, h, w,
                                  KERNELS[kernel], 0, ctypes.byref(found))
    if rc == -3 or rc == -4:
        _check(rc)
    if rc != 0:
        return rc, np.zeros(0, SIGNAL_PIXEL)
    out = np.zeros(found.value, SIGNAL_PIXEL)
    _check(lib().sf_find_fetch(_ptr(out), found.value))
    return 0, out


class Workspace:
    """Device-resident thresholding (sf_ws_*): frames already in HBM, results
    kept there.  `stream` is a hipStream_t (int) to borrow, or None."""

    def __init__(self, stream=None):
        self.h = ctypes.c_void_p()
        _check(lib().sf_ws_create(0, stream, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().sf_ws_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().sf_ws_stream(self.h)

    def threshold(self, dev_ptr, bits, height, width, kernel="fused"):
        _check(lib().sf_ws_threshold(self.h, bits, dev_ptr, height, width,
                                     KERNELS[kernel]))

    def sync(self):
        """(status, count): status 0 or -2."""
        found = ctypes.c_size_t(0)
        rc = lib().sf_ws_sync(self.h, ctypes.byref(found))
        if rc not in (0, -2):
            _check(rc)
        return rc, found.value

    def fetch(self, n):
        out = np.zeros(n, SIGNAL_PIXEL)
        _check(lib().sf_ws_fetch(self.h, _ptr(out), n))
        return out

    def set_timing(self, on):
        lib().sf_ws_set_timing(self.h, int(bool(on)))

    def kernel_ms(self):
        """(threshold_ms, order_ms) of the last threshold, timing on."""
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().sf_ws_kernel_ms(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def fallbacks(self):
        return int(lib().sf_ws_fallbacks(self.h))

    def device_signal(self):
        return lib().sf_ws_device_signal(self.h)

    def device_mask(self):
        wpr = ctypes.c_uint32(0)
        p = lib().sf_ws_device_mask(self.h, ctypes.byref(wpr))
        return p, wpr.value


def synth_desc(kind="4m", **overrides):
    """A frame description: "4m" / "16m" EIGER2 geometry, or (w, h) plain."""
    d = SynthDesc()
    if kind == "4m":
        lib().sf_synth_eiger(ctypes.byref(d), 2, 4)
    elif kind == "16m":
        lib().sf_synth_eiger(ctypes.byref(d), 4, 8)
    elif kind == "9m":
        lib().sf_synth_default(ctypes.byref(d), 3110, 3269)
        d.module_w, d.gap_w, d.module_h, d.gap_h = 1030, 10, 514, 37
    else:
        w, h = kind
        lib().sf_synth_default(ctypes.byref(d), w, h)
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def synth_frame(desc, frame=0, out=None):
    dt = np.uint16 if desc.bits == 16 else np.uint32
    if out is None:
        out = np.empty((desc.height, desc.width), dt)
    rc = lib().sf_synth_frame(ctypes.byref(desc), int(frame), _ptr(out))
    if rc != 0:
        raise ValueError("bad synthetic frame description")
    return out


class Batch:
    """Many device frames per launch (sf_batch_*): any mix of sizes and of
    16- and 32-bit frames, optionally replayed from a captured hipGraph."""

    def __init__(self, stream=None):
        self.h = ctypes.c_void_p()
        _check(lib().sf_batch_create(stream, ctypes.byref(self.h)))
        self.n = 0

    def close(self):
        if self.h:
            lib().sf_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def threshold(self, frames, graph=False):
        """frames: (device pointer, height, width, bits) tuples."""
        arr = (FrameDesc * len(frames))()
        for i, (p, h, w, b) in enumerate(frames):
            arr[i].frame, arr[i].height, arr[i].width, arr[i].bits = p, h, w, b
        _check(lib().sf_batch_threshold(self.h, arr, len(frames), int(bool(graph))))
        self.n = len(frames)

    def sync(self):
        """(counts, statuses) per frame."""
        counts = (ctypes.c_size_t * self.n)()
        status = (ctypes.c_int * self.n)()
        _check(lib().sf_batch_sync(self.h, counts, status))
        return list(counts), list(status)

    def fetch(self, i, n):
        out = np.zeros(n, SIGNAL_PIXEL)
        _check(lib().sf_batch_fetch(self.h, i, _ptr(out), n))
        return out

    def label(self, link_previous=True):
        _check(lib().sf_batch_label(self.h, int(bool(link_previous))))

    def labels(self, i):
        n, nc, nl = ctypes.c_size_t(0), ctypes.c_size_t(0), ctypes.c_size_t(0)
        _check(lib().sf_batch_label_counts(self.h, i, ctypes.byref(n),
                                           ctypes.byref(nc), ctypes.byref(nl)))
        comp_of = np.zeros(n.value, np.uint32)
        comps = np.zeros(nc.value, COMPONENT)
        links = np.zeros(2 * nl.value, np.uint32)
        _check(lib().sf_batch_fetch_labels(self.h, i, _ptr(comp_of), n.value,
                                           _ptr(comps), nc.value, _ptr(links),
                                           nl.value))
        return comp_of, comps, links.reshape(-1, 2)

    def set_timing(self, on):
        lib().sf_batch_set_timing(self.h, int(bool(on)))

    def kernel_ms(self):
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().sf_batch_kernel_ms(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def device_bytes(self):
        return int(lib().sf_batch_device_bytes(self.h))

    def fallbacks(self):
        return int(lib().sf_batch_fallbacks(self.h))


class LabellerError(RuntimeError):
    """dials_spots::Labeller's std::runtime_error, through the C ABI."""


class Labeller:
    """dials_spots::Labeller (src/spots/dials_spots.hh:108-158), with the
    labelling and per-spot reductions on the device.  With `device_pixels`
    the frames' pixels also stay on the device: the host stitch sees
    components only, and the spots' pixel lists are written on the device
    (sf_labeller_create_device); it then takes device frames and host lists
    (add), not add_labelled."""

    def __init__(self, height, width, min_spot_size=3, max_spot_size=1000,
                 max_separation=2.0, two_d=False, device_pixels=False):
        g = Grouping()
        lib().sf_grouping_default(ctypes.byref(g))
        g.min_spot_size, g.max_spot_size = min_spot_size, max_spot_size
        g.max_separation, g.two_d = max_separation, int(bool(two_d))
        self.h = ctypes.c_void_p()
        create = (lib().sf_labeller_create_device if device_pixels
                  else lib().sf_labeller_create)
        rc = create(height, width, ctypes.byref(g), ctypes.byref(self.h))
        if rc == -1:
            raise LabellerError(last_error())
        _check(rc)

    def __del__(self):
        try:
            if self.h:
                lib().sf_labeller_destroy(self.h)
                self.h = ctypes.c_void_p()
        except Exception:
            pass

    def _status(self, rc):
        if rc == -1:
            raise LabellerError(last_error())
        _check(rc)

    def add(self, frame, signal):
        signal = np.ascontiguousarray(signal, SIGNAL_PIXEL)
        self._status(lib().sf_labeller_add(self.h, int(frame), _ptr(signal),
                                           len(signal)))

    def add_workspace(self, frame, ws):
        self._status(lib().sf_labeller_add_ws(self.h, int(frame), ws.h))

    def add_batch(self, first_frame, batch):
        self._status(lib().sf_labeller_add_batch(self.h, int(first_frame), batch.h))

    def add_labelled(self, frame, signal, comp_of, comps, links,
                     links_to_previous=True):
        signal = np.ascontiguousarray(signal, SIGNAL_PIXEL)
        comp_of = np.ascontiguousarray(comp_of, np.uint32)
        comps = np.ascontiguousarray(comps, COMPONENT)
        links = np.ascontiguousarray(links, np.uint32).reshape(-1)
        self._status(lib().sf_labeller_add_labelled(
            self.h, int(frame), _ptr(signal), len(signal), _ptr(comp_of),
            _ptr(comps), len(comps), _ptr(links), len(links) // 2,
            int(bool(links_to_previous))))

    def finish(self):
        _check(lib().sf_labeller_finish(self.h))

    def device_bytes(self):
        """Device memory the staged frames still hold (device_pixels)."""
        return int(lib().sf_labeller_device_bytes(self.h))

    def _records(self, count, data, dtype, copy):
        n = count(self.h)
        if not copy:
            # The library's own host array, read-only; valid until the next
            # add, finish or close.
            if not n:
                return np.zeros(0, dtype)
            buf = (ctypes.c_char * (n * dtype.itemsize)).from_address(data(self.h))
            out = np.frombuffer(buf, dtype, n)
            out.flags.writeable = False
            return out
        out = np.zeros(n, dtype)
        if n:
            ctypes.memmove(_ptr(out), data(self.h), n * dtype.itemsize)
        return out

    def spots(self, copy=True):
        return self._records(lib().sf_labeller_spot_count, lib().sf_labeller_spots,
                             SPOT, copy)

    def pixels(self, copy=True):
        return self._records(lib().sf_labeller_pixel_count, lib().sf_labeller_pixels,
                             PIXEL, copy)

    def counts(self):
        c = lib().sf_labeller_counts(self.h)
        return {f: int(getattr(c, f)) for f in COUNTS_FIELDS}

    def live(self):
        return int(lib().sf_labeller_live(self.h))

    def write_refl(self, path, panel=0, id=0, identifier="", shoeboxes=True):
        """The finished grouping's spots as a DIALS reflection table
        (refl::write, src/spots/refl.hh:70-73)."""
        o = _refl_options(panel, id, identifier, shoeboxes)
        if lib().sf_labeller_write_refl(self.h, str(path).encode(), ctypes.byref(o)) != 0:
            raise RuntimeError(last_error())


def autotune(frames, label=True):
    """sf_autotune over device frames given as (ptr, height, width, bits):
    measures the fused kernel's shape and frames per launch, installs the
    best, and returns the report as a dict."""
    arr = (FrameDesc * len(frames))(*[FrameDesc(p, h, w, b) for p, h, w, b in frames])
    rep = AutotuneReport()
    _check(lib().sf_autotune(ctypes.cast(arr, ctypes.c_void_p), len(frames),
                             int(bool(label)), ctypes.byref(rep)))
    return rep.as_dict()


def autotune_mix(frames):
    """sf_autotune_mix over device frames given as (ptr, height, width, bits):
    the fused kernel's shape measured on this batch and kept for batches of
    its size; the process-wide tuning is left alone.  The report as a dict."""
    arr = (FrameDesc * len(frames))(*[FrameDesc(p, h, w, b) for p, h, w, b in frames])
    rep = AutotuneReport()
    _check(lib().sf_autotune_mix(ctypes.cast(arr, ctypes.c_void_p), len(frames),
                                 ctypes.byref(rep)))
    return rep.as_dict()


def tuning_for_mix(frames):
    """The tuning a batch of these frames launches with, and whether it is a
    shape sf_autotune_mix kept for its size."""
    arr = (FrameDesc * len(frames))(*[FrameDesc(p, h, w, b) for p, h, w, b in frames])
    t = Tuning()
    rc = lib().sf_tuning_for_mix(ctypes.cast(arr, ctypes.c_void_p), len(frames),
                                 ctypes.byref(t))
    if rc < 0:
        _check(rc)
    return t.as_dict(), bool(rc)


def clear_mix_tunings():
    lib().sf_clear_mix_tunings()


class NotAdmitted(SpotfinderError):
    """A server refused a sweep: too little free device memory."""


class Server:
    """The serving loop (sf_server_*): device-resident frames from several
    sweeps, batched as they arrive by a worker thread of the library."""

    def __init__(self, max_batch=64, ring=4, label=True, reserve_bytes=1 << 30):
        cfg = ServerConfig(max_batch, ring, int(bool(label)), reserve_bytes)
        self.h = ctypes.c_void_p()
        self._held = {}  # chunk bytes of requests not yet waited for
        _check(lib().sf_server_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().sf_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def open_sweep(self, height, width, bits=32):
        s = ctypes.c_int(-1)
        rc = lib().sf_server_open_sweep(self.h, height, width, bits, ctypes.byref(s))
        if rc == -6:
            raise NotAdmitted(rc, last_error())
        _check(rc)
        return s.value

    def submit(self, sweep, dev_ptr, out=None):
        """Queues a device frame; with `out` (a SIGNAL_PIXEL array) its ordered
        list is copied there before the request completes."""
        r = ctypes.c_uint64(0)
        _check(lib().sf_server_submit(self.h, sweep, dev_ptr,
                                      None if out is None else _ptr(out),
                                      0 if out is None else len(out),
                                      ctypes.byref(r)))
        return r.value

    def submit_chunk(self, sweep, data, algorithm="bslz4", out=None):
        """Queues a compressed chunk (bytes as HDF5 stored them) of a frame of
        the sweep's shape; decoded on the device ahead of the threshold.  The
        bytes are held until the request is waited for."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        r = ctypes.c_uint64(0)
        _check(lib().sf_server_submit_chunk(
            self.h, sweep, ctypes.cast(buf, ctypes.c_void_p), len(data),
            (algorithm or "").encode(), None if out is None else _ptr(out),
            0 if out is None else len(out), ctypes.byref(r)))
        self._held[r.value] = buf
        return r.value

    def wait(self, request):
        res = ServerResult()
        try:
            _check(lib().sf_server_wait(self.h, request, ctypes.byref(res)))
        finally:
            self._held.pop(request, None)
        return res.as_dict()

    def stats(self):
        st = ServerStats()
        _check(lib().sf_server_stats_get(self.h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        lib().sf_server_stats_reset(self.h)

    def replay(self, sweeps, frames, arrivals):
        """Submits frames[i] (device pointers) for sweeps[i] at arrivals[i]
        seconds (ascending), waits for all; the results in order."""
        n = len(sweeps)
        sw = np.ascontiguousarray(sweeps, np.int32)
        fr = np.ascontiguousarray(frames, np.uint64)
        ar = np.ascontiguousarray(arrivals, np.float64)
        out = (ServerResult * n)()
        _check(lib().sf_server_replay(self.h, _ptr(sw), _ptr(fr), _ptr(ar), n,
                                      ctypes.cast(out, ctypes.c_void_p)))
        return [r.as_dict() for r in out]

    def replay_chunks(self, sweeps, chunks, arrivals, algorithm="bslz4"):
        """replay() for compressed chunks: chunks[i] (bytes) for sweeps[i],
        submitted from host memory at arrivals[i] seconds."""
        n = len(sweeps)
        sw = np.ascontiguousarray(sweeps, np.int32)
        ar = np.ascontiguousarray(arrivals, np.float64)
        held = {}
        desc = (Chunk * max(1, n))()
        alg = (algorithm or "").encode()
        for i, c in enumerate(chunks):
            if id(c) not in held:
                held[id(c)] = ctypes.create_string_buffer(bytes(c), len(c))
            desc[i].data = ctypes.cast(held[id(c)], ctypes.c_void_p)
            desc[i].size = len(c)
            desc[i].algorithm = alg
        out = (ServerResult * max(1, n))()
        _check(lib().sf_server_replay_chunks(self.h, _ptr(sw), ctypes.cast(desc, ctypes.c_void_p),
                                             _ptr(ar), n, ctypes.cast(out, ctypes.c_void_p)))
        return [out[i].as_dict() for i in range(n)]


def workspace_label_enqueue(ws, link_previous=True):
    """Enqueues the labelling of the workspace's last frame; returns at once."""
    _check(lib().sf_ws_label(ws.h, int(bool(link_previous))))


def workspace_label(ws, link_previous=True):
    """Labels the workspace's last frame; returns (comp_of, comps, links,
    linked) copied to the host."""
    _check(lib().sf_ws_label(ws.h, int(bool(link_previous))))
    n, nc, nl = ctypes.c_size_t(0), ctypes.c_size_t(0), ctypes.c_size_t(0)
    linked = ctypes.c_int(0)
    _check(lib().sf_ws_sync_label(ws.h, ctypes.byref(n), ctypes.byref(nc),
                                  ctypes.byref(nl), ctypes.byref(linked)))
    comp_of = np.zeros(n.value, np.uint32)
    comps = np.zeros(nc.value, COMPONENT)
    links = np.zeros(2 * nl.value, np.uint32)
    _check(lib().sf_ws_fetch_labels(ws.h, _ptr(comp_of), n.value, _ptr(comps),
                                    nc.value, _ptr(links), nl.value))
    return comp_of, comps, links.reshape(-1, 2), bool(linked.value)


def workspace_load_signal(ws, signal, height, width):
    signal = np.ascontiguousarray(signal, SIGNAL_PIXEL)
    rc = lib().sf_ws_load_signal(ws.h, _ptr(signal), len(signal), height, width)
    if rc == -1:
        raise ValueError(last_error())
    _check(rc)


# ---- decompression on the device (src/spots/decompress.hh:22-30) ----

class DecompressError(RuntimeError):
    """A chunk that does not describe its frame, or malformed data: the
    reference's decompress::image throws std::runtime_error for these."""


def frame_bytes(height, width, bit_depth):
    """decompress::frame_bytes (decompress.hh:18-20)."""
    return int(lib().sf_frame_bytes(height, width, bit_depth))


def decompress_image(data, algorithm, bit_depth, height, width, frame):
    """decompress::image with the frame left in device memory: ``frame`` is a
    DeviceArray (or a device address and its byte count as a tuple)."""
    ptr, nbytes = (frame.ptr, frame.nbytes) if hasattr(frame, "ptr") else frame
    data = bytes(data)
    rc = lib().sf_decompress_image(data, len(data), (algorithm or "").encode(),
                                   bit_depth, height, width, ptr, nbytes)
    if rc == -1:
        raise DecompressError(last_error())
    _check(rc)


class Decoder:
    """sf_decoder: many chunks decoded at once.  load() checks every chunk and
    uploads them; run() decodes them into their frames (again and again, the
    chunks staying resident); sync() waits and returns each chunk's status."""

    def __init__(self, stream=None):
        self.h = lib().sf_decoder_create(stream)
        if not self.h:
            raise RuntimeError(last_error())
        self._keep = []

    def close(self):
        if self.h:
            lib().sf_decoder_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        return lib().sf_decoder_stream(self.h)

    def load(self, chunks):
        """chunks: (data bytes, algorithm, bit_depth, height, width, frame)
        tuples, frame a DeviceArray or a device address."""
        arr = (Chunk * len(chunks))()
        keep = []
        for i, (data, algo, bd, h, w, frame) in enumerate(chunks):
            data = data if isinstance(data, bytes) else bytes(data)
            keep.append(data)  # read in place: the library copies it in load()
            arr[i].data = ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p)
            arr[i].size = len(data)
            arr[i].algorithm = (algo or "").encode()
            arr[i].bit_depth, arr[i].height, arr[i].width = bd, h, w
            arr[i].frame = frame.ptr if hasattr(frame, "ptr") else frame
        rc = lib().sf_decoder_load(self.h, arr, len(chunks))
        if rc == -1:
            raise DecompressError(last_error())
        _check(rc)
        self.n = len(chunks)

    def load_descriptors(self, arr, n, keep=None):
        """Loads n sf_chunk descriptors as they stand (NxmxSeries.read_chunks'
        output); `keep` holds whatever their data point into."""
        rc = lib().sf_decoder_load(self.h, arr, n)
        if rc == -1:
            raise DecompressError(last_error())
        _check(rc)
        self._keep = keep
        self.n = n

    def run(self):
        _check(lib().sf_decoder_run(self.h))

    def sync(self):
        """Each chunk's status: 0, or -1 where its data were malformed;
        last_error() holds the first failure's message."""
        st = (ctypes.c_int * max(self.n, 1))()
        rc = lib().sf_decoder_sync(self.h, st)
        if rc not in (0, -1):
            _check(rc)
        return list(st)[: self.n]

    def set_timing(self, on):
        lib().sf_decoder_set_timing(self.h, int(bool(on)))

    def kernel_ms(self):
        ms = ctypes.c_double(0)
        _check(lib().sf_decoder_kernel_ms(self.h, ctypes.byref(ms)))
        return ms.value


# ---- the NXmx chunk reader (libspotfinder_nxmx.so) ----

NXMX_LIBRARY = os.path.join(HERE, "libspotfinder_nxmx.so")
_nxmx = None


class NxmxInfo(ctypes.Structure):
    _fields_ = [("images", ctypes.c_uint64), ("height", ctypes.c_uint64),
                ("width", ctypes.c_uint64)]


class NxmxBlock(ctypes.Structure):
    _fields_ = [("filename", ctypes.c_char_p), ("dataset", ctypes.c_char_p),
                ("first", ctypes.c_uint64), ("last", ctypes.c_uint64),
                ("offset", ctypes.c_uint64)]


class NxmxFrame(ctypes.Structure):
    _fields_ = [("number", ctypes.c_uint64), ("height", ctypes.c_uint64),
                ("width", ctypes.c_uint64), ("bit_depth", ctypes.c_uint32),
                ("algorithm", ctypes.c_char_p), ("size", ctypes.c_size_t)]

    def as_dict(self):
        return {"number": self.number, "height": self.height, "width": self.width,
                "bit_depth": self.bit_depth,
                "algorithm": (self.algorithm or b"").decode(), "size": self.size}


class NxmxError(RuntimeError):
    """What series::nxmx throws (std::runtime_error), through the C ABI."""


def nxmx_lib():
    """libspotfinder_nxmx.so, the reader over HDF5 (built where HDF5 is)."""
    global _nxmx
    if _nxmx is not None:
        return _nxmx
    if not os.path.exists(NXMX_LIBRARY):
        raise ImportError(f"{NXMX_LIBRARY} is not built (needs the HDF5 headers)")
    L = ctypes.CDLL(NXMX_LIBRARY)
    vp, sz, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64
    for name, args, res in (
            ("sf_nxmx_open", [ctypes.c_char_p, ctypes.POINTER(vp),
                              ctypes.POINTER(NxmxInfo)], ctypes.c_int),
            ("sf_nxmx_close", [vp], None),
            ("sf_nxmx_block_count", [vp], sz),
            ("sf_nxmx_get_block", [vp, sz, ctypes.POINTER(NxmxBlock)], ctypes.c_int),
            ("sf_nxmx_frame_info", [vp, u64, ctypes.POINTER(NxmxFrame)], ctypes.c_int),
            ("sf_nxmx_read", [vp, u64, vp, sz, ctypes.POINTER(NxmxFrame)], ctypes.c_int),
            ("sf_nxmx_read_chunks", [vp, u64, sz, vp, sz, vp, vp,
                                     ctypes.POINTER(sz), ctypes.POINTER(sz), vp],
             ctypes.c_int),
            ("sf_nxmx_last_error", [], ctypes.c_char_p)):
        f = getattr(L, name)
        f.argtypes, f.restype = args, res
    _nxmx = L
    return L


class NxmxSeries:
    """series::nxmx (src/spots/series.hh:32-97): an NXmx master file's frames
    as raw chunks.  Raises FileNotFoundError while the master is not a
    regular file, NxmxError for anything the file or HDF5 refuses."""

    def __init__(self, master):
        L = nxmx_lib()
        self.h = ctypes.c_void_p()
        info = NxmxInfo()
        rc = L.sf_nxmx_open(str(master).encode(), ctypes.byref(self.h), ctypes.byref(info))
        if rc == 1:
            raise FileNotFoundError(str(master))
        self._status(rc)
        self.images, self.height, self.width = info.images, info.height, info.width

    def _status(self, rc):
        if rc == -1:
            raise ValueError(nxmx_lib().sf_nxmx_last_error().decode())
        if rc == -3:
            raise NxmxError(nxmx_lib().sf_nxmx_last_error().decode())

    def close(self):
        if self.h:
            nxmx_lib().sf_nxmx_close(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def blocks(self):
        L = nxmx_lib()
        out = []
        for i in range(L.sf_nxmx_block_count(self.h)):
            b = NxmxBlock()
            self._status(L.sf_nxmx_get_block(self.h, i, ctypes.byref(b)))
            out.append({"filename": b.filename.decode(), "dataset": b.dataset.decode(),
                        "first": b.first, "last": b.last, "offset": b.offset})
        return out

    def frame_info(self, frame):
        f = NxmxFrame()
        rc = nxmx_lib().sf_nxmx_frame_info(self.h, frame, ctypes.byref(f))
        self._status(rc)
        return None if rc == 1 else f.as_dict()

    def read(self, frame):
        """Reader::read: (chunk bytes, description), or None for a frame
        never written."""
        f = NxmxFrame()
        L = nxmx_lib()
        rc = L.sf_nxmx_frame_info(self.h, frame, ctypes.byref(f))
        self._status(rc)
        if rc == 1:
            return None
        buf = ctypes.create_string_buffer(max(1, f.size))
        rc = L.sf_nxmx_read(self.h, frame, buf, f.size, ctypes.byref(f))
        self._status(rc)
        return buf.raw[: f.size], f.as_dict()

    def read_chunks(self, first, frames):
        """Frames first .. first + len(frames) - 1 (frames: their device
        frames, DeviceArray or address) read into one host buffer: (Chunk
        array, count, buffer, present flags) for Decoder.load_descriptors."""
        L = nxmx_lib()
        n = len(frames)
        ptrs = (ctypes.c_void_p * max(n, 1))(
            *[f.ptr if hasattr(f, "ptr") else f for f in frames])
        chunks = (Chunk * max(n, 1))()
        present = (ctypes.c_int * max(n, 1))()
        loaded, used = ctypes.c_size_t(0), ctypes.c_size_t(0)
        rc = L.sf_nxmx_read_chunks(self.h, first, n, None, 0, ptrs, chunks,
                                   ctypes.byref(loaded), ctypes.byref(used), present)
        if rc != -2:
            self._status(rc)
        buf = (ctypes.c_uint8 * max(1, used.value))()
        rc = L.sf_nxmx_read_chunks(self.h, first, n, buf, used.value, ptrs, chunks,
                                   ctypes.byref(loaded), ctypes.byref(used), present)
        self._status(rc)
        return chunks, loaded.value, buf, list(present)[:n]


# ---- --gpu-force and mxi_max on frames in HBM ----

def narrow_frames(jobs, timed=False):
    """narrow_to_16 (find_spots.cc:162-182) and scan<T> (apps/mxi_max.cc:45-68)
    over device frames in one launch.  jobs: (frame, out16 or None, pixels,
    bits) with frame/out16 DeviceArrays or device addresses.  Returns a list of
    dicts (largest, bad, joins, over, fits), and the device time in ms when
    timed."""
    arr = (NarrowJob * len(jobs))()
    for i, (frame, out, pixels, bits) in enumerate(jobs):
        arr[i].frame = frame.ptr if hasattr(frame, "ptr") else frame
        arr[i].out16 = (out.ptr if hasattr(out, "ptr") else out) if out is not None else None
        arr[i].pixels, arr[i].bits = pixels, bits
    res = (FrameScan * max(len(jobs), 1))()
    ms = ctypes.c_double(0)
    _check(lib().sf_narrow_frames(arr, len(jobs), res, ctypes.byref(ms)))
    out = [{k: int(getattr(r, k)) for k in ("largest", "bad", "joins", "over")}
           | {"fits": bool(r.fits)} for r in list(res)[: len(jobs)]]
    return (out, ms.value) if timed else out


# ---- the spots as a DIALS reflection table (src/spots/refl.cc:223-351) ----

def _refl_options(panel, id, identifier, shoeboxes):
    o = ReflOptions()
    lib().sf_refl_options_default(ctypes.byref(o))
    o.panel, o.id, o.shoeboxes = panel, id, int(bool(shoeboxes))
    o.identifier = identifier.encode() if identifier else None
    return o


def refl_encode(spots, pixels, width, panel=0, id=0, identifier="", shoeboxes=True):
    """The .refl bytes of SPOT / PIXEL arrays (a grouping's output)."""
    spots = np.ascontiguousarray(spots, SPOT)
    pixels = np.ascontiguousarray(pixels, PIXEL)
    o = _refl_options(panel, id, identifier, shoeboxes)
    size = ctypes.c_size_t(0)
    args = (_ptr(spots) if len(spots) else None, len(spots),
            _ptr(pixels) if len(pixels) else None, len(pixels), width, ctypes.byref(o))
    # At most: the headers, 112 bytes a row of plain columns, and each shoebox
    # (29 bytes and 9 a voxel); one call writes it all.
    cap = 4096 + len(identifier or "") + 112 * len(spots)
    if shoeboxes and len(spots):
        b = spots["bbox"].astype(np.int64)
        cap += int((29 + 9 * (b[:, 1] - b[:, 0]) * (b[:, 3] - b[:, 2])
                    * (b[:, 5] - b[:, 4])).sum())
    out = np.empty(cap, np.uint8)
    rc = lib().sf_refl_encode(*args, _ptr(out), cap, ctypes.byref(size))
    if rc == -2:  # not reached for a consistent table; sized exactly, once more
        out = np.empty(size.value, np.uint8)
        rc = lib().sf_refl_encode(*args, _ptr(out), size.value, ctypes.byref(size))
    if rc != 0:
        raise RuntimeError(last_error())
    return out[: size.value].tobytes()


def refl_write(path, spots, pixels, width, panel=0, id=0, identifier="", shoeboxes=True):
    spots = np.ascontiguousarray(spots, SPOT)
    pixels = np.ascontiguousarray(pixels, PIXEL)
    o = _refl_options(panel, id, identifier, shoeboxes)
    if lib().sf_refl_write(str(path).encode(), _ptr(spots) if len(spots) else None,
                           len(spots), _ptr(pixels) if len(pixels) else None,
                           len(pixels), width, ctypes.byref(o)) != 0:
        raise RuntimeError(last_error())
