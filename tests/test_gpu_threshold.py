"""The HIP threshold against the reference, through the C ABI, on the device.

Bit-exact on every field of every SignalPixel (index, value, background float
bits, population), as the reference's own device test demands
(tests/spots/test_dext_gpu.cc:73-110), for both kernels: the fused
single-pass kernel and the three direct-window kernels.  Inputs are the golden
size/edge matrix and EIGER2-shaped synthetic frames at full size (4M, 9M,
16M, 512x512 uint16, and a frame whose values exceed the fused kernel's fast
range so that the exact fallback runs); expected outputs come from the
reference's dext.cc compiled as is.
"""
import hashlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KERNELS = ["fused", "direct"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(params=KERNELS)
def kernel(request, device):
    device.set_kernel(request.param)
    yield request.param
    device.set_kernel("fused")


def test_matrix_bit_exact(device, kernel, matrix):
    for name, (frame, expected) in matrix.items():
        say(kernel, name, frame.shape, frame.dtype)
        rc, got = device.find(frame)
        assert rc == 0, name
        assert len(got) == len(expected), (name, len(got), len(expected))
        assert got.tobytes() == expected.tobytes(), name


def test_synthetic_frames_bit_exact(device, kernel, synth_cases):
    for name, case in synth_cases.items():
        kind = case["kind"] if isinstance(case["kind"], str) else tuple(case["kind"])
        frame = device.synth_frame(device.synth_desc(kind, **case["overrides"]),
                                   case["frame"])
        assert hashlib.sha256(frame.tobytes()).hexdigest() == case["frame_sha256"]
        say(kernel, name, frame.shape, frame.dtype)
        rc, got = device.find(frame)
        assert rc == case["status"], name
        assert len(got) == case["count"], (name, len(got), case["count"])
        assert hashlib.sha256(got.tobytes()).hexdigest() == case["signal_sha256"], name


def test_edge_shapes_against_oracle(device, kernel, oracle_lib):
    rng = np.random.default_rng(99)
    shapes = [(1, 1), (1, 700), (700, 1), (2, 2), (11, 233), (233, 11),
              (21, 232), (64, 255), (64, 256), (64, 257), (300, 465), (3, 1000)]
    for (h, w) in shapes:
        for dt in (np.uint16, np.uint32):
            top = np.iinfo(dt).max
            f = rng.poisson(3.0, (h, w)).astype(np.int64)
            s = rng.uniform(size=(h, w)) < 0.03
            f[s] += rng.integers(50, 5000, s.sum())
            f = np.minimum(f, top - 2)
            f[rng.uniform(size=(h, w)) < 0.02] = top - 1
            f = f.astype(dt)
            say(kernel, "edge", h, w, dt)
            rc, got = device.find(f)
            _, want = oracle_lib.dext(f)
            assert rc == 0 and got.tobytes() == want.tobytes(), (h, w, dt)


def test_pinned_frames_against_oracle(device, kernel, oracle_lib):
    """Frames in pinned host memory (sf_host_alloc), which find pulls with a
    kernel over the link: at the allocation's start, 16 bytes in, and 4 bytes
    in (not 16-byte aligned: the runtime's copy), with byte counts that leave
    a tail of under 16 bytes, each the oracle's list."""
    import ctypes
    rng = np.random.default_rng(123)
    room = 700 * 465 * 4 + 64
    p = device.lib().sf_host_alloc(room)
    assert p
    try:
        raw = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint8)),
                                    shape=(room,))
        for (h, w) in [(1, 1), (11, 233), (233, 11), (64, 257), (700, 465)]:
            for dt in (np.uint16, np.uint32):
                for offset in (0, 16, 4):
                    f = rng.poisson(3.0, (h, w)).astype(np.int64)
                    s = rng.uniform(size=(h, w)) < 0.03
                    f[s] += rng.integers(50, 5000, s.sum())
                    f[rng.uniform(size=(h, w)) < 0.01] = np.iinfo(dt).max
                    f = f.astype(dt)
                    view = raw[offset:offset + f.nbytes].view(dt).reshape(h, w)
                    view[...] = f
                    say(kernel, "pinned", h, w, dt, offset)
                    rc, got = device.find(view)
                    _, want = oracle_lib.dext(f)
                    assert rc == 0 and got.tobytes() == want.tobytes(), (h, w, dt, offset)
    finally:
        device.lib().sf_host_free(p)


def test_large_odd_frames_pinned_and_pageable(device, oracle_lib):
    """A 4097 x 4097 uint16 frame (33 570 818 bytes: two 16 MiB staging pieces
    and a last one ending 2 bytes past a 16-byte word) from pageable memory,
    and from pinned memory pulled whole, each the oracle's list."""
    import ctypes
    rng = np.random.default_rng(321)
    h = w = 4097
    f = rng.poisson(1.5, (h, w)).astype(np.int64)
    s = rng.uniform(size=(h, w)) < 0.002
    f[s] += rng.integers(50, 40000, s.sum())
    f[rng.uniform(size=(h, w)) < 0.0005] = 65535
    f = f.astype(np.uint16)
    _, want = oracle_lib.dext(f)
    say("large odd pageable", f.shape, f.dtype, f.nbytes)
    rc, got = device.find(f)
    assert rc == 0 and got.tobytes() == want.tobytes()
    p = device.lib().sf_host_alloc(f.nbytes)
    assert p
    try:
        view = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint16)),
                                     shape=f.shape)
        view[...] = f
        say("large odd pinned", f.shape, f.dtype, f.nbytes)
        rc, got = device.find(view)
        assert rc == 0 and got.tobytes() == want.tobytes()
    finally:
        device.lib().sf_host_free(p)


def test_all_masked_and_empty(device, kernel):
    for dt in (np.uint16, np.uint32):
        top = np.iinfo(dt).max
        f = np.full((64, 64), top - 1, dt)
        f[::7, ::3] = top
        rc, got = device.find(f)
        assert rc == 0 and len(got) == 0
        rc, got = device.find(np.zeros((128, 300), dt))
        assert rc == 0 and len(got) == 0


def test_over_capacity_returns_minus_two(device, kernel):
    # Alternating columns of 0 and 10000: half the pixels are signal, above
    # the list's one entry per four pixels (eight at 16 bits).
    for dt in (np.uint32, np.uint16):
        f = np.zeros((256, 512), dt)
        f[:, ::2] = 10000
        say(kernel, "over capacity", dt)
        rc, got = device.find(f)
        assert rc == -2 and len(got) == 0
    # and the workspace recovers for the next frame
    f = np.zeros((64, 64), np.uint32)
    f[30, 30] = 500
    rc, got = device.find(f)
    assert rc == 0 and len(got) == 1


def test_bad_arguments(device):
    assert device.find(np.zeros((0, 3), np.uint32))[0] == -1


def test_device_workspace_matches_find(device, synth_cases):
    case = synth_cases["eiger4m_f0"]
    frame = device.synth_frame(device.synth_desc("4m", **case["overrides"]),
                               case["frame"])
    d = device.DeviceArray.from_numpy(frame)
    ws = device.Workspace()
    for kernel in KERNELS:
        say("workspace", kernel)
        ws.threshold(d.ptr, 32, frame.shape[0], frame.shape[1], kernel)
        rc, n = ws.sync()
        assert rc == 0 and n == case["count"]
        got = ws.fetch(n)
        assert hashlib.sha256(got.tobytes()).hexdigest() == case["signal_sha256"]
    ws.close()


def test_fused_fast_path_wide_values(device, oracle_lib):
    """Valid pixels anywhere below 2^18 stay on the fused kernel's fast path
    (no exact re-run) and still match dext bit for bit: squares above 2^31,
    values above 2^16 (their high bits travel in a separate ring), isolated
    bright pixels on the first and last rows of 32-row segments, and the
    EIGER-shaped frames, whose hot pixels under spots exceed 2^16."""
    rng = np.random.default_rng(7)
    ws = device.Workspace()
    frames = []
    for (h, w) in [(64, 96), (300, 465), (130, 700), (97, 233)]:
        f = rng.poisson(1.0, (h, w)).astype(np.uint32)
        s = rng.uniform(size=(h, w)) < 0.01
        f[s] = rng.integers(40000, 1 << 18, s.sum())
        for y in range(0, h, 32):
            for x in rng.integers(0, w, 3):
                f[y, x] = rng.choice([46341, 65535, 65536, (1 << 18) - 1])
            if y + 31 < h:
                f[y + 31, rng.integers(0, w)] = 70000
        f[rng.uniform(size=(h, w)) < 0.01] = 0xFFFFFFFF
        frames.append(("wide", f))
    frames.append(("4m synth", device.synth_frame(device.synth_desc("4m"), 0)))
    for name, f in frames:
        assert f[f < 0xFFFFFFFE].max() < (1 << 18)
        say("fast path", name, f.shape)
        d = device.DeviceArray.from_numpy(f)
        before = ws.fallbacks()
        ws.threshold(d.ptr, 32, f.shape[0], f.shape[1], "fused")
        rc, n = ws.sync()
        got = ws.fetch(n)
        _, want = oracle_lib.dext(f)
        assert ws.fallbacks() == before, (name, "re-ran on the exact path")
        assert rc == 0 and got.tobytes() == want.tobytes(), name
    ws.close()


def striped_frame(rng, h, w, dt, phase, width, mask_phase):
    """Poisson background; bright vertical stripes `width` columns wide and
    up to 40 rows long every 23 columns from column `phase`, and a few bright
    bars across the frame; masked columns every 37 from `mask_phase` and
    masked rows every 41.  Stage-0 positives then fall on and off any lattice
    of columns with every phase, beside masked columns of every phase."""
    top = np.iinfo(dt).max
    f = rng.poisson(2.0, (h, w)).astype(np.int64)
    for x0 in range(phase, w, 23):
        cols = slice(x0, min(w, x0 + width))
        for y in rng.integers(0, h, 3):
            rows = slice(y, min(h, y + rng.integers(3, 40)))
            f[rows, cols] += rng.integers(20, 200)
    for y in rng.integers(0, h, max(1, h // 100)):
        f[y, :] += rng.integers(20, 100)
    f = np.minimum(f, top - 2)
    f[:, mask_phase::37] = top - 1
    f[5::41, :] = top
    return f.astype(dt)


@pytest.mark.parametrize("lane_columns", [4, 8])
def test_lane_widths_and_lattice_against_oracle(device, oracle_lib, lane_columns):
    """The fused kernel with four and with eight columns a lane, bit-exact
    against dext on frames built to exercise its lattice pre-test (stage-0
    positives that miss the lattice columns of some rows) and its counted
    path (masked columns at every lattice phase), on shapes that split into
    several strips, and on a 4M EIGER frame."""
    start = device.get_tuning()
    device.set_tuning(lane_columns=lane_columns)
    try:
        rng = np.random.default_rng(1000 + lane_columns)
        frames = []
        for (h, w) in [(300, 1000), (257, 517), (64, 2100)]:
            for dt in (np.uint16, np.uint32):
                for phase, width in [(0, 1), (1, 2), (2, 3), (3, 4), (5, 2)]:
                    frames.append(((h, w, dt.__name__, phase, width),
                                   striped_frame(rng, h, w, dt, phase, width,
                                                 (phase * 7) % 37)))
        frames.append(("4m synth", device.synth_frame(device.synth_desc("4m"), 1)))
        for name, f in frames:
            say("lane columns", lane_columns, name)
            rc, got = device.find(f)
            _, want = oracle_lib.dext(f)
            assert len(want) <= f.nbytes // 16, name  # within the list's capacity
            assert rc == 0 and got.tobytes() == want.tobytes(), name
    finally:
        device.set_tuning(**start)


def gapped_frame(rng, h, w, dt, widths, band=None, from_row=0):
    """Poisson background with masked column blocks (module gaps, dead
    columns) of the given widths every 41 columns, so that blocks start at
    every phase of a lane's columns; bright blobs right beside each block and
    isolated bright pixels anywhere.  `band` masks a run of rows across the
    frame; `from_row` starts the blocks part way down."""
    top = np.iinfo(dt).max
    f = rng.poisson(1.5, (h, w)).astype(np.int64)
    gaps = [(x, widths[k % len(widths)]) for k, x in enumerate(range(17, w - 14, 41))]
    for x, g in gaps:
        for y in rng.integers(0, h - 6, 5):
            for xx in (x - int(rng.integers(1, 6)), x + g + int(rng.integers(0, 5))):
                ys = slice(y, y + int(rng.integers(1, 6)))
                xs = slice(max(0, xx - 1), min(w, xx + int(rng.integers(1, 4))))
                f[ys, xs] += rng.integers(15, 400)
    n = h * w // 300
    f[rng.integers(0, h, n), rng.integers(0, w, n)] += rng.integers(20, 3000, n)
    f = np.minimum(f, top - 2)
    for x, g in gaps:
        f[from_row:, x:x + g] = top
    if band is not None:
        f[band[0]:band[1], :] = top
    return f.astype(dt)


@pytest.mark.parametrize("lane_columns", [4, 8])
def test_static_masks_against_oracle(device, oracle_lib, lane_columns):
    """Masks that stay put down the frame (the fused kernel's pre-test on
    one planned column per four beside them): masked blocks 1 to 13 columns
    wide at every phase, positives right against them, the frame's edges, a
    horizontal gap band (masks that change from row to row around it) and
    blocks that begin part way down; bit-exact against dext."""
    start = device.get_tuning()
    device.set_tuning(lane_columns=lane_columns)
    try:
        rng = np.random.default_rng(77 + lane_columns)
        cases = []
        for (h, w) in [(190, 1100), (150, 533)]:
            for dt in (np.uint16, np.uint32):
                for widths in ([1, 2, 3], [4, 5, 6, 7], [12, 13, 1, 9]):
                    cases.append(((h, w, dt.__name__, tuple(widths)),
                                  gapped_frame(rng, h, w, dt, widths)))
                cases.append(((h, w, dt.__name__, "band"),
                              gapped_frame(rng, h, w, dt, [12, 2], band=(60, 98))))
                cases.append(((h, w, dt.__name__, "late"),
                              gapped_frame(rng, h, w, dt, [3, 12], from_row=47)))
        for name, f in cases:
            say("static masks", lane_columns, name)
            rc, got = device.find(f)
            _, want = oracle_lib.dext(f)
            assert len(want) > 0, name
            assert rc == 0 and got.tobytes() == want.tobytes(), name
    finally:
        device.set_tuning(**start)


def test_stage_times(device):
    device.profile_stages(True)
    try:
        f = device.synth_frame(device.synth_desc("4m"), 0)
        rc, _ = device.find(f)
        t = device.last_stage_times()
        assert rc == 0 and t["upload"] > 0 and t["fused"] > 0
    finally:
        device.profile_stages(False)


def test_cpp_dropin_interface(device):
    """gpu::find<T> through include/dext_gpu.hh, built by the package
    Makefile as tests/cpp/test_dext_gpu, against the oracle."""
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "tests", "cpp", "test_dext_gpu")
    assert os.path.exists(exe), "build first: make -C dials-metal_amd tests"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_find_from_threads_like_mxi_find(device):
    """gpu::find from eight worker threads at once, each on its own stream,
    in mxi_find's chunks of 4N+4 frames grouped in frame order, half of them
    through the C ABI with the direct kernels chosen per call; then a failure
    forced after a frame's work is enqueued throws, leaves nothing in flight,
    and the next frame is right (tests/cpp/test_find_threads.cc)."""
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "tests", "cpp", "test_find_threads")
    assert os.path.exists(exe), "build first: make -C dials-metal_amd tests"
    say("test_find_threads 8")
    r = subprocess.run([exe, "8"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr


def test_find_kernel_per_call_from_python_threads(device, oracle_lib):
    """sf_find_run_as: threads asking for different kernels at once each get
    the oracle's list, and the process-wide choice is left as it was."""
    import threading
    rng = np.random.default_rng(5)
    frames = []
    for i in range(12):
        f = rng.poisson(2.0, (300 + 7 * i, 400)).astype(np.uint32)
        s = rng.uniform(size=f.shape) < 0.02
        f[s] += rng.integers(50, 3000, s.sum()).astype(np.uint32)
        frames.append(f)
    wants = [oracle_lib.dext(f)[1] for f in frames]
    before = device.get_kernel()
    errors = []

    def work(t):
        try:
            for rep in range(3):
                for i in range(t, len(frames), 4):
                    rc, got = device.find(frames[i], kernel="direct" if t % 2 else "fused")
                    if rc != 0 or got.tobytes() != wants[i].tobytes():
                        errors.append((t, rep, i))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append((t, repr(e)))

    ts = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    assert device.get_kernel() == before
