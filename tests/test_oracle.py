"""The oracle's threshold, pinned.

  * against the reference's own outputs (golden vectors written by
    tests/golden/make_golden.py from the reference's dext.cc compiled as is);
  * against an independent brute force of the reference's semantics
    (src/spots/dext.cc:53-222: clamped windows summed straight, the tests
    evaluated per pixel in numpy's IEEE float32/float64) on tiny frames,
    including the uint16 wrap of the sum of squares;
  * against closed forms: an all-masked frame, a flat frame, a lone bright
    pixel, and a frame whose every background pixel is eroded away (m < 2).
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O


def test_oracle_matches_reference_matrix(oracle_lib, matrix):
    total = 0
    for name, (frame, expected) in matrix.items():
        rc, got = oracle_lib.dext(frame)
        assert rc == 0, name
        assert len(got) == len(expected), name
        assert got.tobytes() == expected.tobytes(), name
        total += len(got)
    assert total > 50000  # the matrix is not trivially empty


def test_oracle_matches_reference_synth(oracle_lib, synth_cases, sf):
    for name, case in synth_cases.items():
        if case["kind"] == "16m":
            continue  # covered on the device; a few seconds on one core
        kind = case["kind"] if isinstance(case["kind"], str) else tuple(case["kind"])
        frame = sf.synth_frame(sf.synth_desc(kind, **case["overrides"]), case["frame"])
        assert hashlib.sha256(frame.tobytes()).hexdigest() == case["frame_sha256"], \
            f"{name}: the synthetic generator no longer makes the golden frame"
        rc, got = oracle_lib.dext(frame)
        assert rc == case["status"]
        assert len(got) == case["count"], name
        assert hashlib.sha256(got.tobytes()).hexdigest() == case["signal_sha256"], name


def test_oracle_matches_reference_build_random():
    ref = O.reference()
    if ref is None:
        pytest.skip("oracle/_ref is not built (no reference sources here)")
    o = O.oracle()
    rng = np.random.default_rng(1234)
    for trial in range(30):
        dt = (np.uint16, np.uint32)[trial % 2]
        h, w = rng.integers(1, 90, 2)
        top = np.iinfo(dt).max
        f = rng.poisson(rng.uniform(0.1, 30), (h, w)).astype(np.int64)
        bright = rng.uniform(size=(h, w)) < 0.05
        f[bright] += rng.integers(10, 60000, bright.sum())
        f = np.minimum(f, top - 2)
        f[rng.uniform(size=(h, w)) < 0.03] = top - 1
        f[rng.uniform(size=(h, w)) < 0.01] = top
        f = f.astype(dt)
        ra, a = o.dext(f)
        rb, b = ref.dext(f)
        assert ra == rb and a.tobytes() == b.tobytes(), (trial, h, w)


def brute_dext(f):
    """dext<T> by definition, one pixel at a time (tiny frames only)."""
    dt = f.dtype
    top = np.iinfo(dt).max
    masked = top - 1
    R = np.float32 if dt == np.uint16 else np.float64
    mod = 1 << (32 if dt == np.uint16 else 64)
    H, W = f.shape
    v = f.astype(np.int64)
    valid = f < masked
    scr0 = np.zeros((H, W), bool)
    for i in range(H):
        for j in range(W):
            r0, r1, c0, c1 = max(0, i - 3), min(H, i + 4), max(0, j - 3), min(W, j + 4)
            win, ok = v[r0:r1, c0:c1], valid[r0:r1, c0:c1]
            m = int(ok.sum())
            s = int(win[ok].sum())
            s2 = int(sum(int(x) * int(x) for x in win[ok])) % mod
            sig = False
            if m >= 2:
                mm, ss, ss2 = R(m), R(s), R(s2)
                lhs = (mm * ss2 - ss * ss) - ss * (mm - R(1))
                rhs = (ss * R(6)) * np.sqrt(R(2) * (mm - R(1)))
                sig = bool(lhs > rhs)
            scr0[i, j] = sig or not valid[i, j]
    scr1 = np.zeros((H, W), bool)
    for i in range(H):
        for j in range(W):
            if not valid[i, j]:
                scr1[i, j] = True
                continue
            r0, r1, c0, c1 = max(0, i - 2), min(H, i + 3), max(0, j - 2), min(W, j + 3)
            scr1[i, j] = bool(scr0[r0:r1, c0:c1].all())
    out = []
    for i in range(H):
        for j in range(W):
            p = int(v[i, j]) if valid[i, j] else 0
            if not (p > 0 and scr1[i, j]):
                continue
            r0, r1, c0, c1 = max(0, i - 5), min(H, i + 6), max(0, j - 5), min(W, j + 6)
            ok = valid[r0:r1, c0:c1] & ~scr1[r0:r1, c0:c1]
            m = int(ok.sum())
            s = int(v[r0:r1, c0:c1][ok].sum())
            mean = R(s) / R(m) if m >= 2 else R(0)
            if R(p) >= mean + R(3) * np.sqrt(mean):
                out.append((i * W + j, p, np.float32(mean), m, 0))
    return np.array(out, dtype=O.SIGNAL_PIXEL)


@pytest.mark.parametrize("dt", [np.uint16, np.uint32])
def test_oracle_matches_brute_force(oracle_lib, dt):
    rng = np.random.default_rng(7 if dt == np.uint16 else 8)
    top = np.iinfo(dt).max
    checked = 0
    for trial in range(24):
        h, w = (int(x) for x in rng.integers(1, 15, 2))
        f = rng.poisson(rng.uniform(0.2, 6), (h, w)).astype(np.int64)
        spikes = rng.uniform(size=(h, w)) < 0.2
        f[spikes] += rng.integers(5, 4000, spikes.sum())
        if trial % 3 == 0:
            # near the top of the range, where uint16's sum of squares wraps
            hot = rng.uniform(size=(h, w)) < 0.5
            f[hot] = rng.integers(60000, 65534, hot.sum())
        f = np.minimum(f, top - 2)
        f[rng.uniform(size=(h, w)) < 0.06] = top - 1
        f[rng.uniform(size=(h, w)) < 0.03] = top
        f = f.astype(dt)
        rc, got = oracle_lib.dext(f)
        assert rc == 0
        want = brute_dext(f)
        assert got.tobytes() == want.tobytes(), (trial, f)
        checked += len(got)
    assert checked > 0


@pytest.mark.parametrize("dt", [np.uint16, np.uint32])
def test_closed_forms(oracle_lib, dt):
    top = np.iinfo(dt).max
    # All masked (both sentinels): nothing.
    f = np.full((64, 64), top - 1, dt)
    f[::3, ::5] = top
    assert oracle_lib.dext(f)[1].size == 0
    # A flat frame has no dispersion: m*s2 - s^2 = 0 < the right-hand side.
    assert oracle_lib.dext(np.full((40, 50), 7, dt))[1].size == 0
    # A lone pixel of value v >= 3 in a field of zeros, away from the edges:
    # stage 0 sets its 7x7 (48 v (v - 1) > 6 v sqrt(96)), the erode leaves the
    # 3x3 centre, stage 2 averages 121 - 9 = 112 zeros: mean 0, signal.
    f = np.zeros((40, 50), dt)
    f[20, 30] = 500
    rc, got = oracle_lib.dext(f)
    assert rc == 0 and len(got) == 1
    assert (got[0]["index"], got[0]["value"], got[0]["background"],
            got[0]["population"]) == (20 * 50 + 30, 500, 0.0, 112)
    # v = 2 fails stage 0 (48 * 2 * 1 = 96 is not above 12 sqrt(96) = 117.6).
    f[20, 30] = 2
    assert oracle_lib.dext(f)[1].size == 0
    # Every pixel eroded in: a 3x3 checkerboard of 100 and 0 is dispersed
    # everywhere (9*50000 - 500^2 - 500*8 = 196000 > 3000*sqrt(16)), so scr1 is
    # all set, stage 2 has no background (m = 0, mean 0) and every non-zero
    # pixel is signal with population 0.
    f = np.zeros((3, 3), dt)
    f[::2, ::2] = 100
    f[1, 1] = 100
    rc, got = oracle_lib.dext(f)
    assert [int(x) for x in got["index"]] == [0, 2, 4, 6, 8]
    assert (got["background"] == 0).all() and (got["population"] == 0).all()


def test_bad_arguments(oracle_lib):
    import ctypes
    f = np.zeros((4, 4), np.uint32)
    found = ctypes.c_size_t(0)
    out = np.zeros(4, O.SIGNAL_PIXEL)
    L = oracle_lib.lib
    assert L.oracle_dext_u32(None, 4, 4, O._p(out), 4, ctypes.byref(found)) == -1
    assert L.oracle_dext_u32(O._p(f), 0, 4, O._p(out), 4, ctypes.byref(found)) == -1
    assert L.oracle_dext_u32(O._p(f), 1 << 20, 1 << 12, O._p(out), 4,
                             ctypes.byref(found)) == -1


def test_capacity_overflow_reports_minus_two(oracle_lib):
    # Alternating 0/10000 columns: dispersed everywhere, eroded everywhere,
    # so every non-zero pixel is signal: half the frame, above a quarter.
    f = np.zeros((64, 64), np.uint32)
    f[:, ::2] = 10000
    rc, got = oracle_lib.dext(f, capacity=64 * 64 // 4)
    assert rc == -2
    rc, got = oracle_lib.dext(f)
    assert rc == 0 and len(got) == 64 * 32


def test_sort_by_index_matches_reference():
    o = O.oracle()
    rng = np.random.default_rng(3)
    for n in (0, 1, 5, 1023, 1024, 5000):
        p = np.zeros(n, O.SIGNAL_PIXEL)
        p["index"] = rng.permutation(max(n * 3, 1))[:n]
        p["value"] = rng.integers(0, 1000, n)
        got = o.sort_by_index(p)
        assert (np.diff(got["index"].astype(np.int64)) > 0).all()
        assert sorted(got.tobytes()[i:i + 16] for i in range(0, 16 * n, 16)) == \
            sorted(p.tobytes()[i:i + 16] for i in range(0, 16 * n, 16))
        ref = O.reference()
        if ref is not None:
            assert ref.sort_by_index(p).tobytes() == got.tobytes()
