"""--gpu-force's narrowing and mxi_max's scan on the device, through the C
ABI, against the oracle (tests/test_narrow.py pins the oracle).

Every field of every frame's scan equals the oracle's, and every narrowed
frame equals the oracle's narrowed frame byte for byte: the reference test's
planted series in one launch, random frames with every marker and the values
at the 16-bit limit at sizes that are not multiples of the eight pixels a
thread takes, and a full EIGER2 16M frame, whose narrowed frame's 16-bit
threshold equals the 32-bit threshold of the original, byte for byte.
"""
import sys

import numpy as np
import pytest

from test_narrow import planted

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(sf, frames, narrow=True):
    devs, outs, jobs = [], [], []
    for f in frames:
        d = sf.DeviceArray.from_numpy(np.ascontiguousarray(f))
        o = sf.DeviceArray(max(2 * f.size, 16)) if (narrow and f.dtype == np.uint32) else None
        devs.append(d)
        outs.append(o)
        jobs.append((d, o, f.size, 8 * f.dtype.itemsize))
    res = sf.narrow_frames(jobs)
    narrowed = [o.to_numpy(np.uint16, (f.size,)).reshape(f.shape) if o else None
                for o, f in zip(outs, frames)]
    for x in devs + [o for o in outs if o]:
        x.free()
    return res, narrowed


def check(orc, frames, res, narrowed):
    for f, r, n16 in zip(frames, res, narrowed):
        want = orc.scan(f)
        assert {k: r[k] for k in want} == want
        if f.dtype == np.uint32:
            fits, worst, f16 = orc.narrow_to_16(f)
            assert r["fits"] == fits and r["largest"] == worst
            if n16 is not None:
                assert n16.tobytes() == f16.tobytes()


def test_planted_series_in_one_launch(device, oracle_lib):
    frames = planted(16) + planted(32) + planted(32, hot=(7, 70000))
    say("planted", len(frames), "frames")
    res, narrowed = run(device, frames)
    check(oracle_lib, frames, res, narrowed)
    s32 = res[12:24]
    assert sum(r["bad"] for r in s32) == 12 and sum(r["joins"] for r in s32) == 3072
    hot = res[24:]
    assert [z for z, r in enumerate(hot) if not r["fits"]] == [7]
    assert hot[7]["largest"] == 70000


def test_random_frames_with_markers_and_ragged_sizes(device, oracle_lib):
    rng = np.random.default_rng(9)
    frames = []
    for trial in range(24):
        bits = 16 if trial % 3 == 0 else 32
        dt = np.uint16 if bits == 16 else np.uint32
        top = np.iinfo(dt).max
        n = int(rng.choice([1, 7, 8, 9, 2047, 2049, 100003, 1 << 20]))
        f = rng.integers(0, 1 << (12 if trial % 2 else bits), size=n,
                         dtype=np.uint64).astype(dt)
        edges = np.array([0xFFFC, 0xFFFD, 0xFFFE, 0xFFFF, top - 2, top - 1, top],
                         np.uint64).astype(dt)
        k = int(rng.integers(0, min(n, 9) + 1))
        f[rng.integers(0, n, k)] = rng.choice(edges, k)
        frames.append(f)
    say("random", [f.size for f in frames])
    res, narrowed = run(device, frames)
    check(oracle_lib, frames, res, narrowed)


def test_16m_frame_narrowed_thresholds_as_the_original(device, oracle_lib):
    sf = device
    # spots up to 3000 counts and no hot pixels, so that the frame fits
    desc = sf.synth_desc("16m", spots_per_frame=2000, persist=1, seed=2024,
                         peak_max=3000.0, hot_fraction=0.0)
    h, w = desc.height, desc.width
    f = sf.synth_frame(desc, 1)
    d = sf.DeviceArray.from_numpy(f)
    o = sf.DeviceArray(2 * f.size)
    (r,) = sf.narrow_frames([(d, o, f.size, 32)])
    fits, worst, f16 = oracle_lib.narrow_to_16(f)
    assert r["fits"] == fits and r["largest"] == worst and fits
    assert o.to_numpy(np.uint16, (h, w)).tobytes() == f16.tobytes()
    ws = sf.Workspace()
    lists = []
    for src, bits in ((d, 32), (o, 16)):
        ws.threshold(src.ptr, bits, h, w, "fused")
        rc, n = ws.sync()
        assert rc == 0
        lists.append(ws.fetch(n))
    ws.close()
    say("16M", len(lists[0]), "signal pixels")
    assert len(lists[0]) > 1000 and lists[0].tobytes() == lists[1].tobytes()
    d.free()
    o.free()
