"""The allocation-free threshold (sf_threshold_noalloc) against the entry
point it varies (sf_ws_threshold + sf_ws_sync + sf_ws_fetch), through the C
ABI, record for record: calls in a row on frames of falling size (16M, 4M,
a 512 x 512 uint16 ROI), a frame beyond the fused kernel's exact range (the
re-run on the direct kernels, whose byte masks borrow the scratch list and
the caller's output), a frame over the list's capacity, the frames it
refuses, and a frame again after all of them.  The library keeps one scratch
array a device, so every call here reuses the last one's memory."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def via_workspace(device, ws, dev, f, bits):
    ws.threshold(dev.ptr, bits, f.shape[0], f.shape[1], "fused")
    rc, n = ws.sync()
    return rc, n, (ws.fetch(n) if rc == 0 else None)


def via_noalloc(device, dev, f, bits, out, records):
    rc, n = device.threshold_noalloc(dev.ptr, bits, f.shape[0], f.shape[1],
                                     out.ptr, records)
    got = out.to_numpy(device.SIGNAL_PIXEL, (n,)) if rc == 0 else None
    return rc, n, got


def test_noalloc_matches_the_workspace_call_after_call(device, oracle_lib):
    rng = np.random.default_rng(12)
    frames = []
    frames.append(("16M", device.synth_frame(device.synth_desc("16m", seed=8), 0)))
    frames.append(("4M", device.synth_frame(device.synth_desc("4m", seed=9), 1)))
    frames.append(("ROI u16", device.synth_frame(device.synth_desc(
        (512, 512), bits=16, seed=10, spots_per_frame=40, ring_radius=150.0), 0)))
    wide = rng.poisson(1.0, (300, 465)).astype(np.uint32)
    s = rng.uniform(size=wide.shape) < 0.01
    wide[s] = rng.integers(40000, 1 << 20, s.sum())
    frames.append(("beyond the fast range", wide))
    dense = np.zeros((256, 512), np.uint32)
    dense[:, ::2] = 10000
    frames.append(("over capacity", dense))
    frames.append(("tiny", np.array([[0, 9000, 0], [0, 0, 0]], np.uint32)))
    frames.append(("4M again", frames[1][1]))

    big = max(device.noalloc_output_records(f.shape[0], f.shape[1], 8 * f.itemsize)
              for _, f in frames)
    out = device.DeviceArray(big * 16)
    ws = device.Workspace()
    before = ws.fallbacks()
    for name, f in frames:
        bits = 8 * f.itemsize
        dev = device.DeviceArray.from_numpy(f)
        say("noalloc", name, f.shape, bits)
        want = via_workspace(device, ws, dev, f, bits)
        records = device.noalloc_output_records(f.shape[0], f.shape[1], bits)
        got = via_noalloc(device, dev, f, bits, out, records)
        assert got[0] == want[0] and got[1] == want[1], (name, got[:2], want[:2])
        if want[0] == 0:
            assert got[2].tobytes() == want[2].tobytes(), name
            _, ref = oracle_lib.dext(f)
            assert got[2].tobytes() == ref.tobytes(), name
        dev.free()
    assert ws.fallbacks() > before  # the wide frame took the exact path
    ws.close()
    out.free()


def test_noalloc_refusals(device):
    dev = device.DeviceArray(16)
    out = device.DeviceArray(16)
    with pytest.raises(ValueError, match="scratch"):
        device.threshold_noalloc(dev.ptr, 32, 6000, 6000, out.ptr, 1 << 30)
    with pytest.raises(ValueError, match="output list"):
        device.threshold_noalloc(dev.ptr, 32, 64, 64, out.ptr, 1)
    with pytest.raises(ValueError):
        device.threshold_noalloc(dev.ptr, 8, 64, 64, out.ptr, 1 << 20)
    with pytest.raises(ValueError):
        device.threshold_noalloc(None, 32, 64, 64, out.ptr, 1 << 20)
    dev.free()
    out.free()
