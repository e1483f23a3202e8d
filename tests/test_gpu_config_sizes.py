"""The benchmark configurations' own sizes, checked record for record.

  * consecutive EIGER2-XE 16M uint32 frames (4148 x 4362, 2000 spots a frame
    persisting over frames): each frame's list against the oracle's
    dext<uint32_t>, then grouped in 2D and in 3D, one frame at a time
    through the workspace (the plain per-frame path) and in batches with the
    pixels on the host and on the device, every spot, pixel and count
    against the oracle's Labeller (src/spots/dials_spots.cc:57-360; variance
    within 1e-9, everything else exact);
  * 4M batches of 32 and 128 frames in one launch, replayed from the
    captured graph with other frames, every frame's list against the oracle.
"""
import sys

import numpy as np
import pytest

from labeller_cases import compare_spots
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


FRAMES_16M = 4


@pytest.fixture(scope="module")
def sweep16(device):
    desc = device.synth_desc("16m", spots_per_frame=2000, persist=1, seed=2026)
    host = [device.synth_frame(desc, z) for z in range(FRAMES_16M)]
    want = [O.oracle().dext(f)[1] for f in host]
    return desc, host, want


def oracle_grouping(desc, lists, two_d):
    ref = O.oracle().labeller(desc.height, desc.width, two_d=two_d)
    for z, sig in enumerate(lists):
        ref.add(z, sig)
    ref.finish()
    return ref


def check_grouping(lab, ref):
    lab.finish()
    spots = lab.spots()
    assert len(spots) > 1000
    compare_spots(spots, ref.spots())
    assert lab.pixels().tobytes() == ref.pixels().tobytes()
    assert lab.counts() == ref.counts()


@pytest.mark.parametrize("mode", ["2d", "3d"])
def test_16m_plain_per_frame(device, sweep16, mode):
    desc, host, want = sweep16
    h, w = desc.height, desc.width
    ws = device.Workspace()
    lab = device.Labeller(h, w, two_d=(mode == "2d"))
    for z, f in enumerate(host):
        d = device.DeviceArray.from_numpy(f)
        say("16M plain", mode, "frame", z)
        ws.threshold(d.ptr, 32, h, w, "fused")
        rc, n = ws.sync()
        assert rc == 0 and n == len(want[z]), (z, n, len(want[z]))
        assert ws.fetch(n).tobytes() == want[z].tobytes(), z
        device.workspace_label(ws, True)
        lab.add_workspace(z, ws)
        d.free()
    check_grouping(lab, oracle_grouping(desc, want, mode == "2d"))
    ws.close()


@pytest.mark.parametrize("device_pixels", [False, True])
@pytest.mark.parametrize("mode", ["2d", "3d"])
def test_16m_batches(device, sweep16, mode, device_pixels):
    """Batches of one then three frames, so that spots cross the batch
    boundary and the second batch's first frame links to the first batch's
    last."""
    desc, host, want = sweep16
    h, w = desc.height, desc.width
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    lab = device.Labeller(h, w, two_d=(mode == "2d"), device_pixels=device_pixels)
    batch = device.Batch()
    for start, stop in ((0, 1), (1, FRAMES_16M)):
        say("16M batch", mode, "device pixels", device_pixels, start, stop)
        batch.threshold([(dev[z].ptr, h, w, 32) for z in range(start, stop)],
                        graph=(start > 0))
        counts, status = batch.sync()
        assert status == [0] * (stop - start)
        for i in range(stop - start):
            assert batch.fetch(i, counts[i]).tobytes() == want[start + i].tobytes()
        batch.label(True)
        lab.add_batch(start, batch)
    check_grouping(lab, oracle_grouping(desc, want, mode == "2d"))
    if device_pixels:
        assert lab.device_bytes() == 0
    batch.close()
    for d in dev:
        d.free()


@pytest.mark.parametrize("n", [32, 128])
def test_4m_batch_graph_replays(device, n):
    """A batch of n 4M frames captured once and replayed with other frames of
    the same shape in the slots (the second replay shifts the frames by
    half a batch), each replay checked frame by frame."""
    desc = device.synth_desc("4m", seed=330 + n, persist=1)
    h, w = desc.height, desc.width
    total = n + n // 2
    host = [device.synth_frame(desc, z) for z in range(total)]
    want = [O.oracle().dext(f)[1] for f in host]
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    batch = device.Batch()
    for rep, first in enumerate((0, n // 2, 0)):
        say("4M batch", n, "replay", rep, "first frame", first)
        batch.threshold([(dev[first + i].ptr, h, w, 32) for i in range(n)], graph=True)
        counts, status = batch.sync()
        assert status == [0] * n
        for i in range(n):
            got = batch.fetch(i, counts[i])
            assert got.tobytes() == want[first + i].tobytes(), (rep, i)
    batch.close()
    for d in dev:
        d.free()
