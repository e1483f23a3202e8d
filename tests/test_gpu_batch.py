"""Batches: many frames per persistent launch, on the device.

  * an interleaved batch of frame sizes and pixel types -- EIGER2 4M, 9M and
    16M uint32 frames, 512 x 512 uint16 ROIs, and the golden edge matrix --
    thresholded in one launch, every frame bit-exact against the reference's
    own output (golden vectors) or the oracle;
  * the same through a captured hipGraph, replayed several times with other
    frames in between eager work, every replay checked;
  * a frame over the list's capacity (-2) and a frame beyond the fused
    kernel's exact range (re-run exactly) inside a batch, the others intact;
  * batch labelling, linked across batch boundaries, grouped into spots, equal
    to the oracle's Labeller on the oracle's threshold.
"""
import hashlib
import sys

import numpy as np
import pytest

from labeller_cases import compare_spots
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def frame_of(device, case):
    kind = case["kind"] if isinstance(case["kind"], str) else tuple(case["kind"])
    f = device.synth_frame(device.synth_desc(kind, **case["overrides"]), case["frame"])
    assert hashlib.sha256(f.tobytes()).hexdigest() == case["frame_sha256"]
    return f


def mixed_frames(device, synth_cases, matrix):
    """(name, host frame, expected check) in an interleaved order."""
    out = []
    names = ["eiger4m_f0", "roi512_u16", "eiger9m_f0", "roi512_u16",
             "eiger16m_f0", "eiger4m_f7_seed3", "roi512_u16", "eiger4m_dark"]
    for nm in names:
        out.append((nm, frame_of(device, synth_cases[nm]), synth_cases[nm]))
    for nm in ("m257x511_2.0_32", "m67x300_12.0_16", "m1x1_0.3_16",
               "m33x65_2.0_32"):
        f, expected = matrix[nm]
        out.append((nm, f, expected))
    return out


def check_frame(name, got, expected):
    if isinstance(expected, dict):
        assert len(got) == expected["count"], name
        assert hashlib.sha256(got.tobytes()).hexdigest() == expected["signal_sha256"], name
    else:
        assert got.tobytes() == expected.tobytes(), name


def run_batch(device, batch, frames, graph):
    dev = [device.DeviceArray.from_numpy(f) for _, f, _ in frames]
    descs = [(d.ptr, f.shape[0], f.shape[1], 16 if f.dtype == np.uint16 else 32)
             for d, (_, f, _) in zip(dev, frames)]
    batch.threshold(descs, graph=graph)
    counts, status = batch.sync()
    for i, (name, f, expected) in enumerate(frames):
        assert status[i] == 0, name
        check_frame(name, batch.fetch(i, counts[i]), expected)
    return dev


def test_item_partitions_cover_every_item(device, oracle_lib):
    """Batches whose fused item counts fall on and around the launch's
    partitioning (chunks of eight items dealt over up to eight counters):
    fewer items than partitions, less than a chunk a partition, ragged last
    chunks, 40-row and automatic segments; every frame's list equal to the
    oracle's."""
    rng = np.random.default_rng(5)
    frames = []
    for k in range(11):
        h, w = int(rng.integers(20, 300)), int(rng.integers(240, 1300))
        f = rng.poisson(1.5, (h, w)).astype(np.int64)
        n = h * w // 400
        f[rng.integers(0, h, n), rng.integers(0, w, n)] += rng.integers(30, 3000, n)
        f[rng.uniform(size=(h, w)) < 0.01] = 0xFFFFFFFF
        frames.append(f.astype(np.uint32))
    want = [oracle_lib.dext(f)[1] for f in frames]
    dev = [device.DeviceArray.from_numpy(f) for f in frames]
    start = device.get_tuning()
    try:
        for seg in (40, 0):
            device.set_tuning(seg_rows=seg)
            batch = device.Batch()
            for n in (1, 2, 3, 5, 8, 11):
                say("partitions", seg, n)
                batch.threshold([(dev[i].ptr, f.shape[0], f.shape[1], 32)
                                 for i, f in enumerate(frames[:n])])
                counts, status = batch.sync()
                for i in range(n):
                    assert status[i] == 0, (seg, n, i)
                    assert batch.fetch(i, counts[i]).tobytes() == want[i].tobytes(), (seg, n, i)
            batch.close()
    finally:
        device.set_tuning(**start)


@pytest.mark.parametrize("graph,kernel", [(False, "fused"), (True, "fused"),
                                          (True, "direct")])
def test_mixed_batch_bit_exact(device, synth_cases, matrix, graph, kernel):
    """With graphs on, each replay takes new frame arrays of the same shapes:
    the fused sequence reads every pointer from the descriptors, and the
    direct kernels, which take theirs as launch arguments, are not captured."""
    frames = mixed_frames(device, synth_cases, matrix)
    device.set_kernel(kernel)
    try:
        batch = device.Batch()
        say("mixed batch", len(frames), "graph", graph, kernel)
        replay_mixed(device, batch, frames, graph)
    finally:
        device.set_kernel("fused")


def replay_mixed(device, batch, frames, graph):
    run_batch(device, batch, frames, graph)
    if graph:
        # Replays with other frames of the same shapes, eager work between.
        for rep in range(3):
            shuffled = frames[:]
            # swap the two 4M frames and the roi copies: same shapes,
            # different contents per slot
            i0, i1 = 0, 5
            shuffled[i0], shuffled[i1] = shuffled[i1], shuffled[i0]
            if rep % 2 == 0:
                frames = shuffled
            say("replay", rep)
            run_batch(device, batch, frames, True)
            ws = device.Workspace()
            f = frames[0][1]
            d = device.DeviceArray.from_numpy(f)
            ws.threshold(d.ptr, 32, f.shape[0], f.shape[1], "fused")
            assert ws.sync()[0] == 0
            ws.close()
    batch.close()


def test_batch_with_overfull_and_wide_frames(device, synth_cases):
    dense = np.zeros((256, 512), np.uint32)
    dense[:, ::2] = 10000
    wide = frame_of(device, synth_cases["wide_values_u32"])
    ok = frame_of(device, synth_cases["eiger4m_f0"])
    frames = [dense, wide, ok]
    dev = [device.DeviceArray.from_numpy(f) for f in frames]
    batch = device.Batch()
    batch.threshold([(d.ptr, f.shape[0], f.shape[1], 32)
                     for d, f in zip(dev, frames)])
    counts, status = batch.sync()
    assert status == [-2, 0, 0]
    check_frame("wide", batch.fetch(1, counts[1]), synth_cases["wide_values_u32"])
    check_frame("ok", batch.fetch(2, counts[2]), synth_cases["eiger4m_f0"])
    assert batch.fallbacks() >= 1
    batch.close()


@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_batch_label_group(device, mode):
    desc = device.synth_desc("4m", spots_per_frame=600, persist=1, seed=91)
    h, w = desc.height, desc.width
    host = [device.synth_frame(desc, z) for z in range(7)]
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    lab = device.Labeller(h, w, two_d=(mode == "2d"))
    ref = O.oracle().labeller(h, w, two_d=(mode == "2d"))
    batch = device.Batch()
    for start, stop in ((0, 3), (3, 6), (6, 7)):
        say(mode, "batch", start, stop)
        batch.threshold([(dev[z].ptr, h, w, 32) for z in range(start, stop)],
                        graph=(start == 3))
        counts, status = batch.sync()
        assert status == [0] * (stop - start)
        batch.label(True)
        lab.add_batch(start, batch)
    for z in range(7):
        _, want = O.oracle().dext(host[z])
        ref.add(z, want)
    lab.finish()
    ref.finish()
    compare_spots(lab.spots(), ref.spots())
    assert lab.pixels().tobytes() == ref.pixels().tobytes()
    assert lab.counts() == ref.counts()
    batch.close()


def test_autotune_installs_a_shape_and_results_do_not_change(device):
    """sf_autotune measures and installs a launch shape; whatever it picks,
    every frame's list is the same as under the starting shape."""
    desc = device.synth_desc("4m", spots_per_frame=500, persist=2, seed=41)
    host = [device.synth_frame(desc, z) for z in range(16)]
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    descs = [(d.ptr, desc.height, desc.width, 32) for d in dev]
    start = device.get_tuning()
    batch = device.Batch()
    batch.threshold(descs)
    counts, status = batch.sync()
    before = [batch.fetch(i, counts[i]) for i in range(len(descs))]
    try:
        say("autotune")
        report = device.autotune(descs, label=True)
        say(report)
        assert report["shapes_tried"] == 42  # 2 lane widths x 3 block sizes x 7 segmentings
        assert report["launch_sizes_tried"] >= 1
        assert report["best"] == device.get_tuning()
        assert report["fused_ms_per_frame"] > 0 and report["step_ms_per_frame"] > 0
        batch.threshold(descs)
        counts2, status2 = batch.sync()
        assert list(counts2) == list(counts) and list(status2) == list(status)
        for i in range(len(descs)):
            assert batch.fetch(i, counts2[i]).tobytes() == before[i].tobytes(), i
    finally:
        device.set_tuning(**start)


def test_mix_shape_is_kept_for_its_size_only(device):
    """sf_autotune_mix measures a shape on a mixed batch and keeps it for
    batches of that size: the process-wide tuning is unchanged, a batch of
    the same frames reports (and launches with) the kept shape, a batch of a
    very different size does not, every frame's list is unchanged under it,
    and clearing forgets it."""
    mix = [device.synth_desc("4m", spots_per_frame=400, persist=1, seed=61),
           device.synth_desc((512, 512), bits=16, seed=62, spots_per_frame=30,
                             ring_radius=150.0),
           device.synth_desc("9m", spots_per_frame=800, persist=1, seed=63)]
    host = [device.synth_frame(d, 0) for d in mix] * 2
    descs_all = [d for d in mix] * 2
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    descs = [(a.ptr, d.height, d.width, d.bits) for a, d in zip(dev, descs_all)]
    start = device.get_tuning()
    device.clear_mix_tunings()
    batch = device.Batch()
    batch.threshold(descs)
    counts, status = batch.sync()
    before = [batch.fetch(i, counts[i]) for i in range(len(descs))]
    try:
        report = device.autotune_mix(descs)
        say("mix shape", report)
        assert report["shapes_tried"] == 42 and report["fused_ms_per_frame"] > 0
        assert device.get_tuning() == start
        shape, kept = device.tuning_for_mix(descs)
        assert kept and shape["seg_rows"] == report["best"]["seg_rows"]
        for k in ("target_waves", "waves_per_block", "lane_columns"):
            assert shape[k] == report["best"][k], k
        assert shape["frames_per_launch"] == start["frames_per_launch"]
        _, kept_small = device.tuning_for_mix(descs[1:2])
        assert not kept_small
        for graph in (False, True):
            batch.threshold(descs, graph=graph)
            counts2, status2 = batch.sync()
            assert list(counts2) == list(counts) and list(status2) == list(status)
            for i in range(len(descs)):
                assert batch.fetch(i, counts2[i]).tobytes() == before[i].tobytes(), (graph, i)
        device.clear_mix_tunings()
        assert device.tuning_for_mix(descs) == (start, False)
    finally:
        device.clear_mix_tunings()
        device.set_tuning(**start)
        batch.close()


def test_batch_size_ramp_keeps_slots_bounded(device):
    """Batches of 1, 2, ..., 8 frames in turn: each links its first frame to
    the previous batch's last, the grouping across all of them equals the
    oracle's, and the batch never holds more than the buffers of the largest
    batch plus one frame (measured against a fresh batch of nine)."""
    desc = device.synth_desc((640, 480), spots_per_frame=60, persist=1, seed=71)
    h, w = desc.height, desc.width
    total = sum(range(1, 9))
    host = [device.synth_frame(desc, z) for z in range(total)]
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    lab = device.Labeller(h, w)
    ref = O.oracle().labeller(h, w)
    batch = device.Batch()
    z = 0
    for n in range(1, 9):
        say("ramp batch", n)
        batch.threshold([(dev[k].ptr, h, w, 32) for k in range(z, z + n)], graph=(n % 2 == 0))
        counts, status = batch.sync()
        assert status == [0] * n
        batch.label(True)
        lab.add_batch(z, batch)
        z += n
    ramp_bytes = batch.device_bytes()
    batch.close()
    for k in range(total):
        ref.add(k, O.oracle().dext(host[k])[1])
    lab.finish()
    ref.finish()
    compare_spots(lab.spots(), ref.spots())
    assert lab.pixels().tobytes() == ref.pixels().tobytes()
    assert lab.counts() == ref.counts()
    nine = device.Batch()
    nine.threshold([(dev[k].ptr, h, w, 32) for k in range(9)])
    nine.sync()
    nine.label(True)
    nine.labels(0)
    assert ramp_bytes <= nine.device_bytes(), (ramp_bytes, nine.device_bytes())
    nine.close()
