"""The benchmark's host-side helpers, on the CPU: the committed traffic
measurements are found for exactly the batch size and launch shapes they were
taken at, the host core count is bounded by the affinity mask, and a request
for more than one GPU is refused before any device is touched."""
import json
import os
import subprocess
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_traffic_found_for_each_measured_shape():
    with open(bench.TRAFFIC) as fh:
        m = json.load(fh)
    frames = m["frames_per_launch"]
    assert m["measurements"]
    for e in m["measurements"]:
        tuning = dict(e["tuning"], frames_per_launch=frames)
        got, src = bench.measured_traffic(frames, tuning)
        assert got == e["traffic_bytes_per_launch"]
        assert src == "profiles/" + os.path.basename(bench.TRAFFIC)
        # every measured shape reads at least the frames' pixels once
        assert got > frames * 4148 * 4362 * 4


def test_traffic_absent_for_other_batches_and_shapes():
    with open(bench.TRAFFIC) as fh:
        m = json.load(fh)
    frames = m["frames_per_launch"]
    tuning = dict(m["measurements"][0]["tuning"])
    assert bench.measured_traffic(frames + 1, tuning) == (None, None)
    other = dict(tuning, lane_columns=8 if tuning["lane_columns"] != 8 else 4)
    assert bench.measured_traffic(frames, other) == (None, None)
    assert bench.measured_traffic(frames, dict(tuning, seg_rows=1000)) == (None, None)


def test_host_cores_within_affinity():
    n, info = bench.host_cores()
    assert 1 <= n <= len(os.sched_getaffinity(0))
    assert info["affinity"] == len(os.sched_getaffinity(0))


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
