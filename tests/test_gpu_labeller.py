"""The device labelling and the grouping built on it, on the device.

  * the device's components, labels, moments and frame links equal the
    plain-Python model's exactly (tests/labelling_model.py);
  * sf.Labeller (device labelling + host grouping) passes the reference
    suite's hand-derived cases and reproduces the reference Labeller's
    golden sweep: spots exact but for the variance (1e-9), pixels and counts
    exact, in 3D and 2D;
  * threshold -> label -> group on full-size EIGER2 frames from HBM equals
    the oracle's threshold grouped by the oracle's Labeller.
"""
import sys

import numpy as np
import pytest

from labeller_cases import compare_spots, run_cases
from labelling_model import label_frame, links_between
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_device_labels_equal_the_model(device, sweep):
    ws = device.Workspace()
    prev = None
    for z in range(10):
        sig = sweep[f"signal_{z}"]
        say("label frame", z, len(sig))
        device.workspace_load_signal(ws, sig, 384, 320)
        comp_of, comps, links, linked = device.workspace_label(ws, True)
        m_comp_of, m_comps, m_rank = label_frame(sig, 320)
        assert (comp_of == m_comp_of).all()
        assert comps.tobytes() == m_comps.tobytes()
        assert linked == (z > 0)
        if prev is not None:
            want = links_between(sig, m_comp_of, prev[0], prev[1])
            got = np.unique(links, axis=0) if len(links) else links
            assert got.tolist() == want.tolist()
        prev = (m_rank, m_comp_of)
    ws.close()


def wide_signal(rng, h, w):
    """An ascending signal list over an h x w frame whose rows hold runs of
    every length from 1 to a few hundred pixels, some ending on and some
    crossing 32-pixel words, 2048-pixel chunks and 6144-pixel rounds of the
    device's row passes, and runs that touch the rows above and below."""
    mask = np.zeros((h, w), bool)
    for y in range(h):
        x = int(rng.integers(0, 40))
        while x < w:
            n = int(rng.choice([1, 2, 5, 31, 32, 33, 64, 97, 300]))
            mask[y, x:x + n] = True
            x += n + int(rng.integers(1, 400))
        for edge in (2047, 2048, 6143, 6144, 12287):
            if edge < w and rng.uniform() < 0.5:
                mask[y, max(0, edge - 20):edge + 20] = True
    idx = np.flatnonzero(mask).astype(np.uint32)
    sig = np.zeros(len(idx), O.SIGNAL_PIXEL)
    sig["index"] = idx
    sig["value"] = rng.integers(1, 1 << 20, len(idx))
    return sig


def test_device_labels_equal_the_model_on_wide_rows(device):
    rng = np.random.default_rng(77)
    h, w = 24, 13000
    ws = device.Workspace()
    prev = None
    for z in range(3):
        sig = wide_signal(rng, h, w)
        say("wide rows", z, len(sig))
        device.workspace_load_signal(ws, sig, h, w)
        comp_of, comps, links, linked = device.workspace_label(ws, True)
        m_comp_of, m_comps, m_rank = label_frame(sig, w)
        assert (comp_of == m_comp_of).all()
        assert comps.tobytes() == m_comps.tobytes()
        if prev is not None:
            want = links_between(sig, m_comp_of, prev[0], prev[1])
            got = np.unique(links, axis=0) if len(links) else links
            assert got.tolist() == want.tolist()
        prev = (m_rank, m_comp_of)
    ws.close()


def test_device_labeller_cases(device):
    run_cases(lambda h, w, **kw: device.Labeller(h, w, **kw),
              device.LabellerError)


@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_device_labeller_golden_sweep(device, sweep, mode):
    lab = device.Labeller(384, 320, two_d=(mode == "2d"))
    for z in range(10):
        lab.add(z, sweep[f"signal_{z}"])
    lab.finish()
    compare_spots(lab.spots(), sweep[f"spots_{mode}"])
    assert lab.pixels().tobytes() == sweep[f"pixels_{mode}"].tobytes()
    c = lab.counts()
    assert [c[k] for k in O.COUNTS_FIELDS] == list(sweep[f"counts_{mode}"])


@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_threshold_label_group_from_hbm(device, mode):
    desc = device.synth_desc("4m", spots_per_frame=600, persist=1, seed=77)
    h, w = desc.height, desc.width
    ws = device.Workspace()
    lab = device.Labeller(h, w, two_d=(mode == "2d"))
    ref = O.oracle().labeller(h, w, two_d=(mode == "2d"))
    for z in range(5):
        frame = device.synth_frame(desc, z)
        d = device.DeviceArray.from_numpy(frame)
        say(mode, "frame", z)
        ws.threshold(d.ptr, 32, h, w, "fused")
        rc, n = ws.sync()
        assert rc == 0
        _ = device.workspace_label(ws, True)
        lab.add_workspace(z, ws)
        _, want = O.oracle().dext(frame)
        ref.add(z, want)
    lab.finish()
    ref.finish()
    compare_spots(lab.spots(), ref.spots())
    assert lab.pixels().tobytes() == ref.pixels().tobytes()
    assert lab.counts() == ref.counts()
    assert len(lab.spots()) > 1000
    ws.close()
