"""The grouping with the pixels kept on the device (sf_labeller_create_device):
the host stitch sees components only, and each frame's pixels are regrouped by
component in HBM and gathered into the spots' pixel lists there.

Everything must equal the host path and the reference's Labeller exactly:
  * the reference suite's hand-derived cases, and its golden sweep (spots
    exact but for the variance, 1e-9; pixels and counts exact), 3D and 2D;
  * a sweep of long-lived components (pixels on for the whole sweep or for
    runs of frames, beside transient ones), which keeps staged frames open
    across many later frames and gives spots several components on a frame;
  * full-size frames thresholded and labelled in batches from HBM, whose
    spots continue across batch boundaries, against the oracle's Labeller
    and against the host path on the same batches.
"""
import sys

import pytest

from labeller_cases import compare_spots, run_cases
from oracle import oracle as O
from test_labeller import long_lived_sweep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_cases(device):
    run_cases(lambda h, w, **kw: device.Labeller(h, w, device_pixels=True, **kw),
              device.LabellerError)


@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_golden_sweep(device, sweep, mode):
    lab = device.Labeller(384, 320, two_d=(mode == "2d"), device_pixels=True)
    for z in range(10):
        lab.add(z, sweep[f"signal_{z}"])
    lab.finish()
    compare_spots(lab.spots(), sweep[f"spots_{mode}"])
    assert lab.pixels().tobytes() == sweep[f"pixels_{mode}"].tobytes()
    c = lab.counts()
    assert [c[k] for k in O.COUNTS_FIELDS] == list(sweep[f"counts_{mode}"])


def test_long_lived_components(device):
    kw = dict(max_spot_size=100000, max_separation=1000.0)
    lab = device.Labeller(128, 128, device_pixels=True, **kw)
    want = O.oracle().labeller(128, 128, **kw)
    for z, px in long_lived_sweep():
        lab.add(z, px)
        want.add(z, px)
    # The hot pixels' components hold every frame they touch until the end.
    assert lab.device_bytes() > 0
    lab.finish()
    want.finish()
    assert lab.device_bytes() == 0
    got = lab.spots()
    assert len(got) > 1000 and (got["n_signal"] == 160).sum() == 10
    compare_spots(got, want.spots())
    assert lab.pixels().tobytes() == want.pixels().tobytes()
    assert lab.counts() == want.counts()


@pytest.mark.parametrize("layout", [((0, 4), (4, 5), (5, 10)),
                                    ((0, 1), (1, 2), (2, 4), (4, 10))])
@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_batches_from_hbm(device, mode, layout):
    """Batches that grow after a smaller one: each batch's slots must leave
    the previous batch's last frame intact, since the next batch's first
    frame links to it."""
    desc = device.synth_desc("4m", spots_per_frame=600, persist=3, seed=63)
    h, w = desc.height, desc.width
    host = [device.synth_frame(desc, z) for z in range(10)]
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    on_device = device.Labeller(h, w, two_d=(mode == "2d"), device_pixels=True)
    on_host = device.Labeller(h, w, two_d=(mode == "2d"))
    ref = O.oracle().labeller(h, w, two_d=(mode == "2d"))
    batch = device.Batch()
    for start, stop in layout:
        say(mode, "batch", start, stop)
        batch.threshold([(dev[z].ptr, h, w, 32) for z in range(start, stop)],
                        graph=(start == layout[-1][0]))
        counts, status = batch.sync()
        assert status == [0] * (stop - start)
        batch.label(True)
        on_device.add_batch(start, batch)
        on_host.add_batch(start, batch)
    for z in range(10):
        _, want = O.oracle().dext(host[z])
        ref.add(z, want)
    on_device.finish()
    on_host.finish()
    ref.finish()
    assert on_device.device_bytes() == 0 and on_host.device_bytes() == 0
    spots = on_device.spots()
    assert len(spots) > 1000
    assert spots.tobytes() == on_host.spots().tobytes()
    assert on_device.pixels().tobytes() == on_host.pixels().tobytes()
    compare_spots(spots, ref.spots())
    assert on_device.pixels().tobytes() == ref.pixels().tobytes()
    assert on_device.counts() == ref.counts()
    batch.close()


def test_host_labelled_frames_are_refused(device):
    lab = device.Labeller(8, 10, device_pixels=True)
    with pytest.raises(device.LabellerError, match="device frames"):
        lab.add_labelled(0, [], [], [], [])
