"""End to end on the device: EIGER2 4M frames resident in HBM through the
threshold, the device labelling and the grouping (pixels on the host, or
kept on the device) into a DIALS reflection table, against the reference's
own CPU path (dext, Labeller and refl::write, compiled from its sources into
oracle/_ref) on the same frames.

Every column of the two tables is byte-identical except the spots' variances,
which the device reduces as exact integer moments and the reference sums in
floating point: those agree to 1e-9 (relative).
"""
import sys

import msgpack
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def columns(blob):
    doc = msgpack.unpackb(blob, raw=False, strict_map_key=False)
    table = doc[2]
    return doc[:2], table["identifiers"], table["nrows"], {
        k: (t, n, b) for k, (t, (n, b)) in table["data"].items()}


@pytest.mark.parametrize("device_pixels", [False, True])
@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_frames_to_reflection_table(device, tmp_path, mode, device_pixels):
    sf = device
    ref = O.reference()
    assert ref is not None, "oracle/_ref is not built"
    desc = sf.synth_desc("4m", spots_per_frame=600, persist=2, seed=91)
    h, w = desc.height, desc.width
    ws = sf.Workspace()
    lab = sf.Labeller(h, w, two_d=(mode == "2d"), device_pixels=device_pixels)
    rlab = ref.labeller(h, w, two_d=(mode == "2d"))
    for z in range(4):
        frame = sf.synth_frame(desc, z)
        d = sf.DeviceArray.from_numpy(frame)
        say(mode, "frame", z)
        ws.threshold(d.ptr, 32, h, w, "fused")
        rc, _ = ws.sync()
        assert rc == 0
        sf.workspace_label(ws, True)
        lab.add_workspace(z, ws)
        rc, want = ref.dext(frame)
        assert rc == 0
        rlab.add(z, want)
        d.free()
    lab.finish()
    rlab.finish()
    ws.close()
    path = tmp_path / "device.refl"
    lab.write_refl(path, id=0, identifier="sweep-1")
    ours = path.read_bytes()
    theirs = ref.refl_bytes(rlab.spots(), rlab.pixels(), w, id=0, identifier="sweep-1")
    head_a, ids_a, n_a, cols_a = columns(ours)
    head_b, ids_b, n_b, cols_b = columns(theirs)
    say(mode, n_a, "spots,", len(ours), "bytes")
    assert head_a == head_b and ids_a == ids_b and n_a == n_b and n_a > 300
    assert list(cols_a) == list(cols_b)
    for name in cols_a:
        if name == "xyzobs.px.variance":
            a = np.frombuffer(cols_a[name][2])
            b = np.frombuffer(cols_b[name][2])
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=0)
        else:
            assert cols_a[name] == cols_b[name], name
