"""From an NXmx file to spots' threshold on the device: an EIGER2 4M series
written as bslz4 chunks over two data files (one frame stored raw, one never
written), its chunks read by the NXmx reader in one call, decoded by the
device decoder into frames in HBM in one launch, and those frames thresholded
in one batch -- every decoded frame byte for byte the frame it was made
from, every list equal to the oracle's dext<uint32_t>."""
import sys

import numpy as np
import pytest

from test_nxmx import frames_of, write_data, write_master, writer  # noqa: F401

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_nxmx_chunks_to_device_frames_to_threshold(tmp_path, writer, device, oracle_lib):
    from spotfinder_hip import codec
    sf = device
    try:
        sf.nxmx_lib()
    except ImportError as e:
        pytest.skip(str(e))
    desc = sf.synth_desc("4m", seed=12, persist=1)
    h, w = desc.height, desc.width
    frames = [sf.synth_frame(desc, z) for z in range(6)]
    chunks = [codec.bslz4_encode(f) for f in frames]
    chunks[2] = frames[2].tobytes()
    say("writing the series")
    write_data(writer, tmp_path / "d1.h5", "/entry/data/data", chunks[:3], h, w, 32,
               masks=[0, 0, 1])
    write_data(writer, tmp_path / "d2.h5", "/entry/data/data",
               [chunks[3], None, chunks[5]], h, w, 32)
    write_master(writer, tmp_path / "m.h5", 6, h, w, 32,
                 [("d1.h5", "/entry/data/data", 0, 3, 0),
                  ("d2.h5", "/entry/data/data", 3, 3, 0)])
    series = sf.NxmxSeries(tmp_path / "m.h5")
    dev = [sf.DeviceArray(h * w * 4) for _ in range(6)]
    arr, n, buf, present = series.read_chunks(0, dev)
    assert present == [1, 1, 1, 1, 0, 1] and n == 5
    say("decoding", n, "chunks")
    dec = sf.Decoder()
    dec.load_descriptors(arr, n, keep=buf)
    dec.run()
    assert dec.sync() == [0] * n
    have = [z for z in range(6) if present[z]]
    for z in have:
        assert dev[z].to_numpy(np.uint32, (h, w)).tobytes() == frames[z].tobytes(), z
    say("thresholding")
    batch = sf.Batch()
    batch.threshold([(dev[z].ptr, h, w, 32) for z in have], graph=True)
    counts, status = batch.sync()
    assert status == [0] * len(have)
    for i, z in enumerate(have):
        _, want = oracle_lib.dext(frames[z])
        assert batch.fetch(i, counts[i]).tobytes() == want.tobytes(), z
    batch.close()
    dec.close()
    series.close()
