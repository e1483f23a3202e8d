"""The serving loop (sf_server_*) on the device.

  * alone and in the mix: the frames of three sweeps (two EIGER2 4M uint32,
    one 512 x 512 uint16 ROI), and of four of the serving configuration's
    sizes (9M, 16M and 4M uint32, a uint16 ROI), served one at a time, then
    all submitted at once so that the worker batches them together across
    sweeps; every
    frame's ordered list, signal count, component count and link count are
    identical both ways, the lists equal the oracle's threshold, and the mix
    really was served in shared batches;
  * admission: a sweep that would leave less free device memory than the
    reserve is refused (-6);
  * a request can be waited for once;
  * compressed chunks from host memory (bitshuffle-LZ4, LZ4, raw) among
    device frames serve every frame as its device frame does; a chunk that
    does not describe its frame is refused, one with malformed data fails
    alone;
  * frames that take the exact path while the batch behind them, already on
    the device, links to them are served as when served alone.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SWEEPS = [("4m", dict(seed=31, spots_per_frame=600, persist=2)),
          ("4m", dict(seed=32, spots_per_frame=300, persist=1)),
          ((512, 512), dict(bits=16, seed=33, spots_per_frame=40,
                            ring_radius=150.0))]
SWEEPS_LARGE = [("9m", dict(seed=41, spots_per_frame=1200, persist=1)),
                ("16m", dict(seed=42, spots_per_frame=2000, persist=1)),
                ("4m", dict(seed=43, spots_per_frame=600, persist=1)),
                ((512, 512), dict(bits=16, seed=44, spots_per_frame=40,
                                  ring_radius=150.0))]
FRAMES = 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_data(device, sweeps, frames_per_sweep):
    out = []
    for kind, kw in sweeps:
        d = device.synth_desc(kind, **kw)
        frames = [device.synth_frame(d, z) for z in range(frames_per_sweep)]
        out.append((d, frames, [device.DeviceArray.from_numpy(f) for f in frames]))
    return out


@pytest.fixture(scope="module")
def data(device):
    return make_data(device, SWEEPS, FRAMES)


@pytest.fixture(scope="module")
def data_large(device):
    return make_data(device, SWEEPS_LARGE, 3)


def capacity(device, d):
    return d.height * d.width * (d.bits // 8) // 16


def serve(device, data, together, ring=None):
    frames = len(data[0][1])
    srv = device.Server(max_batch=64, ring=ring or frames + 1, label=True)
    ids = [srv.open_sweep(d.height, d.width, d.bits) for d, _, _ in data]
    results = {}
    order = [(s, z) for z in range(frames) for s in range(len(data))]
    pending = []
    for s, z in order:
        d, _, dev = data[s]
        out = np.zeros(capacity(device, d), device.SIGNAL_PIXEL)
        r = srv.submit(ids[s], dev[z].ptr, out)
        if together:
            pending.append((s, z, r, out))
        else:
            res = srv.wait(r)
            results[(s, z)] = (res, out[: res["signal_pixels"]].copy())
    for s, z, r, out in pending:
        res = srv.wait(r)
        results[(s, z)] = (res, out[: res["signal_pixels"]].copy())
    stats = srv.stats()
    srv.close()
    return results, stats


def test_each_result_is_the_same_alone_and_in_the_mix(device, data, oracle_lib):
    alone_and_mixed(device, data, oracle_lib)


def test_large_sweeps_alone_and_in_the_mix(device, data_large, oracle_lib):
    alone_and_mixed(device, data_large, oracle_lib)


def alone_and_mixed(device, data, oracle_lib):
    say("server: alone")
    alone, st_alone = serve(device, data, together=False)
    say("server: in the mix", st_alone)
    mixed, st_mixed = serve(device, data, together=True)
    say("server: mixed", st_mixed)
    assert st_alone["mean_batch"] == 1.0
    assert st_mixed["mean_batch"] > 1.0 and st_mixed["batches"] < len(mixed)
    for key, (res_a, list_a) in alone.items():
        res_m, list_m = mixed[key]
        s, z = key
        assert res_a["status"] == 0 and res_m["status"] == 0, key
        for field in ("signal_pixels", "components", "links"):
            assert res_a[field] == res_m[field], (key, field)
        assert list_a.tobytes() == list_m.tobytes(), key
        _, want = oracle_lib.dext(data[s][1][z])
        assert list_a.tobytes() == want.tobytes(), key
        # Where spots persist over frames, every frame after a sweep's first
        # links to its previous frame.
        if z > 0 and data[s][0].persist >= 2:
            assert res_a["links"] > 0, key


def test_admission_by_free_memory(device):
    srv = device.Server(reserve_bytes=1 << 62)
    with pytest.raises(device.NotAdmitted):
        srv.open_sweep(4362, 4148, 32)
    srv.close()


def test_a_request_is_waited_for_once(device, data):
    srv = device.Server(label=False)
    d, _, dev = data[2]
    sweep = srv.open_sweep(d.height, d.width, d.bits)
    r = srv.submit(sweep, dev[0].ptr)
    assert srv.wait(r)["status"] == 0
    with pytest.raises(device.SpotfinderError):
        srv.wait(r)
    srv.close()


def serve_chunks(device, data, encoders):
    """All frames submitted at once, frame z of sweep s as a compressed chunk
    from host memory (encoders[(s + z) % len]) or, for None, as its device
    frame; results and lists by (sweep, frame)."""
    frames = len(data[0][1])
    srv = device.Server(max_batch=64, ring=frames + 1, label=True)
    ids = [srv.open_sweep(d.height, d.width, d.bits) for d, _, _ in data]
    pending = []
    for z in range(frames):
        for s, (d, host, dev) in enumerate(data):
            out = np.zeros(capacity(device, d), device.SIGNAL_PIXEL)
            enc = encoders[(s + z) % len(encoders)]
            if enc is None:
                r = srv.submit(ids[s], dev[z].ptr, out)
            else:
                alg, chunk = enc(host[z])
                r = srv.submit_chunk(ids[s], chunk, alg, out)
            pending.append((s, z, r, out))
    results = {}
    for s, z, r, out in pending:
        res = srv.wait(r)
        results[(s, z)] = (res, out[: res["signal_pixels"]].copy())
    stats = srv.stats()
    srv.close()
    return results, stats


def test_chunks_are_served_like_device_frames(device, data):
    """Compressed chunks (bitshuffle-LZ4, LZ4 and stored raw) submitted from
    host memory among device frames: decoded on the server's stream ahead of
    the threshold, they give every frame the same list, signal, component and
    link counts as the device frames served one at a time."""
    from spotfinder_hip import codec
    alone, _ = serve(device, data, together=False)
    encoders = [lambda f: ("bslz4", codec.bslz4_encode(f)), None,
                lambda f: ("lz4", codec.lz4_encode(f)),
                lambda f: ("", f.tobytes())]
    say("server: chunks")
    got, st = serve_chunks(device, data, encoders)
    say("server: chunks", st)
    assert st["mean_batch"] > 1.0
    for key, (res_a, list_a) in alone.items():
        res_c, list_c = got[key]
        assert res_c["status"] == 0, key
        for field in ("signal_pixels", "components", "links"):
            assert res_a[field] == res_c[field], (key, field)
        assert list_a.tobytes() == list_c.tobytes(), key


def test_chunk_refused_or_malformed(device, data):
    """A chunk whose filter header does not describe the sweep's frame is
    refused at submission with the reference's message; one whose LZ4 data
    are garbage completes with status -1 and its message, and the requests
    around it are served."""
    from spotfinder_hip import codec
    d, host, dev = data[0]
    srv = device.Server(ring=8, label=True)
    sweep = srv.open_sweep(d.height, d.width, d.bits)
    small = codec.bslz4_encode(host[0][:-1])
    with pytest.raises(device.SpotfinderError, match="filter header says"):
        srv.submit_chunk(sweep, small, "bslz4")
    good = codec.bslz4_encode(host[0])
    # The first block's length stays; its LZ4 sequence becomes 0xFF bytes.
    bad = bytearray(good)
    bad[16:64] = b"\xff" * 48
    r0 = srv.submit_chunk(sweep, good, "bslz4")
    r1 = srv.submit_chunk(sweep, bytes(bad), "bslz4")
    r2 = srv.submit(sweep, dev[1].ptr)
    assert srv.wait(r0)["status"] == 0
    with pytest.raises(device.SpotfinderError, match="decompression failed|truncated|size"):
        srv.wait(r1)
    assert srv.wait(r2)["status"] == 0
    srv.close()


def test_exact_path_frames_in_overlapped_batches(device, oracle_lib):
    """Frames with a valid pixel beyond the fused kernel's exact range (re-run
    by the direct kernels when their batch is read) in a sweep whose next
    frames are already on the device in the batch behind: those are labelled
    again against the re-run frame, so every count and list equals serving
    the frames one at a time."""
    d = device.synth_desc("4m", seed=51, spots_per_frame=800, persist=2)
    host = [device.synth_frame(d, z) for z in range(6)]
    rng = np.random.default_rng(5)
    for z in (1, 3):
        f = host[z]
        ys, xs = rng.integers(0, d.height, 20), rng.integers(0, d.width, 20)
        ok = f[ys, xs] < 0xFFFFFFFE
        f[ys[ok], xs[ok]] = 300000
    other = device.synth_desc((512, 512), bits=16, seed=52, spots_per_frame=40,
                              ring_radius=150.0)
    data = [(d, host, [device.DeviceArray.from_numpy(f) for f in host])]
    ohost = [device.synth_frame(other, z) for z in range(6)]
    data.append((other, ohost, [device.DeviceArray.from_numpy(f) for f in ohost]))
    alone, _ = serve(device, data, together=False)
    # A ring of three holds a sweep's previous frame and two in flight: the
    # batch behind the one being read then holds the sweep's next frame.
    mixed, st = serve(device, data, together=True, ring=3)
    say("server: exact path mixed", st)
    assert st["batches"] >= 6
    for key, (res_a, list_a) in alone.items():
        res_m, list_m = mixed[key]
        for field in ("status", "signal_pixels", "components", "links"):
            assert res_a[field] == res_m[field], (key, field)
        assert list_a.tobytes() == list_m.tobytes(), key
        s, z = key
        _, want = oracle_lib.dext(data[s][1][z])
        assert list_a.tobytes() == want.tobytes(), key
