"""Benchmark of the spot-finding hot path on one MI355X.

Workload: EIGER2-XE 16M frames, 4148 x 4362 uint32, synthetic -- Poisson
background 0.8 with a solvent ring (+1.5 at r = 600 px, sigma 40), 2000
Gaussian spots per frame (sigma 0.6-1.5 px, peak 20-5000) persisting 1-3
frames, 8 x 4 module gaps masked, 0.01 % hot pixels at 65535.  The path:
the dispersion threshold (7x7 / erode 5x5 / 11x11) into an ordered
SignalPixel list and a strong-pixel bitmask, then four-connected labelling
with exact per-component moments (2D), all on the device.

Frames are generated on the host, uploaded once and resident in HBM before
the timed region.  A step is one batch of frames (--frames) through one
persistent threshold launch, the batched ordering and the batched labelling.
Two batch handles, each on its own stream, alternate (--in-flight 2, the
default): a step is queued before the one before it is waited for, and every
step's results are waited for and checked (--in-flight 1 waits for each step
before the next).  `value` is frames per second over the timed steps.
Beside it: the plain per-frame mode (one frame per launch sequence, waiting
for each), the fused kernel's achieved HBM bandwidth against the 8 TB/s peak
from HIP events on its stream, and the reference's own CPU threshold
(compiled from its sources, oracle/_ref) on this host's cores over a bounded
sample.  The
`decompress` object measures the step before the path: 16M frames stored as
bslz4 chunks (the EIGER's bitshuffle-LZ4 HDF5 filter), resident in HBM and
decoded into frames on the device.  The `narrow` object measures
--gpu-force's narrowing of those frames to 16 bits with mxi_max's scan, and
the `refl` object the reflection-table writer beside the reference's.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dials-metal_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = ("16M frames/s + achieved HBM GB/s (bit-exact mask); "
          "serving p50/p99 frame latency")


TRAFFIC = os.path.join(ROOT, "profiles", "fused_kernel_16m_batch32_traffic.json")


def measured_traffic(frames, tuning):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC measurements (FETCH_SIZE and WRITE_SIZE scaled by a device copy of
    known size), when one was taken at this batch size and launch shape."""
    try:
        with open(TRAFFIC) as fh:
            m = json.load(fh)
    except (OSError, ValueError):
        return None, None
    if m.get("frames_per_launch") != frames:
        return None, None
    for e in m.get("measurements", []):
        if all(tuning.get(k) == v for k, v in e["tuning"].items()):
            return e["traffic_bytes_per_launch"], "profiles/" + os.path.basename(TRAFFIC)
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """The host cores this process can run on, as mxi_find's default -j
    (std::thread::hardware_concurrency(), find_spots.cc:276-278) would use
    them: the CPUs in this process's affinity mask, bounded by a cgroup CPU
    quota where one is set; with what was read, for the record."""
    info = {"cpu_count": os.cpu_count()}
    n = len(os.sched_getaffinity(0))
    info["affinity"] = n
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            quota, period = fh.read().split()
        if quota != "max":
            q = -(-int(quota) // int(period))
            info["cgroup_quota_cpus"] = q
            n = min(n, max(1, q))
    except (OSError, ValueError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, info


def cpu_baseline(frames_host):
    """The reference's dext<uint32_t> (oracle/_ref, compiled from the
    reference's own sources) on every host core this process may use, one
    frame per thread at a time as mxi_find -j N spreads them, over a bounded
    sample: two frames a thread, cycled from the benchmark's frames."""
    from oracle import oracle as O
    threads, info = host_cores()
    lib, kind = O.reference(), "reference"
    if lib is None:
        lib, kind = O.oracle(), "port"
    total = 2 * threads
    sample = np.stack(frames_host)
    signal, secs = lib.threads_u32(sample, threads, total=total)
    return {"value": total / secs, "unit": "frames/s", "cores": threads, "kind": kind,
            "host": info,
            "sample": f"{total} of the benchmark's 16M uint32 frames (its "
                      f"{len(frames_host)} cycled) through dext<uint32_t> on "
                      f"{threads} threads, one frame per thread at a time: "
                      f"{secs:.2f} s wall, {signal} signal pixels; threshold "
                      f"only, no grouping"}


def plain_host_fed(sf, frames_host, reps=2):
    """gpu::find's own path (dext_cuda.cu:762-849 in the reference): a host
    frame copied to the device, thresholded and ordered there, and its list
    copied back, one frame at a time on one thread; frames in pinned memory
    (the reference's FrameBuffer takes host_alloc's, find_spots.cc:351-367)
    and in ordinary pageable memory."""
    out = {}
    n = len(frames_host)
    pinned = []
    for f in frames_host[: min(n, 8)]:
        p = sf.lib().sf_host_alloc(f.nbytes)
        if not p:
            raise RuntimeError("sf_host_alloc failed")
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)),
                                    shape=f.shape)
        arr[...] = f
        pinned.append((p, arr))
    try:
        for name, frames in (("pinned", [a for _, a in pinned]),
                             ("pageable", frames_host[: len(pinned)])):
            for f in frames[:2]:
                sf.find(f)
            t0 = time.perf_counter()
            count = 0
            for _ in range(reps):
                for f in frames:
                    rc, _ = sf.find(f)
                    if rc != 0:
                        raise RuntimeError(f"plain host-fed: status {rc}")
                    count += 1
            out[name] = count / (time.perf_counter() - t0)
    finally:
        for p, _ in pinned:
            sf.lib().sf_host_free(p)
    return {"value": out["pinned"], "unit": "frames/s",
            "pageable_frames_per_s": out["pageable"],
            "workload": "gpu::find on 16M uint32 host frames: upload, fused "
                        "threshold, ordering, list copied back, one frame at a "
                        "time, unlabelled; value with the frames in pinned memory"}


def serving(sf, seconds, seed, scale=3.0):
    """Six concurrent sweeps -- two each of EIGER2 4M, 9M and 16M uint32 --
    emitting frames by Poisson arrival at 100-750 Hz each (rates drawn from a
    seeded generator), replayed into the serving loop on the same schedule:
    continuous batching (every frame that has arrived, across sweeps, in one
    launch, two launches in flight) and one frame per launch; then the same
    sweeps at `scale` times the rates, batched; and, host-fed, with every
    frame arriving as its bitshuffle-LZ4 chunk in host memory (decoded on the
    device ahead of the threshold), at half and at the full rates.  Threshold, ordering and
    2D labelling linked to each sweep's previous frame; device frames resident
    in HBM (four distinct frames per sweep, cycled)."""
    from spotfinder_hip import codec
    rng = np.random.default_rng(seed)
    kinds = ["4m", "4m", "9m", "9m", "16m", "16m"]
    rates = rng.uniform(100.0, 750.0, len(kinds))
    descs, frames, chunks = [], [], []
    for i, kind in enumerate(kinds):
        d = sf.synth_desc(kind, seed=1000 + i,
                          spots_per_frame=int(rng.integers(300, 2000)), persist=2)
        descs.append(d)
        host = [sf.synth_frame(d, z) for z in range(4)]
        frames.append([sf.DeviceArray.from_numpy(f) for f in host])
        chunks.append([codec.bslz4_encode(f) for f in host])

    def schedule(k):
        ev_rng = np.random.default_rng(seed + 1)
        events = []
        for i, r in enumerate(rates * k):
            t = ev_rng.exponential(1.0 / r)
            while t < seconds:
                events.append((t, i))
                t += ev_rng.exponential(1.0 / r)
        events.sort()
        return events

    def run(events, max_batch, host_fed):
        srv = sf.Server(max_batch=max_batch, ring=8, label=True)
        ids = [srv.open_sweep(d.height, d.width, 32) for d in descs]
        taken = [0] * len(kinds)

        def replay(evs):
            items = []
            for _, i in evs:
                src = chunks if host_fed else frames
                items.append(src[i][taken[i] % 4])
                taken[i] += 1
            sweeps = [ids[i] for _, i in evs]
            arrivals = [t for t, _ in evs]
            if host_fed:
                res = srv.replay_chunks(sweeps, items, arrivals)
            else:
                res = srv.replay(sweeps, [f.ptr for f in items], arrivals)
            bad = [r for r in res if r["status"] != 0]
            if bad:
                raise RuntimeError(f"serving: {len(bad)} frames refused")

        # Untimed: the first quarter second of the schedule, in which each
        # sweep's slots and (host-fed) frame buffers and decoders are made.
        replay([e for e in events if e[0] < 0.25])
        srv.reset_stats()
        replay(events)
        st = srv.stats()
        srv.close()
        out = {k: round(st[k], 4) for k in
               ("p50_ms", "p99_ms", "max_ms", "mean_ms", "frames_per_s", "mean_batch")}
        out["offered_frames_per_s"] = round(len(events) / seconds, 1)
        return out

    base, fast, half = schedule(1.0), schedule(scale), schedule(0.5)
    raw = sum(sum(f.nbytes for f in fr) for fr in frames)
    packed = sum(sum(len(c) for c in cs) for cs in chunks)
    out = {"sweeps": kinds, "rates_hz": [round(float(r), 1) for r in rates],
           "seconds": seconds, "frames": len(base), "target_p99_ms": 2.0,
           "chunk_ratio": round(packed / raw, 4)}
    for name, events, max_batch, host_fed in (
            ("batched", base, 64, False), ("plain_per_frame", base, 1, False),
            (f"batched_x{scale:g}", fast, 64, False),
            ("host_fed_chunks_x0.5", half, 64, True),
            ("host_fed_chunks", base, 64, True)):
        out[name] = run(events, max_batch, host_fed)
        log(f"serving {name}: {out[name]}")
    return out


def uneven_batch(sf, reps, graph):
    """One launch over an interleaved stream of 4M, 9M and 16M uint32 frames
    and 512 x 512 uint16 ROIs in the ratio 3:2:4:8, backgrounds with Poisson
    means 0.2-5 and 100-5000 spots per frame: threshold, ordering and 2D
    labelling, under a launch shape measured for this mix (sf_autotune_mix);
    frames/s and the fused kernel's achieved HBM bandwidth over the batch's
    own bytes, and its time under the big batches' shape beside it."""
    rng = np.random.default_rng(17)
    kinds = ["4m"] * 3 + ["9m"] * 2 + ["16m"] * 4 + ["roi"] * 8
    rng.shuffle(kinds)
    dev, descs, byts = [], [], 0.0
    for i, kind in enumerate(kinds):
        bg = float(rng.uniform(0.2, 5.0))
        if kind == "roi":
            d = sf.synth_desc((512, 512), bits=16, seed=500 + i, background=bg,
                              spots_per_frame=int(rng.integers(100, 400)),
                              ring_radius=150.0)
        else:
            d = sf.synth_desc(kind, seed=500 + i, background=bg,
                              spots_per_frame=int(rng.integers(100, 5000)))
        f = sf.synth_frame(d, 0)
        dev.append(sf.DeviceArray.from_numpy(f))
        descs.append((dev[-1].ptr, d.height, d.width, d.bits))
    # The mix's own launch shape, measured on these frames and kept for
    # batches of their size (sf_autotune_mix), beside the shape the big
    # batches' autotuning installed.
    batch = sf.Batch()
    batch.set_timing(True)
    batch.threshold(descs)
    batch.sync()
    before_ms = batch.kernel_ms()[0]
    batch.set_timing(False)
    tuned = sf.autotune_mix(descs)
    batch.threshold(descs, graph=graph)
    batch.label(True)
    counts, status = batch.sync()
    for d, c in zip(descs, counts):
        px = d[1] * d[2]
        byts += px * d[3] / 8 + 16 * c + px / 8
    sf.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.threshold(descs, graph=graph)
        batch.label(True)
        batch.sync()
    secs = (time.perf_counter() - t0) / reps
    batch.set_timing(True)
    batch.threshold(descs)
    batch.sync()
    kms = batch.kernel_ms()[0]
    batch.set_timing(False)
    return {"frames": len(kinds), "mix": "4M:9M:16M:ROI16 = 3:2:4:8",
            "frames_per_s": len(kinds) / secs, "ms_per_batch": 1e3 * secs,
            "fused_ms": kms, "fused_GBps": byts / (kms * 1e-3) / 1e9,
            "fused_ms_under_batch_shape": before_ms, "mix_shape": tuned["best"],
            "refused": int(sum(1 for s in status if s))}


def decompress(sf, frames_host, distinct=4, per_step=32, steps=10):
    """Device decompression of EIGER2 16M uint32 frames stored as bslz4 chunks
    (bitshuffle's default 8192-byte blocks): `distinct` frames encoded on the
    host, `per_step` chunks loaded once (resident in HBM), then each step
    decodes all of them into their own frames.  Also the rate with the chunks
    uploaded every step (PCIe included): two decoders in turn, as the serving
    loop runs them, and one decoder waited for each step; and the oracle's
    decoder on one frame on one host thread."""
    from spotfinder_hip import codec
    f0 = frames_host[0]
    h, w = f0.shape
    log(f"encoding {distinct} 16M frames as bslz4")
    chunks = [codec.bslz4_encode(frames_host[i]) for i in range(distinct)]
    outs = [sf.DeviceArray(f0.nbytes) for _ in range(per_step)]
    load = [(chunks[i % distinct], "bslz4", 32, h, w, outs[i]) for i in range(per_step)]
    dec = sf.Decoder()
    dec.load(load)
    dec.run()
    if dec.sync() != [0] * per_step:
        raise RuntimeError("a chunk failed to decode: " + sf.last_error())
    for i in range(distinct):
        if not np.array_equal(outs[i].to_numpy(np.uint32, (h, w)), frames_host[i]):
            raise RuntimeError("a decoded frame differs from its original")
    t0 = time.perf_counter()
    for _ in range(steps):
        dec.run()
        dec.sync()
    resident = per_step * steps / (time.perf_counter() - t0)
    dec.set_timing(True)
    kms = []
    for _ in range(3):
        dec.run()
        dec.sync()
        kms.append(dec.kernel_ms())
    dec.set_timing(False)
    kms = float(np.median(kms))
    t0 = time.perf_counter()
    for _ in range(3):
        dec.load(load)
        dec.run()
        dec.sync()
    host_serial = per_step * 3 / (time.perf_counter() - t0)
    # The same with two decoders in turn (each into frames of its own), as the
    # serving loop runs them: one's host staging and upload overlap the other's
    # decoding; each is waited for, and its statuses checked, before it loads
    # again.
    dec2 = sf.Decoder()
    outs2 = [sf.DeviceArray(f0.nbytes) for _ in range(per_step)]
    load2 = [ld[:5] + (outs2[i],) for i, ld in enumerate(load)]
    pair = [(dec, load), (dec2, load2)]
    n_runs = 8
    t0 = time.perf_counter()
    for i in range(n_runs):
        d, ld = pair[i % 2]
        if i >= 2 and d.sync() != [0] * per_step:
            raise RuntimeError("a chunk failed to decode: " + sf.last_error())
        d.load(ld)
        d.run()
    for d, _ in pair:
        if d.sync() != [0] * per_step:
            raise RuntimeError("a chunk failed to decode: " + sf.last_error())
    host = per_step * n_runs / (time.perf_counter() - t0)
    if not np.array_equal(outs2[1].to_numpy(np.uint32, (h, w)), frames_host[1]):
        raise RuntimeError("a decoded frame differs from its original")
    dec2.close()
    for o in outs2:
        o.free()
    comp = sum(len(c) for c, *_ in load)
    traffic = traffic_source = None
    try:  # rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled, for this launch size
        with open(os.path.join(ROOT, "profiles", "decode_bslz4_16m_traffic.json")) as fh:
            m = json.load(fh)
        if m.get("chunks_per_launch") == per_step:
            traffic = m["traffic_bytes_per_launch"]
            traffic_source = "profiles/decode_bslz4_16m_traffic.json"
    except (OSError, ValueError):
        pass
    alg = comp + per_step * f0.nbytes  # each chunk read once, each frame written once
    achieved = alg / (kms * 1e-3) / 1e9
    dec.close()
    for o in outs:
        o.free()
    from oracle import oracle as O
    t0 = time.perf_counter()
    rc, out, _ = O.oracle().decompress_image(chunks[0], "bslz4", 32, h, w)
    cpu_s = time.perf_counter() - t0
    assert rc == 0 and out == frames_host[0].tobytes()
    return {
        "workload": f"{per_step} EIGER2 16M uint32 frames as bslz4 chunks "
                    "(bitshuffle default 8192-byte blocks), resident in HBM, "
                    "decoded into frames in HBM",
        "value": resident, "unit": "frames/s",
        "ms_per_step": 1e3 * per_step / resident,
        "compression_ratio": per_step * f0.nbytes / comp,
        "kernel_ms": kms,
        "host_upload_included_frames_per_s": host,
        "host_upload_included_serial_frames_per_s": host_serial,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": "decode_bslz4 (one wave per 8192-byte block)",
                     "bytes_per_frame": alg / per_step},
        "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "frames/s", "cores": 1,
                         "kind": "port",
                         "sample": "one 16M bslz4 chunk through oracle/decompress.c"},
    }


def narrow(sf, frames_host, per_step=32, steps=20):
    """--gpu-force's narrowing of 32-bit frames to 16 bits with mxi_max's
    scan, on `per_step` 16M frames resident in HBM in one launch: HBM bytes 4
    read and 2 written a pixel.  The oracle's narrow_to_16 on one frame on one
    host thread beside it."""
    f0 = frames_host[0]
    dev = [sf.DeviceArray.from_numpy(frames_host[i % len(frames_host)])
           for i in range(per_step)]
    outs = [sf.DeviceArray(2 * f0.size) for _ in range(per_step)]
    jobs = [(d, o, f0.size, 32) for d, o in zip(dev, outs)]
    res = sf.narrow_frames(jobs)
    from oracle import oracle as O
    fits, worst, f16 = O.oracle().narrow_to_16(frames_host[0])
    if res[0]["largest"] != worst or outs[0].to_numpy(np.uint16, f0.shape).tobytes() != f16.tobytes():
        raise RuntimeError("the device narrowing differs from the oracle's")
    # Untimed launches first (about 50 ms of device work): the leg follows the
    # lightly loaded serving replays, and a few launches measured alone can
    # catch the device before its clocks are back up.
    for _ in range(40):
        sf.narrow_frames(jobs)
    ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ms.append(sf.narrow_frames(jobs, timed=True)[1])
    wall = (time.perf_counter() - t0) / steps
    kms = float(np.median(ms))
    alg = 6.0 * f0.size * per_step
    achieved = alg / (kms * 1e-3) / 1e9
    traffic = traffic_source = None
    try:  # rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled, for this launch size
        with open(os.path.join(ROOT, "profiles", "narrow_scan_16m_traffic.json")) as fh:
            m = json.load(fh)
        if m.get("frames_per_launch") == per_step:
            traffic = m["traffic_bytes_per_launch"]
            traffic_source = "profiles/narrow_scan_16m_traffic.json"
    except (OSError, ValueError):
        pass
    t0 = time.perf_counter()
    O.oracle().narrow_to_16(frames_host[0])
    cpu_s = time.perf_counter() - t0
    for x in dev + outs:
        x.free()
    return {
        "workload": f"{per_step} EIGER2 16M uint32 frames resident in HBM, narrowed "
                    "to uint16 with the largest valid count, marker and over-limit "
                    "counts, one launch",
        "value": per_step / (kms * 1e-3), "unit": "frames/s",
        "wall_frames_per_s": per_step / wall, "kernel_ms": kms,
        "fits_16_bits": bool(res[0]["fits"]), "largest_count": res[0]["largest"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source,
                     "kernel": "narrow_scan", "bytes_per_frame": alg / per_step},
        "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "frames/s", "cores": 1,
                         "kind": "port",
                         "sample": "one 16M frame through oracle/narrow.c"},
    }


def pipeline_3d(sf, dev, h, w, batches=4):
    """The whole spot-finding chain in 3D, twice: grouped by the host stitch
    (the pixels copied back per frame), and grouped on the device (the
    stitch, spots and spot pixel list in HBM)."""
    host, host_rec = pipeline_3d_run(sf, dev, h, w, batches, False)
    device, dev_rec = pipeline_3d_run(sf, dev, h, w, batches, True)
    # The two groupings must agree record for record.
    device["same_as_host_stitch"] = (
        host_rec[0].tobytes() == dev_rec[0].tobytes()
        and host_rec[1].tobytes() == dev_rec[1].tobytes()
        and host["counts"] == device["counts"])
    out = {"host_stitch": host, "device_stitch": device}
    out["cpu_baseline"] = grouping_baseline(sf, dev, h, w, batches, device["spots"])
    return out


def grouping_baseline(sf, dev, h, w, batches, spots):
    """The reference's own Labeller (oracle/_ref, compiled from its sources) on
    one host thread, over the same consecutive frames' signal lists as the
    device pipeline groups: the grouping step alone."""
    from oracle import oracle as O
    lib, kind = O.reference(), "reference"
    if lib is None:
        lib, kind = O.oracle(), "port"
    b = sf.Batch()
    b.threshold([(d.ptr, h, w, 32) for d in dev], graph=False)
    counts, _ = b.sync()
    lists = [b.fetch(i, counts[i]) for i in range(len(dev))]
    b.close()
    lab = lib.labeller(h, w)
    t0 = time.perf_counter()
    for step in range(batches):
        for i, signal in enumerate(lists):
            lab.add(step * len(dev) + i, signal)
    lab.finish()
    secs = time.perf_counter() - t0
    n = batches * len(dev)
    return {"value": n / secs, "unit": "frames/s", "cores": 1, "kind": kind,
            "sample": f"the same {n} frames' signal lists through the reference "
                      "Labeller (3D), grouping only",
            "spots": len(lab.spots()), "same_spot_count": len(lab.spots()) == spots}


def pipeline_3d_run(sf, dev, h, w, batches, device_pixels):
    """`batches` steps of the benchmark's 32 resident frames (threshold,
    ordering and labelling on the device), each batch's components stitched
    across frames into spots by the grouping (sf_labeller_add_batch), as
    consecutive frames of one sweep."""
    descs = [(d.ptr, h, w, 32) for d in dev]
    b = sf.Batch()
    # Untimed: batches alternate between two slot ranges, so two batches
    # allocate both and capture their graphs.
    for _ in range(2):
        b.threshold(descs, graph=True)
        b.label(True)
        b.sync()
    lab = sf.Labeller(h, w, device_pixels=device_pixels)
    t_dev = t_host = 0.0
    for step in range(batches):
        t0 = time.perf_counter()
        b.threshold(descs, graph=True)
        b.label(True)
        b.sync()
        t1 = time.perf_counter()
        lab.add_batch(step * len(dev), b)
        t_dev += t1 - t0
        t_host += time.perf_counter() - t1
    t0 = time.perf_counter()
    lab.finish()
    # The spots and their pixels in the library's host arrays (the device
    # grouping copies each batch's back while the next one runs; reading
    # them waits for the copies): part of the timed chain.
    lab.spots(copy=False), lab.pixels(copy=False)
    t_host += time.perf_counter() - t0
    spots, pixels = lab.spots(), lab.pixels()
    n = batches * len(dev)
    b.close()
    return {"workload": f"{n} consecutive 16M frames (the benchmark's 32, cycled) "
                        "through threshold, labelling and 3D grouping into spots, "
                        "spots and pixels copied to host memory",
            "value": n / (t_dev + t_host), "unit": "frames/s",
            "device_ms_per_frame": 1e3 * t_dev / n,
            "grouping_ms_per_frame": 1e3 * t_host / n,
            "spots": len(spots), "pixels": len(pixels),
            "counts": lab.counts()}, (spots, pixels)


def refl(sf, repeat=64):
    """The spots as a DIALS reflection table with shoeboxes: the reference
    Labeller's 2D spots of the golden sweep repeated `repeat` times, encoded by
    this library and by the reference's own refl::write (oracle/_ref), on one
    host thread each."""
    z = np.load(os.path.join(ROOT, "tests", "golden", "labeller_sweep.npz"))
    spots = np.resize(z["spots_2d"], len(z["spots_2d"]) * repeat)
    pixels, width = z["pixels_2d"], int(z["desc_width"])
    secs = float("inf")
    for _ in range(3):  # the best of three, for both writers
        t0 = time.perf_counter()
        ours = sf.refl_encode(spots, pixels, width)
        secs = min(secs, time.perf_counter() - t0)
    out = {"workload": f"{len(spots)} spots with shoeboxes (the golden sweep's "
                       f"2D spots, repeated), encoded in memory",
           "value": len(spots) / secs, "unit": "spots/s", "bytes": len(ours),
           "MB_per_s": len(ours) / secs / 1e6}
    from oracle import oracle as O
    ref = O.reference()
    if ref is not None:
        rsecs = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            theirs = ref.refl_bytes(spots, pixels, width)
            rsecs = min(rsecs, time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": len(spots) / rsecs, "unit": "spots/s",
                               "cores": 1, "kind": "reference",
                               "sample": "the same spots through refl::write to a file",
                               "identical": theirs == ours}
    return out


def copy_calibration(sf, nbytes=1 << 30, copies=4):
    """Plain device-to-device copies: each reads and writes `nbytes`, so the
    HBM read and write counters of these dispatches give the factor from
    counter units to bytes."""
    src, dst = sf.DeviceArray(nbytes), sf.DeviceArray(nbytes)
    src.copy_from(np.arange(nbytes // 4, dtype=np.uint32))
    sf.synchronize()
    t0 = time.perf_counter()
    for _ in range(copies):
        dst.copy_device(src)
    sf.synchronize()
    secs = time.perf_counter() - t0
    print(json.dumps({"copy_bytes": nbytes, "copies": copies,
                      "copy_GBps": 2 * nbytes * copies / secs / 1e9}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=32,
                    help="frames per batch (one step), resident in HBM")
    ap.add_argument("--in-flight", type=int, default=2,
                    help="batch handles queued at once in the timed steps "
                         "(each on its own stream)")
    ap.add_argument("--graph", type=int, default=1,
                    help="replay the batch's threshold from a captured hipGraph")
    ap.add_argument("--no-label", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--autotune", type=int, default=1,
                    help="measure the kernel shape on the benchmark's frames "
                         "first (sf_autotune) and run with the best")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-decompress", action="store_true")
    ap.add_argument("--no-narrow", action="store_true")
    ap.add_argument("--no-refl", action="store_true")
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only device-to-device copies of 1 GiB, the known "
                         "traffic against which HBM counters are scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path "
              f"(asked for --gpus {args.gpus}). Frames would shard as "
              f"contiguous frame ranges, one per GPU, with the 3D grouping "
              f"joined across each range boundary", file=sys.stderr)
        sys.exit(2)

    import spotfinder_hip as sf
    if not sf.available():
        raise SystemExit("no HIP device: " + sf.unavailable_reason())
    if args.copy_calibration:
        copy_calibration(sf)
        return

    desc = sf.synth_desc("16m", spots_per_frame=2000, persist=1, seed=2024)
    h, w = desc.height, desc.width
    log(f"generating {args.frames} frames of {w} x {h} uint32")
    frames_host = [sf.synth_frame(desc, z) for z in range(args.frames)]
    dev = [sf.DeviceArray.from_numpy(f) for f in frames_host]
    descs = [(d.ptr, h, w, 32) for d in dev]
    sf.synchronize()
    label = not args.no_label

    tuned = None
    if args.autotune:
        tuned = sf.autotune(descs, label=not args.no_label)
        log(f"autotune: {tuned}")

    # Batch handles in flight (two by default): each has its own stream and
    # slots, so with two the next step is queued before the last one is
    # waited for; every step's results are still waited for and checked.
    batches = [sf.Batch() for _ in range(max(1, args.in_flight))]
    batch = batches[0]

    def submit(b):
        b.threshold(descs, graph=bool(args.graph))
        if label:
            b.label(True)

    def finish(b):
        c, status = b.sync()
        if any(status):
            raise RuntimeError(f"a frame was refused: {status}")
        return c

    def run(n):
        c = []
        k = len(batches)
        for i in range(n):
            if i >= k:
                c = finish(batches[i % k])
            submit(batches[i % k])
        for i in range(max(0, n - k), n):
            c = finish(batches[i % k])
        return c

    counts = run(args.warmup)
    sf.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    sf.synchronize()
    elapsed = time.perf_counter() - t0
    frames_done = args.steps * len(dev)

    # The fused kernel's duration, from HIP events on the batch's stream.
    batch.set_timing(True)
    kms = []
    for _ in range(3):
        batch.threshold(descs)
        batch.sync()
        kms.append(batch.kernel_ms()[0])
    batch.set_timing(False)
    kernel_ms = float(np.median(kms))
    # Frames re-run on the exact direct kernels because a valid pixel was too
    # large for the fused kernel's fast sums; on this workload, none.
    fallbacks = sum(b.fallbacks() for b in batches)
    for b in batches:  # their slots are not needed by the sections below
        b.close()
    mean_signal = float(np.mean(counts))
    n_px = h * w
    bytes_per_frame = n_px * 4 + 16 * mean_signal + n_px / 8.0
    achieved = len(dev) * bytes_per_frame / (kernel_ms * 1e-3) / 1e9
    traffic, traffic_source = measured_traffic(len(dev), sf.get_tuning())

    line = {
        "metric": METRIC,
        "value": frames_done / elapsed,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (EIGER2-XE 16M generator, seed 2024, resident in HBM)",
        "config": {
            "workload": "EIGER2-XE 16M uint32 frames (4148 x 4362), batched: "
                        "one persistent fused threshold launch per batch, "
                        "ordered SignalPixel lists, strong-pixel bitmasks"
                        + (", 2D labelling with exact per-component moments"
                           if label else ""),
            "frame": [h, w], "pixel": "uint32", "frames_per_step": len(dev),
            "graph": bool(args.graph), "batches_in_flight": len(batches),
            "tuning": sf.get_tuning(),
            "autotune": tuned,
            "mean_signal_pixels": mean_signal},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": traffic_source,
            "kernel": "fused_kernel (one launch per batch)",
            "ms_per_launch": kernel_ms,
            "bytes_per_frame": bytes_per_frame,
            "fallback_frames": fallbacks},
    }

    if not args.no_plain:
        # One frame per launch sequence, waiting for each: gpu::find's shape
        # with the frame already in HBM.
        ws = sf.Workspace()
        for d in dev[:2]:  # captures the threshold and labelling graphs
            ws.threshold(d.ptr, 32, h, w, "fused")
            if label:
                sf.workspace_label_enqueue(ws)
            ws.sync()
        t0 = time.perf_counter()
        for d in dev:
            ws.threshold(d.ptr, 32, h, w, "fused")
            if label:
                sf.workspace_label_enqueue(ws)
            rc, _ = ws.sync()
        plain = len(dev) / (time.perf_counter() - t0)
        ws.close()
        line["plain_per_frame"] = {"value": plain, "unit": "frames/s",
                                   "labelled": label}
        line["plain_per_frame_h2d"] = plain_host_fed(sf, frames_host)
        log(f"plain per frame, host-fed: {line['plain_per_frame_h2d']}")

    if not args.no_uneven:
        line["uneven_batch"] = uneven_batch(sf, reps=5, graph=bool(args.graph))
        log(f"uneven batch: {line['uneven_batch']}")
    if not args.no_serve:
        line["serving"] = serving(sf, args.serve_seconds, seed=2025)

    if not args.no_decompress:
        line["decompress"] = decompress(sf, frames_host)
        log(f"decompress: {line['decompress']}")
    if not args.no_narrow:
        line["narrow"] = narrow(sf, frames_host)
        log(f"narrow: {line['narrow']}")
    if not args.no_pipeline:
        line["pipeline_3d"] = pipeline_3d(sf, dev, h, w)
        log(f"pipeline_3d: {line['pipeline_3d']}")
    if not args.no_refl:
        line["refl"] = refl(sf)
        log(f"refl: {line['refl']}")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(frames_host)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
