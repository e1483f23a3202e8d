"""Benchmark of the spot-finding hot path on one MI355X.

Workload (EIGER2-XE 16M sweep): 4148 x 4362 uint32 synthetic frames, Poisson
background 0.8 with a solvent ring, 2000 spots per frame persisting 1-3
frames, 8 x 4 module gaps masked, 0.01 % hot pixels; the dispersion threshold
(7x7 / erode 5x5 / 11x11), ordered SignalPixel list and strong-pixel bitmask,
one frame per launch sequence ("plain per-frame" mode).  Frames are generated
on the host, uploaded once, and are resident in HBM when the timed region
starts.  A step is one pass over the batch of frames.

Prints one JSON line:  frames/s (value), the fused kernel's achieved HBM
bandwidth against the 8 TB/s peak (roofline), and the reference's CPU path on
this host's cores over a bounded sample of the same frames (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dials-metal_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "16M frames/s + achieved HBM GB/s (bit-exact mask); serving p50/p99 frame latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(frames_host, h, w):
    """The reference's dext<uint32_t> (oracle/_ref, compiled from the
    reference's own sources) on this host's cores, a bounded sample."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    lib, kind = O.reference(), "reference"
    if lib is None:
        lib, kind = O.oracle(), "port"
    n = min(len(frames_host), max(threads, 8))
    sample = np.stack(frames_host[:n])
    signal, secs = lib.threads_u32(sample, threads)
    return {"value": n / secs, "unit": "frames/s", "cores": threads, "kind": kind,
            "sample": f"{n} of the benchmark's 16M frames, dext<uint32_t> on "
                      f"{threads} threads, one frame per thread at a time "
                      f"({secs:.2f} s wall, {signal} signal pixels)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=16,
                    help="distinct frames resident in HBM, one step's batch")
    ap.add_argument("--kernel", default="fused", choices=["fused", "direct"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}); see DESIGN.md for how frames "
              f"would shard", file=sys.stderr)
        sys.exit(2)

    import spotfinder_hip as sf
    if not sf.available():
        raise SystemExit("no HIP device: " + sf.unavailable_reason())

    desc = sf.synth_desc("16m", spots_per_frame=2000, persist=1, seed=2024)
    h, w = desc.height, desc.width
    log(f"generating {args.frames} frames of {w} x {h} uint32")
    frames_host = [sf.synth_frame(desc, z) for z in range(args.frames)]
    dev = [sf.DeviceArray.from_numpy(f) for f in frames_host]
    sf.synchronize()

    ws = sf.Workspace()
    ws.set_timing(True)
    counts = []

    def step(record):
        kms = []
        for i, d in enumerate(dev):
            ws.threshold(d.ptr, 32, h, w, args.kernel)
            rc, n = ws.sync()
            if rc != 0:
                raise RuntimeError(f"frame {i}: status {rc}")
            if record:
                kms.append(ws.kernel_ms()[0])
                counts.append(n)
        return kms

    for _ in range(args.warmup):
        step(False)
    ws.set_timing(False)
    sf.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    sf.synchronize()
    elapsed = time.perf_counter() - t0

    # Kernel durations, measured live with HIP events on the kernel's stream.
    ws.set_timing(True)
    kms = step(True)
    kernel_ms = float(np.mean(kms))
    mean_signal = float(np.mean(counts))
    n_px = h * w
    algo_bytes = n_px * 4 + 16 * mean_signal + n_px / 8.0
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9

    frames_done = args.steps * len(dev)
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
        "data": "synthetic",
        "config": {"workload": "EIGER2-XE 16M sweep frames, plain per-frame "
                               "threshold (one launch sequence per frame)",
                   "frame": [h, w], "pixel": "uint32",
                   "frames_per_step": len(dev), "kernel": args.kernel,
                   "mean_signal_pixels": mean_signal},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None,
                     "kernel": f"fused_kernel<uint32_t>, {kernel_ms:.4f} ms/frame"},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(frames_host, h, w)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
