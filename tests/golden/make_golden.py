"""Writes the golden vectors under tests/golden/ from the reference's own CPU
code (oracle/_ref/libspots_ref.so, compiled from the reference's
src/spots/dext.cc, signal_order.cc and dials_spots.cc by `make -C oracle ref`).

Run here, where the reference sources exist:  python tests/golden/make_golden.py
The outputs are data only: input frames (or their SHA-256 where the frames
are regenerated from the synthetic generator), and the reference's outputs.

  threshold_matrix.npz   48 frames in the size/edge matrix of the reference's
                         tests/spots/test_dext_fused.cc (8 sizes x 3 means x
                         u16/u32, masked and saturated pixels, a masked row and
                         column, a zero patch, spots on edges and corners),
                         with dext<T>'s SignalPixel list for each.
  threshold_synth.json   EIGER2-shaped synthetic frames (4M, 16M, 9M, 512^2
                         u16): frame SHA-256, signal count, SHA-256 of the list.
  labeller_sweep.npz     a 10-frame 384x320 sweep with persistent spots: the
                         reference Labeller's spots, pixels and counts in 3D
                         and 2D, with the thresholded input lists.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dials-metal_amd"))

from oracle import oracle as O  # noqa: E402
import spotfinder_hip as sf  # noqa: E402

MATRIX_SIZES = [(1, 1), (5, 7), (32, 32), (33, 65), (100, 37), (67, 300),
                (257, 511), (64, 96)]
MATRIX_MEANS = [0.3, 2.0, 12.0]


def matrix_frame(h, w, seed, mean, dtype):
    """A frame with the features of the reference's fused-tile test matrix."""
    rng = np.random.default_rng(seed)
    top = np.iinfo(dtype).max
    f = rng.poisson(mean, (h, w)).astype(np.int64)
    spots = h * w // 400 + 3
    for s in range(spots):
        ci = 0.0 if s == 0 else (h - 1.0 if s == 1 else rng.uniform(0, h))
        cj = 0.0 if s == 0 else (w - 1.0 if s == 1 else rng.uniform(0, w))
        peak = 20.0 + 500.0 * rng.uniform()
        for di in range(-4, 5):
            for dj in range(-4, 5):
                i, j = int(ci) + di, int(cj) + dj
                if 0 <= i < h and 0 <= j < w:
                    f[i, j] += int(peak * np.exp(-(di * di + dj * dj) / 3.0))
    f = np.minimum(f, top - 2)
    x = rng.uniform(size=(h, w))
    f[x < 0.005] = top - 1
    f[(x >= 0.005) & (x < 0.007)] = top
    if h > 20:
        f[h // 2, :] = top - 1
    if w > 20:
        f[:, w // 3] = top - 1
    f[:6, :9] = 0
    return f.astype(dtype)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


SYNTH_CASES = [
    # name, kind, overrides, frame number
    ("eiger4m_f0", "4m", {}, 0),
    ("eiger4m_f7_seed3", "4m", {"seed": 3}, 7),
    ("eiger4m_dense", "4m", {"spots_per_frame": 5000, "background": 5.0, "seed": 11}, 2),
    ("eiger4m_dark", "4m", {"spots_per_frame": 100, "background": 0.2,
                            "ring_mean": 0.0, "seed": 12}, 1),
    ("eiger16m_f0", "16m", {"spots_per_frame": 2000, "persist": 1}, 0),
    ("eiger9m_f0", "9m", {"seed": 5, "background": 2.0, "spots_per_frame": 1000}, 0),
    ("roi512_u16", (512, 512), {"bits": 16, "seed": 9, "spots_per_frame": 60,
                                "ring_radius": 100.0, "background": 1.5}, 0),
    ("wide_values_u32", (640, 480), {"seed": 21, "peak_min": 200000.0,
                                     "peak_max": 3000000.0, "spots_per_frame": 40}, 0),
]


def main():
    ref = O.reference()
    if ref is None:
        raise SystemExit("build the reference first: make -C oracle ref")

    frames, outputs, keys = {}, {}, []
    seed = 1
    for (h, w) in MATRIX_SIZES:
        for mean in MATRIX_MEANS:
            for dt in (np.uint16, np.uint32):
                name = f"m{h}x{w}_{mean}_{np.dtype(dt).itemsize * 8}"
                f = matrix_frame(h, w, seed, mean, dt)
                rc, out = ref.dext(f)
                assert rc == 0
                frames[name] = f
                outputs[name] = out
                keys.append(name)
            seed += 1
    np.savez_compressed(
        os.path.join(HERE, "threshold_matrix.npz"),
        names=np.array(keys),
        **{f"frame_{k}": v for k, v in frames.items()},
        **{f"signal_{k}": v for k, v in outputs.items()})

    synth = {}
    for name, kind, over, fno in SYNTH_CASES:
        d = sf.synth_desc(kind, **over)
        f = sf.synth_frame(d, fno)
        rc, out = ref.dext(f)
        synth[name] = {
            "kind": kind if isinstance(kind, str) else list(kind),
            "overrides": over, "frame": fno, "status": rc,
            "frame_sha256": sha(f), "count": int(len(out)),
            "signal_sha256": sha(out),
            "max_valid": int(f[f < np.iinfo(f.dtype).max - 1].max()),
        }
        print(name, synth[name]["count"])
    with open(os.path.join(HERE, "threshold_synth.json"), "w") as fh:
        json.dump(synth, fh, indent=1, sort_keys=True)

    # A short sweep with persistent spots, grouped by the reference Labeller.
    d = sf.synth_desc((384, 320), seed=4, spots_per_frame=30, persist=1,
                      ring_radius=120.0, peak_min=30.0, peak_max=3000.0)
    lists = []
    for z in range(10):
        f = sf.synth_frame(d, z)
        rc, out = ref.dext(f)
        assert rc == 0
        lists.append(out)
    data = {"desc_width": 320, "desc_height": 384}
    for z, out in enumerate(lists):
        data[f"signal_{z}"] = out
    for mode, two_d in (("3d", False), ("2d", True)):
        lab = ref.labeller(384, 320, two_d=two_d)
        for z, out in enumerate(lists):
            lab.add(z, out)
        lab.finish()
        data[f"spots_{mode}"] = lab.spots()
        data[f"pixels_{mode}"] = lab.pixels()
        c = lab.counts()
        data[f"counts_{mode}"] = np.array([c[k] for k in O.COUNTS_FIELDS],
                                          np.uint64)
        print(mode, len(data[f"spots_{mode}"]), c)
    np.savez_compressed(os.path.join(HERE, "labeller_sweep.npz"), **data)


if __name__ == "__main__":
    main()
