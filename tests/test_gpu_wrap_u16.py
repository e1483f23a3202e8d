"""16-bit frames whose 7x7 sum of squares wraps its 32-bit accumulator.

For uint16 frames the reference sums the window in uint32 (accumulator_t,
src/spots/dext.hh:18-38) and evaluates the test in float on the wrapped value;
tests/spots/test_dext_squares.cc:1-37 writes down where that wrap begins (one
pixel at 65535, two at 46341, a uniform window at 9360 counts).  These frames
go well past it: plateaus of 10 000-65 533 counts, single pixels at the top
of the valid range, and masked sentinels (0xFFFE, 0xFFFF) beside and inside
the bright areas, so that the count, the sum and the wrapped sum of squares
all change where the masked pixels leave the window.  The test first checks,
in exact integers, that the frames really reach the wrap, then compares both
device kernels (fused and direct, one frame at a time and in a batch) with
the oracle's dext<uint16_t>, every field of every SignalPixel.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def window_square_sums(frame):
    """Exact 7x7 sums of squares over valid pixels (clamped windows), as
    Python integers in an int64 array: the value the uint32 accumulator
    wraps."""
    v = frame.astype(np.int64)
    v[frame >= 0xFFFE] = 0
    sq = np.pad(v * v, 3)
    c = np.cumsum(np.cumsum(sq, 0), 1)
    c = np.pad(c, ((1, 0), (1, 0)))
    h, w = frame.shape
    return c[7:7 + h, 7:7 + w] - c[:h, 7:7 + w] - c[7:7 + h, :w] + c[:h, :w]


def bright_u16(h, w, seed):
    rng = np.random.default_rng(seed)
    f = rng.poisson(3.0, (h, w)).astype(np.int64)
    # Plateaus: every window inside one of 10 000 counts a pixel or more is
    # past 2^32 (49 * 10 000^2 = 4.9e9).
    for _ in range(max(4, h * w // 6000)):
        y, x = rng.integers(0, h), rng.integers(0, w)
        dy, dx = rng.integers(3, 14), rng.integers(3, 14)
        f[y:y + dy, x:x + dx] = rng.integers(10000, 65534, (min(dy, h - y), min(dx, w - x)))
    # Lone pixels at the top of the valid range, and pairs at 46 341.
    k = rng.uniform(size=(h, w)) < 0.002
    f[k] = rng.choice([46341, 65533, 65532, 40000], k.sum())
    # Sentinels next to and inside the bright areas.
    for sentinel in (0xFFFE, 0xFFFF):
        m = rng.uniform(size=(h, w)) < 0.01
        f[m] = sentinel
    # Masked module-gap-like columns and rows.
    f[:, w // 3] = 0xFFFF
    f[h // 2, :] = 0xFFFE
    return f.astype(np.uint16)


SHAPES = [(120, 160), (300, 465), (512, 512), (97, 1031)]


@pytest.fixture(scope="module")
def frames():
    out = [bright_u16(h, w, 1000 + i) for i, (h, w) in enumerate(SHAPES)]
    for f in out:
        s2 = window_square_sums(f)
        # The domain this test is for: many windows past the wrap, some by
        # several multiples of 2^32.
        assert (s2 >= (1 << 32)).sum() > 100, f.shape
        assert s2.max() > 20 * (1 << 32), f.shape
    return out


@pytest.mark.parametrize("kernel", ["fused", "direct"])
def test_wrapped_square_sums_one_frame_at_a_time(device, oracle_lib, frames, kernel):
    device.set_kernel(kernel)
    try:
        for f in frames:
            say("u16 wrap", kernel, f.shape)
            rc, got = device.find(f)
            _, want = oracle_lib.dext(f)
            assert rc == 0
            assert len(want) > 20, f.shape
            assert len(got) == len(want), (f.shape, len(got), len(want))
            assert got.tobytes() == want.tobytes(), f.shape
    finally:
        device.set_kernel("fused")


@pytest.mark.parametrize("graph", [False, True])
def test_wrapped_square_sums_in_a_batch(device, oracle_lib, frames, graph):
    """The same frames in one launch beside a 4M uint32 frame, so that the
    16-bit and 32-bit paths share the launch."""
    big = device.synth_frame(device.synth_desc("4m", seed=5), 0)
    host = frames + [big] + frames[:2]
    dev = [device.DeviceArray.from_numpy(f) for f in host]
    batch = device.Batch()
    for rep in range(2 if graph else 1):
        say("u16 wrap batch, graph", graph, "replay", rep)
        batch.threshold([(d.ptr, f.shape[0], f.shape[1], 16 if f.dtype == np.uint16 else 32)
                         for d, f in zip(dev, host)], graph=graph)
        counts, status = batch.sync()
        for i, f in enumerate(host):
            assert status[i] == 0, i
            _, want = oracle_lib.dext(f)
            assert batch.fetch(i, counts[i]).tobytes() == want.tobytes(), (i, f.shape)
    batch.close()
