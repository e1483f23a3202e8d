"""--gpu-force's narrowing (find_spots.cc:162-182) and mxi_max's scan
(apps/mxi_max.cc:45-68): the oracle, pinned without a GPU.

- The reference's own test of them (python/tests/test_max.py:41-76), restated
  on planted frames: a 16-bit series and the same counts in 32 bits give the
  same largest count; a dead pixel on 12 frames and a 256-pixel tile join on 12
  frames count 12 and 3072, neither as a count; a count of 70000 is refused,
  named; and the narrowed frames threshold to the 32-bit threshold's spots,
  byte for byte.
- An independent vectorised formulation over random frames with every marker
  and the values at the 16-bit limit.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


def planted(bits, frames=12, hot=None, seed=5):
    rng = np.random.default_rng(seed)
    dt = np.uint16 if bits == 16 else np.uint32
    top = np.iinfo(dt).max
    out = []
    for z in range(frames):
        f = rng.poisson(3.0, size=(64, 300)).astype(dt)
        f[10:20, 40:44] += np.uint32(1200).astype(dt)  # a spot
        f[:, 100:101] = 0
        f[0:64, 100] = top          # a tile join's column: 64 pixels...
        f[0:64, 200] = top          # ...four of them, 256 pixels a frame
        f[0:64, 201] = top
        f[0:64, 202] = top
        f[5, 7] = top - 1           # one dead pixel a frame
        if hot is not None and z == hot[0]:
            f[30, 30] = hot[1]
        out.append(f)
    return out


def numpy_scan(f):
    top = np.iinfo(f.dtype).max
    real = np.where(f >= top - 1, 0, f.astype(np.uint64))
    return {"largest": int(real.max()), "bad": int((f == top - 1).sum()),
            "joins": int((f == top).sum()), "over": int((real >= 0xFFFD).sum())}


def test_the_reference_tests_planted_series(orc):
    s16 = planted(16)
    s32 = planted(32)
    t16 = [orc.scan(f) for f in s16]
    t32 = [orc.scan(f) for f in s32]
    assert max(t["largest"] for t in t16) == max(t["largest"] for t in t32)
    assert sum(t["bad"] for t in t32) == 12
    assert sum(t["joins"] for t in t32) == 3072
    hot = planted(32, hot=(7, 70000))
    results = [orc.narrow_to_16(f) for f in hot]
    assert [z for z, r in enumerate(results) if not r[0]] == [7]
    assert results[7][1] == 70000


def test_narrowed_frames_threshold_as_the_32_bit_frames(orc):
    """find_spots.cc:783-795: with --gpu-force the 16-bit threshold of the
    narrowed frame gives the 32-bit threshold's spots, byte for byte."""
    for f in planted(32, frames=3):
        fits, worst, f16 = orc.narrow_to_16(f)
        assert fits and worst < 0xFFFD
        assert np.array_equal(f16, (f & 0xFFFF).astype(np.uint16))
        rc32, a = orc.dext(f)
        rc16, b = orc.dext(f16)
        assert rc32 == rc16 == 0 and len(a) > 10 and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("bits", [16, 32])
def test_scan_against_numpy(orc, bits):
    rng = np.random.default_rng(bits)
    dt = np.uint16 if bits == 16 else np.uint32
    top = np.iinfo(dt).max
    for trial in range(20):
        f = rng.integers(0, 1 << min(bits, 20), size=int(rng.integers(1, 5000)),
                         dtype=np.uint64).astype(dt)
        edges = np.array([0xFFFC, 0xFFFD, 0xFFFE, 0xFFFF, top - 2, top - 1, top],
                         np.uint64).astype(dt)
        k = int(rng.integers(0, 6))
        f[rng.integers(0, f.size, k)] = rng.choice(edges, k)
        assert orc.scan(f) == numpy_scan(f), (bits, trial)
        if bits == 32:
            fits, worst, f16 = orc.narrow_to_16(f)
            want = numpy_scan(f)["largest"]
            assert worst == want and fits == (want < 0xFFFD)
            assert np.array_equal(f16, (f & 0xFFFF).astype(np.uint16))


def test_the_16_bit_limit(orc):
    base = np.zeros(16, np.uint32)
    for v, fits in [(0xFFFC, True), (0xFFFD, False), (0xFFFE, False),
                    (0xFFFFFFFD, False), (0xFFFFFFFE, True), (0xFFFFFFFF, True)]:
        f = base.copy()
        f[3] = v
        ok, worst, f16 = orc.narrow_to_16(f)
        assert ok == fits, hex(v)
        assert f16[3] == (v & 0xFFFF)
