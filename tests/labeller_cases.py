"""The grouping's hand-derived cases (test infrastructure).

Restated from the reference's tests/spots/test_dials_spots.cc:71-412: each
case adds frames to a labeller and states what must come out, with the
expected centroids and variances worked by hand there.  `run_cases(make)`
runs every case against labellers made by make(height, width, **options),
so the same cases check the oracle, the reference build, the host-side
grouping and the device labeller."""
import numpy as np

SIG = np.dtype([("index", "<u4"), ("value", "<u4"), ("background", "<f4"),
                ("population", "<u2"), ("reserved", "<u2")])
H, W = 8, 10


def px(pairs):
    a = np.zeros(len(pairs), SIG)
    for k, (i, v) in enumerate(pairs):
        a[k]["index"], a[k]["value"] = i, v
    return a


def flat(indices, value):
    return px([(i, value) for i in indices])


def close(a, b):
    assert abs(a - b) < 1e-9, (a, b)


def case_three_in_a_row(make):
    lab = make(H, W, max_separation=0.0)
    lab.add(0, px([(0, 1), (1, 2), (2, 1)]))
    lab.finish()
    s = lab.spots()
    assert len(s) == 1
    s = s[0]
    assert s["n_signal"] == 3
    close(s["intensity"], 4.0)
    close(s["intensity_variance"], 4.0)
    close(s["position"][0], 1.5)
    close(s["position"][1], 0.5)
    close(s["position"][2], 0.5)
    close(s["variance"][0], 0.8 / 4.0 + 1.0 / 12.0)
    close(s["variance"][1], 1.0 / 12.0)
    close(s["variance"][2], 1.0 / 12.0)
    assert list(s["bbox"]) == [0, 3, 0, 1, 0, 1]
    assert len(lab.pixels()) == 3
    c = lab.counts()
    assert c["groups"] == 1 and c["too_small"] == 0


def case_streak_3d(make):
    lab = make(H, W, max_separation=0.0)
    for z in range(3):
        lab.add(z, flat([0, 1, 2], 1))
    lab.finish()
    s = lab.spots()
    assert len(s) == 1
    s = s[0]
    assert s["n_signal"] == 9
    close(s["intensity"], 9.0)
    close(s["position"][0], 1.5)
    close(s["position"][2], 1.5)
    close(s["variance"][0], 0.75 / 9.0 + 1.0 / 12.0)
    close(s["variance"][2], 0.75 / 9.0 + 1.0 / 12.0)
    assert s["bbox"][4] == 0 and s["bbox"][5] == 3


def case_streak_2d(make):
    lab = make(H, W, max_separation=0.0, two_d=True)
    for z in range(3):
        lab.add(z, flat([0, 1, 2], 1))
    lab.finish()
    s = lab.spots()
    assert len(s) == 3
    close(s[1]["position"][2], 1.5)
    assert s[1]["bbox"][5] - s[1]["bbox"][4] == 1


def case_gap_frame(make):
    lab = make(H, W, max_separation=0.0)
    for z in (0, 1, 3, 4):
        lab.add(z, flat([0, 1, 2], 1))
    lab.finish()
    s = lab.spots()
    assert len(s) == 2
    assert s[0]["n_signal"] == 6 and s[1]["n_signal"] == 6
    assert s[1]["bbox"][4] == 3


def case_connectivity(make):
    def count(frames):
        lab = make(H, W, max_separation=0.0, min_spot_size=1)
        for z, idx in frames:
            lab.add(z, flat(idx, 1))
        lab.finish()
        return len(lab.spots())
    assert count([(0, [9, 10])]) == 2      # a row wrap is not an edge
    assert count([(0, [0, 10])]) == 1      # the pixel above is
    assert count([(0, [0, 11])]) == 2      # a diagonal is not
    assert count([(0, [0]), (1, [10])]) == 2  # nor a diagonal in z


def case_streaming(make):
    lab = make(H, W, max_separation=0.0)
    lab.add(0, flat([0, 1, 2], 1))
    assert len(lab.spots()) == 0 and lab.live() == 3
    lab.add(1, flat([40, 41, 42], 1))
    assert len(lab.spots()) == 1 and lab.live() == 3
    lab.finish()
    assert len(lab.spots()) == 2 and lab.live() == 0


def case_size_filters(make):
    lab = make(H, W, max_separation=0.0)
    lab.add(0, flat([0, 1], 1))
    lab.add(1, flat([40, 41, 42], 1))
    lab.finish()
    c = lab.counts()
    assert len(lab.spots()) == 1
    assert c["too_small"] == 1 and c["groups"] == 2 and c["signal_pixels"] == 5
    lab = make(H, W, max_separation=0.0, max_spot_size=5)
    lab.add(0, flat([0, 1, 2, 3, 4, 5], 1))
    lab.finish()
    assert len(lab.spots()) == 0 and lab.counts()["too_large"] == 1


def case_separation(make):
    lab = make(H, W, max_separation=2.0)
    lab.add(0, px([(0, 1000)] + [(x, 1) for x in range(1, 10)]))
    lab.finish()
    assert len(lab.spots()) == 1 and lab.counts()["separated"] == 0
    lab = make(H, W, max_separation=2.0)
    lab.add(0, px([(0, 100), (1, 100), (2, 1), (3, 1), (4, 1), (5, 1),
                   (6, 90), (7, 90)]))
    lab.finish()
    c = lab.counts()
    assert len(lab.spots()) == 0 and c["separated"] == 1 and c["groups"] == 1


def case_single_pixel(make):
    lab = make(H, W, max_separation=0.0, min_spot_size=1)
    lab.add(0, flat([35], 700))
    lab.finish()
    s = lab.spots()
    assert len(s) == 1
    close(s[0]["position"][0], 5.5)
    close(s[0]["position"][1], 3.5)
    close(s[0]["variance"][0], 1.0 / 12.0)
    close(s[0]["variance"][1], 1.0 / 12.0)
    close(s[0]["intensity"], 700.0)


def case_order_errors(make, error):
    lab = make(H, W, max_separation=0.0)
    lab.add(5, flat([0, 1, 2], 1))
    for frame, idx in ((4, [0]), (5, [0])):
        try:
            lab.add(frame, flat(idx, 1))
            raise AssertionError("a frame out of order was accepted")
        except error:
            pass
    for bad in ([2, 1], [H * W]):
        lab = make(H, W, max_separation=0.0)
        try:
            lab.add(0, flat(bad, 1))
            raise AssertionError("a bad pixel list was accepted")
        except error:
            pass


def case_long_streak(make):
    lab = make(H, W, max_separation=0.0)
    frames = 200
    for z in range(frames):
        lab.add(z, flat([44, 45, 46, 70, 71, 72], 3))
    lab.finish()
    s = lab.spots()
    assert len(s) == 2
    assert s[0]["n_signal"] == 3 * frames
    assert s[0]["bbox"][5] - s[0]["bbox"][4] == frames
    assert len(lab.pixels()) == 6 * frames


CASES = [case_three_in_a_row, case_streak_3d, case_streak_2d, case_gap_frame,
         case_connectivity, case_streaming, case_size_filters, case_separation,
         case_single_pixel, case_long_streak]


def run_cases(make, error):
    for case in CASES:
        case(make)
    case_order_errors(make, error)


def compare_spots(got, want, tol=1e-9):
    """Field by field: exact except the variance, within `tol`."""
    assert len(got) == len(want), (len(got), len(want))
    for f in ("bbox", "n_signal", "first", "intensity", "intensity_variance",
              "position"):
        bad = np.nonzero(~np.all((got[f] == want[f]).reshape(len(got), -1), axis=1))[0]
        assert len(bad) == 0, (f, bad[:5], got[bad[:3]], want[bad[:3]])
    assert np.abs(got["variance"] - want["variance"]).max(initial=0) <= tol
