"""The grouping, without a device.

The oracle's Labeller and the reference's own (oracle/_ref) against the
reference test suite's hand-derived cases and the golden sweep; then the
library's host-side grouping (sf_labeller_add_labelled), fed the device's
contract by a plain-Python model of the per-frame labelling, against the same
cases and the reference's golden output."""
import numpy as np
import pytest

from labeller_cases import compare_spots, run_cases
from labelling_model import ModelFeeder
from oracle import oracle as O


def oracle_make(h, w, **kw):
    return O.oracle().labeller(h, w, **kw)


def test_oracle_labeller_cases():
    run_cases(oracle_make, O.LabellerError)


def test_reference_labeller_cases():
    ref = O.reference()
    if ref is None:
        pytest.skip("oracle/_ref is not built (no reference sources here)")
    run_cases(lambda h, w, **kw: ref.labeller(h, w, **kw), O.LabellerError)


@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_oracle_labeller_matches_golden_sweep(sweep, mode):
    lab = O.oracle().labeller(384, 320, two_d=(mode == "2d"))
    for z in range(10):
        lab.add(z, sweep[f"signal_{z}"])
    lab.finish()
    got = lab.spots()
    want = sweep[f"spots_{mode}"]
    # The oracle sums in the reference's own order: bit-exact, variance too
    # (field by field: the record's padding is not initialised).
    compare_spots(got, want, tol=0.0)
    assert lab.pixels().tobytes() == sweep[f"pixels_{mode}"].tobytes()
    c = lab.counts()
    assert [c[k] for k in O.COUNTS_FIELDS] == list(sweep[f"counts_{mode}"])


class ModelLabeller:
    """sf.Labeller driven through the device contract by the Python model."""

    def __init__(self, h, w, **kw):
        import spotfinder_hip as sf
        self.lab = sf.Labeller(h, w, **kw)
        self.feed = ModelFeeder(self.lab, w)
        self.h, self.w = h, w

    def add(self, frame, signal):
        import spotfinder_hip as sf
        signal = np.ascontiguousarray(signal, sf.SIGNAL_PIXEL)
        idx = signal["index"].astype(np.int64)
        if len(idx) and (idx.max() >= self.h * self.w or (np.diff(idx) <= 0).any()):
            # what the device path refuses before labelling
            raise sf.LabellerError("bad pixel list")
        self.feed.add(frame, signal)

    def __getattr__(self, name):
        return getattr(self.lab, name)


def test_host_grouping_cases(sf):
    run_cases(lambda h, w, **kw: ModelLabeller(h, w, **kw), sf.LabellerError)


@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_host_grouping_matches_golden_sweep(sf, sweep, mode):
    lab = ModelLabeller(384, 320, two_d=(mode == "2d"))
    for z in range(10):
        lab.add(z, sweep[f"signal_{z}"])
    lab.finish()
    compare_spots(lab.spots(), sweep[f"spots_{mode}"])
    assert lab.pixels().tobytes() == sweep[f"pixels_{mode}"].tobytes()
    c = lab.counts()
    assert [c[k] for k in O.COUNTS_FIELDS] == list(sweep[f"counts_{mode}"])


def test_host_grouping_refuses_out_of_order_frames(sf):
    lab = sf.Labeller(8, 10)
    feed = ModelFeeder(lab, 10)
    feed.add(3, np.zeros(0, sf.SIGNAL_PIXEL))
    with pytest.raises(sf.LabellerError):
        feed.add(3, np.zeros(0, sf.SIGNAL_PIXEL))


def test_bad_grouping_options(sf):
    with pytest.raises(sf.LabellerError):
        sf.Labeller(0, 10)
    with pytest.raises(sf.LabellerError):
        sf.Labeller(8, 10, min_spot_size=5, max_spot_size=4)
    with pytest.raises(sf.LabellerError):
        sf.Labeller(8, 10, min_spot_size=0)


def long_lived_sweep(frames=160, h=128, w=128, hot=10, warm=600, transient=2400,
                     seed=5):
    """Frames whose components outlive their frames' pixel buffers: `hot`
    pixels on in every frame, `warm` pixels each on for a run of 5 to 15
    frames, and `transient` pixels in alternating halves of the frame, which
    never touch the next frame's."""
    import spotfinder_hip as sf
    rng = np.random.default_rng(seed)
    band = np.arange(h * w).reshape(h, w)
    hot_at = rng.choice(band[:, :16].ravel(), hot, replace=False)
    warm_at = rng.choice(band[:, 16:32].ravel(), warm, replace=False)
    start = rng.integers(0, frames, warm)
    stop = start + rng.integers(5, 16, warm)
    for z in range(frames):
        half = band[(z % 2) * (h // 2):(z % 2 + 1) * (h // 2), 32:].ravel()
        idx = np.concatenate([hot_at, warm_at[(start <= z) & (z < stop)],
                              rng.choice(half, transient, replace=False)])
        idx.sort()
        px = np.zeros(len(idx), sf.SIGNAL_PIXEL)
        px["index"] = idx
        px["value"] = rng.integers(5, 60, len(idx))
        yield z, px


def test_host_grouping_long_lived_components(sf):
    """Components open across most of the sweep: their frames' pixel buffers
    are released early and their pixels held apart, and that store is
    rewritten as the warm pixels' components close. The spots, their pixels
    and the counts equal the oracle's labeller."""
    kw = dict(max_spot_size=100000, max_separation=1000.0)
    ours = ModelLabeller(128, 128, **kw)
    want = O.oracle().labeller(128, 128, **kw)
    for z, px in long_lived_sweep():
        ours.add(z, px)
        want.add(z, px)
    ours.finish()
    want.finish()
    got = ours.spots()
    assert len(got) > 1000
    assert (got["n_signal"] == 160).sum() == 10  # the hot pixels, whole sweep
    compare_spots(got, want.spots())
    assert ours.pixels().tobytes() == want.pixels().tobytes()
    assert ours.counts() == want.counts()


@pytest.mark.parametrize("fault", ["link", "sizes", "component", "order"])
def test_refused_frame_changes_nothing(sf, sweep, fault):
    """A frame the host grouping refuses (sf_labeller_add_labelled) leaves
    the spots, pixels and counts as they were, and the same frame offered
    again correctly, then the rest of the sweep, groups exactly as the
    reference's Labeller does with no refusal at all."""
    from labelling_model import label_frame, links_between
    lab = sf.Labeller(384, 320)
    prev = None
    for z in range(10):
        sig = np.ascontiguousarray(sweep[f"signal_{z}"], sf.SIGNAL_PIXEL)
        comp_of, comps, rank = label_frame(sig, 320)
        links = np.zeros((0, 2), np.uint32)
        if prev is not None:
            links = links_between(sig, comp_of, prev[0], prev[1])
        if z == 4:
            before = (lab.spots().tobytes(), lab.pixels().tobytes(), lab.counts(),
                      lab.live())
            bad_comp_of, bad_comps, bad_links = comp_of.copy(), comps.copy(), links.copy()
            frame = z
            if fault == "link":
                bad_links = np.vstack([links, [[0, 100000]]]).astype(np.uint32)
            elif fault == "sizes":
                bad_comps["n"][0] += 1
            elif fault == "component":
                bad_comp_of[-1] = len(comps)
            else:
                frame = 2  # before the last frame added
            with pytest.raises(sf.LabellerError):
                lab.add_labelled(frame, sig, bad_comp_of, bad_comps, bad_links, True)
            assert (lab.spots().tobytes(), lab.pixels().tobytes(), lab.counts(),
                    lab.live()) == before
        lab.add_labelled(z, sig, comp_of, comps, links, prev is not None)
        prev = (rank, comp_of)
    lab.finish()
    compare_spots(lab.spots(), sweep["spots_3d"])
    assert lab.pixels().tobytes() == sweep["pixels_3d"].tobytes()
    c = lab.counts()
    assert [c[k] for k in O.COUNTS_FIELDS] == list(sweep["counts_3d"])
