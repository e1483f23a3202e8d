import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
for p in (ROOT, os.path.join(ROOT, "dials-metal_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a gfx950 (MI355X) device; run with -m gpu")


@pytest.fixture(scope="session")
def matrix():
    """The golden size/edge matrix: name -> (frame, reference signal list)."""
    z = np.load(os.path.join(GOLDEN, "threshold_matrix.npz"))
    return {str(n): (z[f"frame_{n}"], z[f"signal_{n}"]) for n in z["names"]}


@pytest.fixture(scope="session")
def synth_cases():
    import json
    with open(os.path.join(GOLDEN, "threshold_synth.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def sweep():
    return dict(np.load(os.path.join(GOLDEN, "labeller_sweep.npz")))


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    return O.oracle()


@pytest.fixture(scope="session")
def sf():
    import spotfinder_hip
    spotfinder_hip.lib()
    return spotfinder_hip


@pytest.fixture(scope="session")
def device(sf):
    if not sf.available():
        pytest.fail("no usable HIP device: " + sf.unavailable_reason())
    return sf
