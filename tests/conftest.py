import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "spectral-init_amd"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DATASETS = [
    "blobs_50", "blobs_500", "blobs_5000", "moons_200", "circles_300",
    "near_dupes_100", "disconnected_200", "blobs_connected_200", "blobs_connected_2000",
]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def load_fixture(name):
    return dict(np.load(os.path.join(GOLDEN, "fixtures", f"{name}.npz"), allow_pickle=False))


@pytest.fixture(scope="session")
def fixtures():
    return {name: load_fixture(name) for name in DATASETS}
