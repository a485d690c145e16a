import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def book():
    s = np.load(os.path.join(GOLDEN, "book_spheres.npy"))
    m = np.load(os.path.join(GOLDEN, "book_materials.npy"))
    return s, m


@pytest.fixture(scope="session")
def ctx():
    """One device context per test session (the reference's Renderer context)."""
    from venusaur_amd import api
    c = api.Context(0)
    yield c
    c.close()
