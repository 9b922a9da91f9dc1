import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the gfx950 library)")
    import xdem_ref
    xdem_ref.set_threads(min(8, os.cpu_count() or 1))


def load_fixture(name: str):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    return d["pos"].copy(), d["vel"].copy()


@pytest.fixture(scope="session")
def sphere_drop():
    return load_fixture("sphere_drop_particles")


@pytest.fixture(scope="session")
def particle_block():
    return load_fixture("particleposition")
