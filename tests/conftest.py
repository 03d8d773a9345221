import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def ctx():
    from banshee_amd import hip
    c = hip.Context(0)
    yield c


@pytest.fixture(scope="session")
def cornell():
    from banshee_amd.scenes import cornell_box
    return cornell_box()


@pytest.fixture(scope="session")
def terrain100k():
    from banshee_amd.scenes import terrain_for_tris
    from golden.make_golden import TERRAIN, TERRAIN_TRIS
    return terrain_for_tris(TERRAIN_TRIS, seed=1, width=TERRAIN["W"], height=TERRAIN["H"])
