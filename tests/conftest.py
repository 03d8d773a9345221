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


def scene_of(ctx, sc, layout):
    """A committed device scene in BVH layout 2, 4 (float 128-B nodes) or "4q" (quantised 64-B
    nodes), checked to have got that layout."""
    from banshee_amd import hip
    s = hip.Scene(ctx, sc, width=4 if layout == "4q" else layout, quantized=layout == "4q")
    info = s.info()
    assert info["width"] == (4 if layout == "4q" else layout) and info["quantized"] == (layout == "4q")
    assert info["node_bytes"] == (128 if layout == 4 else 64)
    return s


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
