import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["tri2d", "bump2d", "box3d", "box3d_super"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(params=CASES)
def golden(request):
    import numpy as np
    return request.param, np.load(os.path.join(GOLDEN, request.param + ".npz"))


def golden_mesh(g):
    from bes_amd import meshgen as mg
    return mg.MeshData(int(g["ndim"]), g["nodes"], g["cells"], g["bfaces"], g["btype"])
