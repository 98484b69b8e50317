import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded once; the GPU tests fail loudly without it."""
    from serac_amd import capi

    lib = capi.load()
    capi.call("sx_init", 0)
    return lib


def load_mesh_fixture(name):
    import json

    import numpy as np

    with open(os.path.join(ROOT, "tests", "golden", name + ".json")) as fh:
        d = json.load(fh)
    return np.array(d["vertices"], dtype=np.float64), np.array(d["hexes"], dtype=np.int32)


@pytest.fixture(scope="session")
def patch_mesh():
    """data/meshes/patch3D_hexes.mesh of the reference: 7 distorted hexes."""
    return load_mesh_fixture("patch3D_hexes")
