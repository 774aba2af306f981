import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture(scope="session")
def confocal():
    """First 20 frames (256x256 uint8) of the reference's confocal test stack."""
    with np.load(os.path.join(GOLDEN, "confocal_20.npz")) as z:
        return z["frames"]


@pytest.fixture(scope="session")
def oracle_lib():
    """The C restatement, built on demand (CPU only)."""
    from oracle import port

    if not os.path.isfile(port._LIB_PATH):
        port.build()
    return port


@pytest.fixture(scope="session")
def hip():
    """The HIP core on a GPU; fails (never skips) when the library or GPU is missing."""
    from fastddm_amd import _core_hip

    n = _core_hip.get_num_devices()
    if n < 1:
        pytest.fail("no AMD GPU visible to the HIP library")
    _core_hip.set_device(0)
    return _core_hip
