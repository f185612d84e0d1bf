import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmfsv_hip.so)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded; GPU tests must run through it (no fallback)."""
    from mfcm_amd import mfsv
    lib = mfsv.lib()
    if mfsv.device_count() < 1:
        pytest.fail("no HIP device visible to libmfsv_hip.so")
    return lib
