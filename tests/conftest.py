import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libcoralhip.so")


@pytest.fixture(scope="session")
def gpu_lib():
    """The built library, initialised on a gfx950 device (fails loudly otherwise)."""
    import coral_amd

    lib = coral_amd.load()
    status = lib.coral_init()
    assert status == 0, f"coral_init failed: {lib.coral_last_error().decode()}"
    return lib


def allclose_ref(x, y, rtol, atol):
    """The reference tests' element check: |x-y| <= atol + rtol*max(|x|,|y|)
    (coral-aarch64/tests/level3/sgemm_tests.rs:64-81)."""
    x = np.asarray(x)
    y = np.asarray(y)
    return np.all(np.abs(x - y) <= atol + rtol * np.maximum(np.abs(x), np.abs(y)))
