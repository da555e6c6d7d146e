import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "reference_values.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu_lib():
    """Loads libqed.so and initializes the device; a missing library or GPU is an error
    in a GPU test (there is no CPU fallback)."""
    from pyqcm_amd import _lib
    _lib.call("qed_init", 0)
    return _lib
