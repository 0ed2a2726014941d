import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libphysis_rt_hip on the device)")


@pytest.fixture(scope="session")
def rt():
    """The HIP runtime, initialised once per test session."""
    from physis_amd import runtime
    runtime.init(["pytest"])
    yield runtime
