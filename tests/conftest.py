import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmp_gp_hip.so")


@pytest.fixture(scope="session")
def fixture_dir():
    return os.path.join(ROOT, "tests", "golden")
