import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def reference_dir():
    if not os.path.isdir(os.path.join(REFERENCE, "scenes")):
        pytest.skip("reference checkout not present (only in the build container)")
    return REFERENCE
