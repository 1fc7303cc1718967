import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdpa_hip.so)")


@pytest.fixture(scope="session")
def ctx():
    """One dpah context for the GPU tests.  No fallback: a missing library or
    GPU fails the test."""
    import dpa_amd
    c = dpa_amd.Context(0, kernel_timing=True)
    yield c
    c.close()
