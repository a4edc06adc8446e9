import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")


@pytest.fixture(scope="session", autouse=True)
def _built():
    """Build the oracle if missing (CPU only; the HIP library is built by __graft_entry__.build())."""
    import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
