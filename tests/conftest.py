import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PACKAGE = os.path.join(ROOT, "dials-metal-spotfinder_amd")
for path in (ROOT, PACKAGE):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def sf():
    """The product binding; on a machine with a GPU the library must load and a device
    must be present (no fallback)."""
    import spotfinder

    if not spotfinder.available():
        pytest.fail("no HIP device available for a gpu-marked test")
    spotfinder.report_windows(False)
    return spotfinder
