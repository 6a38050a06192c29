import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "rust-sloth_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def sloth():
    """The GPU library, initialised on device 0.  Fails (never falls back)
    when the library or the device is missing."""
    import sloth_amd

    sloth_amd.init(0)
    return sloth_amd
