import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run on the MI355X box)")


@pytest.fixture(scope="session")
def gpu():
    """The product library with a visible gfx950 device; skips if there is no GPU."""
    import convstencil_amd as cs
    from convstencil_amd import _lib
    _lib.load()  # raises if the library is not built: never a silent fallback
    if cs.device_count() < 1:
        pytest.skip("no gfx950 device visible")
    return cs
