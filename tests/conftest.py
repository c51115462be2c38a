import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def pytest_sessionstart(session):
    """torch (used by a few GPU tests for device buffers) carries its own HIP
    runtime; it must initialise the device before libchapsim_hip does, as in
    bench.py, or it reports no GPU."""
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


@pytest.fixture(scope="session")
def po():
    import pyoracle

    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def chs():
    from chapsim2_amd import _lib

    return _lib.load()
