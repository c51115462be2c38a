import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def po():
    import pyoracle

    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def chs():
    from chapsim2_amd import _lib

    return _lib.load()
