import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "tests"), os.path.join(REPO, "relativity-simulator_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and librtsim.so")
    config.addinivalue_line("markers", "ref: needs oracle/_ref/libref.so (reference headers at build time)")


@pytest.fixture(scope="session")
def ctx():
    import rtsim
    c = rtsim.Context(0)
    yield c
    c.close()
