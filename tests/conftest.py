import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "vina-cuda_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libvdock.so")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


@pytest.fixture(scope="session")
def ctx():
    import vdock
    c = vdock.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def module_ctx():
    """A device context of one test module's own.  Ligand ids count from 0 in every context and
    key the docking random streams (with the seed), so the statistical tests that dock through it
    draw the same streams whatever other tests ran before them."""
    import vdock
    c = vdock.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def host_ctx():
    import vdock
    c = vdock.Context(-1)
    yield c
    c.close()
