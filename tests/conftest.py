import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); calls the HIP library through its C ABI")


def pytest_collection_modifyitems(config, items):
    # run order inside a module is kept; nothing is skipped silently here
    pass


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def rimp2():
    import rimp2_gpu_amd as r
    r.lib()
    return r


@pytest.fixture(scope="module")
def ctx(rimp2):
    c = rimp2.Context(0)
    yield c
    c.close()
