import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oct-mps_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

K_TRIPLETS = 4096


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def table():
    import octmps as O
    import oracle_py as OR
    return OR.Table(O.safeprimes(3 + 3 * K_TRIPLETS), K_TRIPLETS)


@pytest.fixture(scope="session")
def ctx():
    import octmps as O
    c = O.Context(0)
    yield c
    c.close()
