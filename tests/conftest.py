import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ethersphcpu.jl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: full-size workload")


@pytest.fixture(scope="session")
def ctx():
    from ethersph import _lib

    c = _lib.Context(0)
    yield c
    c.close()
