import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "syntheticscrnaseq.jl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built libscrna.so")
    config.addinivalue_line("markers", "slow: longer statistical test")


@pytest.fixture(scope="session")
def ctx():
    import scrna_amd as sc
    return sc.default_context()
