import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); calls the HIP library through its C ABI")


@pytest.fixture(scope="session")
def ref_fixtures():
    with open(os.path.join(HERE, "golden", "ref_fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle_tables():
    from oracle import oracle as O
    return O.Tables(16, 5)
