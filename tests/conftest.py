import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("op-pic_amd", "oracle", os.path.join("tests", "golden")):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def golden_path(name):
    return os.path.join(ROOT, "tests", "golden", name)


@pytest.fixture(scope="session")
def root():
    return ROOT
