import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpbd.so")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    o.lib()
    return o
