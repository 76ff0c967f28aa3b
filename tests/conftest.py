import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built library")


@pytest.fixture(scope="session")
def ctx():
    import hilbertjoin_amd as hj
    c = hj.Context(0)
    yield c
    c.close()
