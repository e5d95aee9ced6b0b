import os
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the library's tuning cache: one file per test session, not the user's
os.environ.setdefault("VDB_TUNE_CACHE",
                      os.path.join(tempfile.gettempdir(), "vdb_tune_tests_%d.json" % os.getpid()))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")
    config.addinivalue_line("markers", "slow: long-running")


GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
