import os
import sys

import pytest

# every thread insertion and removal in the tests is followed by the
# structural checks of the whole ARG (abort on a failure)
os.environ.setdefault("AWAMD_CHECK_TREES", "1")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def aw():
    import argweaver_amd.argweaverc as A
    return A
