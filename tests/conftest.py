import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session", autouse=True)
def _libs_built():
    import nigiri_amd
    if not os.path.exists(nigiri_amd.LIB_PATH):
        from nigiri_amd import build
        build.build()
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    yield
