import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(TESTS, "golden", "data")
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def ctx():
    import onesolver_amd

    c = onesolver_amd.Context(0)
    print(f"\n[gpu] context on {c.device_name()}", flush=True)
    yield c
    c.close()
