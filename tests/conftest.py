import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def pytest_collection_modifyitems(config, items):
    # GPU tests print what they run before each GPU step, so that a crash
    # names its case.
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(pytest.mark.usefixtures("_announce"))


@pytest.fixture
def _announce(request):
    print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
