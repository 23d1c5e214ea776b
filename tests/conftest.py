import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the HIP kernels through the C ABI")


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the case before it touches the GPU, so a failure that ends the process is attributable.
    if request.node.get_closest_marker("gpu") is not None:
        print("\n[case] %s" % request.node.nodeid, flush=True)
    yield
