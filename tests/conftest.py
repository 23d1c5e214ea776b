import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    # tuned scheduling knobs come from a cache of the session's own, not the user's
    if "PYDENSECRF_AMD_TUNE_CACHE" not in os.environ:
        import tempfile
        os.environ["PYDENSECRF_AMD_TUNE_CACHE"] = os.path.join(tempfile.mkdtemp(prefix="dcrf_tune_"), "tune.json")
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the HIP kernels through the C ABI")


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the case before it touches the GPU, so a failure that ends the process is attributable.
    if request.node.get_closest_marker("gpu") is not None:
        print("\n[case] %s" % request.node.nodeid, flush=True)
    yield
