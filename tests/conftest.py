import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "origami-sort_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(autouse=True)
def device_error_word_clear(request):
    """After every GPU test: each open library context's device error word
    (ogs_device_errors) is clear, so no merge tile was skipped on an
    inconsistent split table during the test's sorts."""
    yield
    if request.node.get_closest_marker("gpu") is None:
        return
    import ogs
    for c in ogs.live_contexts():
        flags = c.device_errors()
        assert flags == 0, f"device error word 0x{flags:x} set during {request.node.nodeid}"
