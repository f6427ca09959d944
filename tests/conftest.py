import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running")
    # The native libraries are built in-tree by __graft_entry__.build(); on a
    # fresh CPU checkout build them once here (never on a GPU box, where the
    # prebuilt objects travel with the tree).
    lib = os.path.join(ROOT, "nanogi_amd", "libngi.so")
    orc = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(orc):
        from oracle import build as ob
        ob.build()
    if not os.path.exists(lib):
        from nanogi_amd import build as nb
        nb.build()


@pytest.fixture(scope="session")
def has_gpu():
    import nanogi_amd as ngi
    return ngi.device_count() > 0


@pytest.fixture(autouse=True)
def _require_gpu(request):
    if request.node.get_closest_marker("gpu"):
        import nanogi_amd as ngi
        if ngi.device_count() <= 0:
            pytest.skip("no HIP device")
