import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


def gpu_available() -> bool:
    try:
        import graphwalker_amd as gw

        e = gw.Engine()
        e.close()
        return True
    except Exception:
        return False


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the case before it touches the GPU, so a run that dies mid-test
    # still says which case it was.
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield


@pytest.fixture(scope="session")
def engine():
    import graphwalker_amd as gw

    e = gw.Engine()
    info = e.device_info()
    print(f"\n[gpu] {info}", flush=True)
    yield e
    e.close()
