import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running parity case")


def pytest_runtest_setup(item):
    # print what names the case before anything touches the GPU
    if item.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-case] {item.nodeid}", flush=True)


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def capi():
    from icarat2_amd import _capi
    _capi.lib()
    return _capi


@pytest.fixture(scope="session")
def gpu(capi):
    if not capi.device_available():
        pytest.fail("GPU test run without a usable HIP device")
    return capi
