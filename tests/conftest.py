import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on device 0)")


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def hip():
    """The HIP runtime of the product library; fails (never skips) without a GPU."""
    from vk_compute_mipmaps_amd import hip_rt

    n = hip_rt.device_count()
    assert n >= 1, "no HIP device visible"
    hip_rt.check(hip_rt.rt().hipSetDevice(0), "hipSetDevice(0)")
    return hip_rt


@pytest.fixture(scope="session")
def stream(hip):
    s = hip.stream_create()
    yield s
    hip.stream_sync(s)
    hip.stream_destroy(s)
