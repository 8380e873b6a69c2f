import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    """Context on device 0.  Fails loudly if the HIP library is missing."""
    from voidray_amd import _native as N
    from voidray_amd.device import Context
    N.lib()
    return Context.default()


def announce(msg):
    """Names the GPU case about to run (flushed, so it survives a lost box)."""
    print(f"\n[gpu-case] {msg}", flush=True)
