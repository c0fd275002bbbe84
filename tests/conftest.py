"""Shared test setup: import paths and the `gpu` marker.

GPU tests (marked `gpu`) call libtwave.so through the C-ABI on a HIP device;
everything else runs on the CPU.  Results are checked against the CPU oracle (oracle/).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tpu-wave_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU run (full-length reference comparison)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def twave_mod():
    import twave
    twave.lib()
    return twave


@pytest.fixture
def progress(capfd):
    """Print a progress line past pytest's output capture: the long full-size
    tests (minutes each) report every phase as it ends."""
    def say(*msg):
        with capfd.disabled():
            print(*msg, flush=True)
    return say
