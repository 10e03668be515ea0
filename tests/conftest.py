"""Shared fixtures.  Tests marked `gpu` need an MI355X (gfx950) and the
built libnanomap_amd.so; everything else runs on the CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs a gfx950 GPU and libnanomap_amd.so")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def nm():
    import nanomap_amd
    nanomap_amd.lib()
    return nanomap_amd


def gpu_device_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu(nm):
    """Skip-free guard: a `gpu` test on a machine without a device fails
    loudly instead of passing on a fallback."""
    import torch
    assert torch.cuda.is_available(), "gpu test needs a HIP device"
    return nm
