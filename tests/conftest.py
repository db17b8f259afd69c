"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and call the HIP
library through its C ABI; everything else runs on the CPU."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libparhde.so")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is visible; fail loudly if it is and the library is missing."""
    from icpp_2020_parhde_amd import _native
    lib = _native.parhde_lib()  # raises when not built: no silent fallback
    import ctypes
    c = ctypes.c_int(0)
    lib.phde_device_count(ctypes.byref(c))
    if c.value < 1:
        pytest.skip("no GPU visible")
    return lib


@pytest.fixture(autouse=True)
def _name_the_case(request):
    # flush the case name before any GPU work so a failing box still tells us which
    if request.node.get_closest_marker("gpu"):
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
