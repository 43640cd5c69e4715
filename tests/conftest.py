import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
for _p in (TESTS, ROOT / "oracle"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs HIP kernels through the C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    return oracle_lib.load()


@pytest.fixture(scope="session")
def engine():
    from mppi_generic_vendor_amd import capi

    return capi.load()


@pytest.fixture(scope="session")
def hip():
    from mppi_generic_vendor_amd import hip as hipmod

    n = hipmod.device_count()
    if n < 1:
        raise RuntimeError("GPU tests need a visible HIP device")
    return hipmod
