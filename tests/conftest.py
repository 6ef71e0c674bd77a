import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs")


@pytest.fixture(scope="session")
def pf():
    from _support import load_pf
    return load_pf()


@pytest.fixture(scope="session")
def gpu(pf):
    """The library with a visible device; fails (never skips) without one."""
    n = pf.lib().cuda_get_device_count()
    assert n >= 1, "no HIP device visible"
    print(f"[gpu] {pf.device_info()}", flush=True)
    return pf
