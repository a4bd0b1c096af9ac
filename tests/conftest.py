import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "luminis-mc_amd"), ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libluminis_hip.so on the device)")
    config.addinivalue_line("markers", "slow: longer CPU-side statistical runs")


def gpu_available() -> bool:
    try:
        import luminis_mc._lib as L
        return L.lib().lmc_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device is visible to libluminis_hip.so")
