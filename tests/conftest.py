import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libwovengpu.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_py
    oracle_py.lib()
    return oracle_py


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to count devices; no HIP call is made here)
        return torch.cuda.device_count() > 0
    except Exception:
        return False
