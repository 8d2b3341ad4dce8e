import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libhx.so")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle import hx_oracle
    hx_oracle.lib()
    return hx_oracle


@pytest.fixture(scope="session")
def ctx():
    import torch
    # GPU tests fail (never skip) without a device: they run only under -m gpu.
    assert torch.cuda.is_available(), "GPU test selected but torch sees no ROCm device"
    import hippie_amd
    return hippie_amd.Context.default(0)
