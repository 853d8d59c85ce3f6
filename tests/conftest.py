import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libradeonrays.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def ctx():
    import torch

    from radeonrays_sdk_amd import api

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = api.Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()


@pytest.fixture(scope="session")
def dev():
    import torch

    return torch.device("cuda", 0)
