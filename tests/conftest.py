import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from tests._oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import kernels_amd
    # fail loudly if the library is missing; a host fault during the GPU
    # tests prints its native frames before pytest's Python traceback
    kernels_amd.check(kernels_amd.lib().prk_crash_backtrace(1))
    return torch.device("cuda:0")
