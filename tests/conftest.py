import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # The oracle is plain C: build it here if it is missing (seconds).
    so = os.path.join(ROOT, "oracle", "liboracle_fac.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=False,
                       capture_output=True)


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    return torch


@pytest.fixture(scope="session")
def gpu_lib(torch_gpu):
    from samrai_amd import fac
    return fac.lib()
