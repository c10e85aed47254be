"""The device per-patch kernels (Fortran-signature entry points sfac_*)
against the reference's golden vectors: bit-exact."""
import numpy as np
import pytest

import kernel_cases as kc
from test_oracle_kernels import GOLDEN, _id

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", GOLDEN, ids=[_id(c) for c in GOLDEN])
def test_device_kernel_matches_reference_golden(gpu_lib, torch_gpu, case):
    out = kc.run_gpu(gpu_lib, case, torch_gpu)
    for k, exp in case["expected"].items():
        assert np.array_equal(out[k], exp), (k, np.max(np.abs(out[k] - exp)))
