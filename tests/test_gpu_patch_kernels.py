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


F77_KINDS = ("compflux", "rbgs", "compres", "ewing", "settype1", "settype2", "settype3")
F77_CASES = [c for c in GOLDEN if c["kernel"] in F77_KINDS]


@pytest.mark.parametrize("case", F77_CASES, ids=[_id(c) for c in F77_CASES])
def test_fortran_link_twins_match_reference_golden(gpu_lib, torch_gpu, case):
    """The link-compatible twins (include/samrai_f77.h: the Fortran names and
    by-reference arguments, device arrays, host scalars and maxres) give the
    reference's golden outputs bit for bit."""
    out = kc.run_ref(gpu_lib, case, torch_gpu)
    for k, exp in case["expected"].items():
        assert np.array_equal(out[k], exp), (k, np.max(np.abs(out[k] - exp)))
