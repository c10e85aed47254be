"""The C oracle against the reference's own Fortran kernels.

Golden vectors (tests/golden/kernels_ref.*) were written by the flang-built
reference routines of solv_cellpoissonfacops{2d,3d}.f and
solv_cartesianrobinbchelper{2d,3d}.f; the oracle must reproduce them bit
for bit, including the single-precision Ewing weights and the 2/3 corner
factor, the patch-relative 2-D colouring and the truncating parity test for
negative indices.
"""
import json
import os

import numpy as np
import pytest

import kernel_cases as kc
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def load_golden():
    with open(os.path.join(HERE, "golden", "kernels_ref.json")) as f:
        meta = json.load(f)["cases"]
    store = np.load(os.path.join(HERE, "golden", "kernels_ref.npz"), allow_pickle=False)
    cases = []
    for m in meta:
        c = dict(m)
        c["arrays"] = {k: store[v] for k, v in m["inputs"].items()}
        c["expected"] = {k: store[v] for k, v in m["expected"].items()}
        cases.append(c)
    return cases


GOLDEN = load_golden()


def _id(c):
    return "%s-%dd-%s" % (c["kernel"], c["dim"], c.get("loc", c.get("offset", "")))


@pytest.mark.parametrize("case", GOLDEN, ids=[_id(c) for c in GOLDEN])
def test_oracle_matches_reference_golden(case):
    out = kc.run_oracle(oracle.oracle_lib(), case)
    for k, exp in case["expected"].items():
        assert np.array_equal(out[k], exp), (k, np.max(np.abs(out[k] - exp)))


def test_golden_covers_all_kernels():
    kinds = {(c["kernel"], c["dim"]) for c in GOLDEN}
    for dim in (2, 3):
        for k in ("compflux", "rbgs", "compres", "ewing", "settype1", "settype2"):
            assert (k, dim) in kinds
    assert ("settype3", 3) in kinds
    locs = {(c["dim"], c["loc"]) for c in GOLDEN if c["kernel"] == "ewing"}
    assert locs == {(3, l) for l in range(6)} | {(2, l) for l in range(4)}


@pytest.mark.skipif(oracle.ref_lib() is None, reason="reference kernels not built here")
def test_oracle_matches_reference_fresh_seed():
    lib = oracle.ref_lib()
    for c in kc.make_cases(seed=777):
        a = kc.run_ref(lib, c)
        b = kc.run_oracle(oracle.oracle_lib(), c)
        for k in a:
            assert np.array_equal(a[k], b[k]), (c["kernel"], c["dim"], k)
