"""Colour-fused sweeps of the levels above the coarsest (one launch per
sweep over all patches, same-level neighbours read in place of the ghost
fills between colours) against the patch-wise colour passes and the oracle.

Bar: bit for bit.  The smoothed error on every level, ghosts included, equals
the oracle's (smoothErrorByRedBlack, CellPoissonFACOps.C:1338-1514) for every
tile shape, odd and even sweep counts, constant and variable D, C = 0 /
constant / cell data, Dirichlet and Robin boundaries, the Ewing and the
refine-operator coarse-fine discretizations; whole FAC solves give the same
residual history and solution bits with the fused sweeps on and off."""
import numpy as np
import pytest

import hier_cases as hc
from test_gpu_solve import c2_like_case

pytestmark = pytest.mark.gpu

PARAMS = dict(max_cycles=4, num_pre_sweeps=2, num_post_sweeps=2,
              prolongation_method="CONSERVATIVE_LINEAR_REFINE",
              coarse_solver_max_iterations=20, coarse_solver_tolerance=1e-12)
SPECS = {"constD": dict(dvar=False, cmode=0),
         "varD-varC-robin": dict(dvar=True, cmode=2, robin=True),
         "constC": dict(dvar=False, cmode=1, cconst=-0.5)}


def _c2like():
    return c2_like_case(2)


_c2like.__name__ = "c2like"
CASES = [hc.case_3d, hc.case_3d_boundary, hc.case_3d_ratio4, _c2like]


def _make(mk, spec_name, params, seed=5):
    spec = dict(SPECS[spec_name])
    out = mk()
    if isinstance(out, tuple):
        case, data = out
        spec["dvar"] = True
    else:
        case = out
        data = hc.make_data(case, seed=seed)
    if spec.pop("robin", False):
        spec["bc"] = hc.bc_robin(case["dim"])
    return case, hc.Pair(case, params, data, **spec)


@pytest.mark.parametrize("cf", ["Ewing", "LINEAR_REFINE"])
@pytest.mark.parametrize("spec", list(SPECS))
@pytest.mark.parametrize("mk", CASES, ids=lambda f: f.__name__)
def test_fused_level_smoothing_matches_oracle(gpu_lib, mk, spec, cf):
    case, pr = _make(mk, spec, dict(PARAMS, cf_discretization=cf))
    nl = len(case["levels"])
    nvar = 5
    for v in range(nvar):
        pr.s.set_fused_levels(True, v)
        for ln in range(1, nl):
            assert pr.s.fused_level_variant(ln) == v
    # sweep counts 1, 2, 3 in turn on every finer level, tile shape varying
    step = 0
    for sweeps in (1, 2, 3):
        for ln in range(nl - 1, 0, -1):
            v = step % nvar
            step += 1
            print("fused smooth", case["name"], spec, cf, ln, sweeps, "variant", v, flush=True)
            pr.s.set_fused_levels(True, v)
            pr.s.smooth(pr.u, pr.f, ln, sweeps)
            pr.o.smooth("u", "f", ln, sweeps)
            pr.compare(pr.u, "u", levels=list(range(nl)))


@pytest.mark.parametrize("spec", ["constD", "varD-varC-robin"])
@pytest.mark.parametrize("mk", CASES, ids=lambda f: f.__name__)
def test_fused_level_solves_are_bit_identical(gpu_lib, mk, spec):
    """Whole solves (graphs on): colour passes vs fused sweeps of every tile
    shape and the autotuned one."""
    ref = None
    for on, v in [(False, 0), (True, -1), (True, 0), (True, 1), (True, 2), (True, 3),
                  (True, 4)]:
        case, pr = _make(mk, spec, dict(PARAMS, max_cycles=3))
        pr.s.set_fused_levels(on, v)
        if on and v >= 0:
            assert all(pr.s.fused_level_variant(ln) == v for ln in range(1, len(case["levels"])))
        if not on:
            assert all(pr.s.fused_level_variant(ln) == -1 for ln in range(len(case["levels"])))
        print("fused solve", case["name"], spec, on, v, flush=True)
        pr.s.solve_system(pr.u, pr.f)
        u = [pr.u.get(ln, p) for ln in range(len(case["levels"]))
             for p in range(len(case["levels"][ln]))]
        got = (pr.s.number_of_iterations(), np.array(pr.s.history()), u)
        if ref is None:
            ref = got
            continue
        assert got[0] == ref[0]
        np.testing.assert_array_equal(got[1], ref[1])
        for a, b in zip(got[2], ref[2]):
            np.testing.assert_array_equal(a, b)
