"""Level-batched GPU FACOperatorStrategy operations against the oracle.

Each operation of the GPU CellPoissonFACOps (ghost fills, red-black
smoothing, composite residual with the Ewing fix and flux coarsening,
restriction, prolongation, weighted RMS norm) runs on the same seeded data
as the oracle; cell values must agree bit for bit (both follow the reference
expression order; the GPU library is built with -ffp-contract=off), norms to
1e-13 relative (parallel summation order).
"""
import numpy as np
import pytest

import hier_cases as hc

pytestmark = pytest.mark.gpu

SPECS = [dict(dvar=False, cmode=0), dict(dvar=True, cmode=2), dict(dvar=False, cmode=1,
                                                                    cconst=-0.5)]
PARAMS = dict(max_cycles=4, num_pre_sweeps=2, num_post_sweeps=2,
              prolongation_method="CONSERVATIVE_LINEAR_REFINE",
              coarse_solver_max_iterations=20, coarse_solver_tolerance=1e-12)


def _pair(mk, spec, seed=1, bc=None, params=None):
    case = mk()
    data = hc.make_data(case, seed=seed)
    kw = dict(spec)
    return hc.Pair(case, params or PARAMS, data, bc=bc, **kw)


@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_ghost_fill(gpu_lib, mk):
    case = mk()
    pr = _pair(mk, SPECS[0], bc=hc.bc_robin(case["dim"]))
    for ln in range(len(case["levels"])):
        for wc in ((0, 1) if ln > 0 else (0,)):
            for hom in (0, 1):
                print("ghost fill", case["name"], ln, wc, hom, flush=True)
                pr.s.ghost_fill(pr.u, ln, wc, hom)
                pr.o.ghost_fill("u", ln, wc, hom)
                pr.compare(pr.u, "u", levels=[ln])


CF_METHODS = ["CONSTANT_REFINE", "LINEAR_REFINE", "CONSERVATIVE_LINEAR_REFINE"]


@pytest.mark.parametrize("cf", CF_METHODS)
@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_ghost_fill_cf_discretization(gpu_lib, mk, cf):
    """Coarse-fine ghosts by the cf_discretization refine operator; the
    linear operators first fill the coarser levels' ghosts (recursively),
    so every level is compared."""
    case = mk()
    pr = _pair(mk, SPECS[0], bc=hc.bc_robin(case["dim"]),
               params=dict(PARAMS, cf_discretization=cf))
    for ln in range(1, len(case["levels"])):
        for hom in (0, 1):
            print("cf ghost fill", case["name"], cf, ln, hom, flush=True)
            pr.s.ghost_fill(pr.u, ln, 1, hom)
            pr.o.ghost_fill("u", ln, 1, hom)
            pr.compare(pr.u, "u", levels=list(range(ln + 1)))


@pytest.mark.parametrize("cf", CF_METHODS)
@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_smooth_and_residual_cf_discretization(gpu_lib, mk, cf):
    """Smoothing and composite residual without the Ewing flux fix, with the
    C/F ghosts of each refine operator (variable D and C)."""
    case = mk()
    pr = _pair(mk, SPECS[1], params=dict(PARAMS, cf_discretization=cf))
    lmax = len(case["levels"]) - 1
    for ln in range(lmax, 0, -1):
        print("cf smooth/residual", case["name"], cf, ln, flush=True)
        pr.s.smooth(pr.u, pr.f, ln, 2)
        pr.o.smooth("u", "f", ln, 2)
        pr.compare(pr.u, "u", levels=[ln])
        for hom in (1, 0):
            pr.s.residual(pr.r, pr.u, pr.f, ln, hom)
            pr.o.residual("r", "u", "f", ln, hom)
            pr.compare(pr.r, "r", levels=[ln])


@pytest.mark.parametrize("spec", SPECS, ids=["constD", "varD-varC", "constC"])
@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_smooth(gpu_lib, mk, spec):
    case = mk()
    pr = _pair(mk, spec)
    for ln in range(len(case["levels"])):
        print("smooth", case["name"], ln, flush=True)
        pr.s.smooth(pr.u, pr.f, ln, 3)
        pr.o.smooth("u", "f", ln, 3)
        pr.compare(pr.u, "u", levels=[ln])


@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_smooth_with_tolerance_matches_sweep_count(gpu_lib, mk):
    case = mk()
    pr = _pair(mk, SPECS[0])
    # a tolerance reached after a few sweeps: the early exit uses the last
    # patch's max residual of each colour (CellPoissonFACOps.C:1432-1507)
    n_o, mr_o = pr.o.smooth("u", "f", 0, 1, 0.0)  # tol >= 0: maxres is computed
    assert mr_o > 0
    tol = mr_o * 0.3
    pr2 = _pair(mk, SPECS[0])
    n_g, mr_g = pr2.s.smooth(pr2.u, pr2.f, 0, 200, tol)
    n_o, mr_o = pr2.o.smooth("u", "f", 0, 200, tol)
    assert n_g == n_o and n_g < 200
    assert mr_g == mr_o
    pr2.compare(pr2.u, "u", levels=[0])


@pytest.mark.parametrize("spec", SPECS, ids=["constD", "varD-varC", "constC"])
@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_composite_residual_and_restriction(gpu_lib, mk, spec):
    case = mk()
    pr = _pair(mk, spec)
    lmax = len(case["levels"]) - 1
    for hom in (1, 0):
        for ln in range(lmax, -1, -1):
            print("residual", case["name"], ln, hom, flush=True)
            pr.s.residual(pr.r, pr.u, pr.f, ln, hom)
            pr.o.residual("r", "u", "f", ln, hom)
            pr.compare(pr.r, "r", levels=[ln])
            pr.compare(pr.u, "u", levels=[ln])
            if ln > 0:
                pr.s.restrict_residual(pr.r, pr.r, ln - 1)
                pr.o.restrict("r", "r", ln - 1)
                pr.compare(pr.r, "r", levels=[ln - 1])
    g = pr.s.rms_norm(pr.r)
    o = pr.o.rms("r")
    assert abs(g - o) <= 1e-13 * abs(o)


@pytest.mark.parametrize("method", ["CONSERVATIVE_LINEAR_REFINE", "LINEAR_REFINE",
                                    "CONSTANT_REFINE"])
@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_restrict_solution_and_prolong(gpu_lib, mk, method):
    case = mk()
    pr = _pair(mk, SPECS[0], bc=hc.bc_robin(case["dim"]),
               params=dict(PARAMS, prolongation_method=method))
    lmax = len(case["levels"]) - 1
    for ln in range(lmax - 1, -1, -1):
        print("restrict solution", case["name"], ln, flush=True)
        pr.s.restrict_solution(pr.u, ln)
        pr.o.restrict_solution("u", ln)
        pr.compare(pr.u, "u", levels=[ln])
    for ln in range(1, lmax + 1):
        print("prolong", case["name"], ln, method, flush=True)
        pr.s.prolong_and_correct(pr.u, ln)
        pr.o.prolong("u", ln)
        pr.compare(pr.u, "u", levels=[ln - 1, ln])
