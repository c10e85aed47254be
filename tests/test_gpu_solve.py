"""End-to-end FAC solves: GPU CellPoissonFACSolver against the oracle FAC.

Same hierarchy, data and input keys; the residual history must agree to
1e-12 relative per cycle and the solution per cell to 1e-12 relative
(north star: FP64 path within 1e-12 per cell of the reference kernels).
"""
import numpy as np
import pytest

import hier_cases as hc

pytestmark = pytest.mark.gpu

PARAMS = dict(max_cycles=40, residual_tol=1e-300, relative_residual_tol=1e-10,
              num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_choice="redblack",
              coarse_solver_tolerance=1e-12, coarse_solver_max_iterations=200,
              prolongation_method="CONSERVATIVE_LINEAR_REFINE")


def smooth_data(case, seed):
    """u = 0, f = -dim pi^2 prod sin(pi x) at cell centres (SAMRAI sign)."""
    data = hc.make_data(case, seed=seed)
    dim = case["dim"]
    for ln, boxes in enumerate(case["levels"]):
        dx = hc.level_dx(case, ln)
        for p, (lo, hi) in enumerate(boxes):
            X = np.meshgrid(*[(np.arange(lo[d], hi[d] + 1) + 0.5) * dx[d] for d in range(dim)],
                            indexing="ij")
            data["f"][ln][p] = -dim * np.pi ** 2 * np.prod([np.sin(np.pi * x) for x in X], axis=0)
            data["u"][ln][p] = np.zeros_like(data["u"][ln][p])
    return data


@pytest.mark.parametrize("spec", [dict(dvar=False), dict(dvar=True)], ids=["constD", "varD"])
@pytest.mark.parametrize("mk", hc.ALL + [hc.case_2d_c0], ids=lambda f: f.__name__)
def test_solve_matches_oracle(gpu_lib, mk, spec):
    case = mk()
    data = smooth_data(case, 3)
    pr = hc.Pair(case, PARAMS, data, **spec)
    print("solve", case["name"], spec, flush=True)
    conv_g = pr.s.solve_system(pr.u, pr.f)
    hist_g = pr.s.history()
    conv_o, it_o, res_o, hist_o = pr.o.solve()
    assert conv_g == conv_o
    assert pr.s.number_of_iterations() == it_o
    np.testing.assert_allclose(hist_g, hist_o, rtol=1e-12, atol=0)
    pr.compare(pr.u, "u", interior_only=True, rtol=1e-12)
