"""Mixed precision: the coarsest-level solve of the error equation in FP32.

With coarse precision 32 the single-rectangle coarsest level is smoothed
(red-black) or solved (multigrid) in FP32 storage and arithmetic, while the
FAC iteration, its composite residuals, the finer levels and u stay FP64.
The outer iteration still tests the FP64 residual, so the solve must reach
the same relative tolerance (1e-10) and its solution must equal the FP64
oracle's converged solution to the accuracy that tolerance fixes: both
differ from the exact discrete solution by at most ||A^-1|| times the
residual, a few 1e-10 at these data, so the per-level max difference is
held to 1e-8 of max |u|.  The FP32 path must also change the arithmetic
(histories differ from the FP64 run) and switching back to 64 bits must
restore the FP64 bits.
"""
import numpy as np
import pytest

import hier_cases as hc
from test_gpu_solve import smooth_data

pytestmark = pytest.mark.gpu

PARAMS = dict(max_cycles=60, residual_tol=1e-300, relative_residual_tol=1e-10,
              num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_tolerance=1e-12,
              coarse_solver_max_iterations=200,
              prolongation_method="CONSERVATIVE_LINEAR_REFINE")
MG = dict(coarse_solver_choice="mg", coarse_solver_max_iterations=20,
          coarse_solver_tolerance=1e-10)
RB = dict(coarse_solver_choice="redblack")


def _level_max_diff(pr, case):
    out = []
    for ln, boxes in enumerate(case["levels"]):
        m = 0.0
        for p in range(len(boxes)):
            g = pr.u.get(ln, p)
            o = pr.o.get("u", ln, p)
            shp = hc.cell_shape(boxes[p], 1, case["dim"])
            inner = (slice(1, -1),) * case["dim"]
            d = (g.reshape(shp, order="F") - o.reshape(shp, order="F"))[inner]
            m = max(m, float(np.abs(d).max()))
        out.append(m)
    return out


@pytest.mark.parametrize("choice", ["redblack", "mg"])
@pytest.mark.parametrize("spec", [dict(dvar=False), dict(dvar=True, cmode=2, robin=True)],
                         ids=["constD", "varD-varC-robin"])
@pytest.mark.parametrize("mk", [hc.case_3d, hc.case_3d_boundary, hc.case_3d_wide0,
                                hc.case_2d_c0], ids=lambda f: f.__name__)
def test_fp32_coarse_solve_reaches_the_fp64_solution(gpu_lib, mk, spec, choice):
    case = mk()
    spec = dict(spec)
    if spec.pop("robin", False):
        spec["bc"] = hc.bc_robin(case["dim"])
    data = smooth_data(case, 3)
    params = dict(PARAMS, **(MG if choice == "mg" else RB))
    pr = hc.Pair(case, params, data, **spec)
    assert pr.s.fused_coarsest_active()
    pr.s.set_coarse_precision(32)
    print("mixed", case["name"], spec.get("dvar"), choice, flush=True)
    conv_g = pr.s.solve_system(pr.u, pr.f)
    hist32 = list(pr.s.history())
    conv_o, it_o, res_o, hist_o = pr.o.solve()
    assert conv_o and conv_g
    assert pr.s.number_of_iterations() <= it_o + 3
    assert hist32[-1] <= 1e-10 * pr.o.rms_f() * (1 + 1e-12)
    scale = max(float(np.abs(pr.o.get("u", ln, p)).max())
                for ln in range(len(case["levels"])) for p in range(len(case["levels"][ln])))
    diffs = _level_max_diff(pr, case)
    print("  cycles fp32 %d fp64 %d, per-level max |u32 - u64| / max|u|:" %
          (pr.s.number_of_iterations(), it_o), ["%.2e" % (d / scale) for d in diffs], flush=True)
    assert max(diffs) <= 1e-8 * scale
    # the precision switch changes the arithmetic, and switching back
    # restores the FP64 bits of a fresh solver
    pr64 = hc.Pair(case, params, data, **spec)
    pr64.s.solve_system(pr64.u, pr64.f)
    assert list(pr64.s.history()) != hist32
    pr.s.set_coarse_precision(64)
    for ln, boxes in enumerate(case["levels"]):
        for p in range(len(boxes)):
            pr.u.put(ln, p, data["u"][ln][p])
    pr.s.solve_system(pr.u, pr.f)
    assert list(pr.s.history()) == list(pr64.s.history())


def test_coarse_precision_rejects_other_widths(gpu_lib):
    from samrai_amd import fac
    case = hc.case_3d()
    pr = hc.Pair(case, dict(PARAMS, **RB), smooth_data(case, 3))
    with pytest.raises(fac.SolverError):
        pr.s.set_coarse_precision(16)
