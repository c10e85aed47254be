"""Mixed precision on every level: the FAC iteration's error vector in FP32
or BF16 storage (smoothing, the error's ghost fills and the prolongation in
FP32 arithmetic, D in a low-precision copy), residuals, coarse-fine flux
corrections, restriction and u += e in FP64.

* Level smoothing (the low-precision colour passes with the Ewing fix and
  the ghost fills) against the oracle's FP64 smoothing of the same
  (rounded) error and residual: max difference per level within the
  storage precision times max |e| (FP32 4e-6, BF16 2e-2; a few sweeps
  compound the unit roundoff 6e-8 / 4e-3).
* Whole solves to the FP64 relative residual 1e-10 (the convergence test
  stays FP64): the solution equals the oracle's FP64 solution per level to
  the accuracy that tolerance fixes (1e-8 of max |u|, as for the FP32
  coarse solve), the precision changes the arithmetic, and 64 bits restores
  the FP64 bits.
"""
import numpy as np
import pytest

import hier_cases as hc
from test_gpu_solve import c2_like_case, smooth_data

pytestmark = pytest.mark.gpu

PARAMS = dict(max_cycles=80, residual_tol=1e-300, relative_residual_tol=1e-10,
              num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_choice="mg",
              coarse_solver_tolerance=1e-10, coarse_solver_max_iterations=20,
              prolongation_method="CONSERVATIVE_LINEAR_REFINE")
SMOOTH_TOL = {32: 4e-6, 16: 2e-2}
ROUND = {32: lambda a: a.astype(np.float32).astype(np.float64),
         16: lambda a: _bf16(a)}


def _bf16(a):
    """Round float64 values to bfloat16 (nearest, ties to even) and back."""
    f = np.asarray(a, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _case(mk, spec):
    out = mk()
    spec = dict(spec)
    if isinstance(out, tuple):
        case, data = out
        spec["dvar"] = True
    else:
        case = out
        data = smooth_data(case, 3)
    if spec.pop("robin", False):
        spec["bc"] = hc.bc_robin(case["dim"])
    return case, data, spec


def _interior_max(case, ln, get):
    m = 0.0
    for p, box in enumerate(case["levels"][ln]):
        shp = hc.cell_shape(box, 1, case["dim"])
        v = get(p).reshape(shp, order="F")[(slice(1, -1),) * case["dim"]]
        m = max(m, float(np.abs(v).max()))
    return m


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("cf", ["Ewing", "LINEAR_REFINE"])
@pytest.mark.parametrize("spec", [dict(dvar=False), dict(dvar=True, cmode=2, robin=True)],
                         ids=["constD", "varD-varC-robin"])
@pytest.mark.parametrize("mk", [hc.case_3d, hc.case_3d_boundary, hc.case_2d],
                         ids=lambda f: f.__name__)
def test_low_precision_smoothing_close_to_fp64(gpu_lib, mk, spec, cf, bits):
    case, _, spec = _case(mk, spec)
    data = hc.make_data(case, seed=11)
    pr = hc.Pair(case, dict(PARAMS, cf_discretization=cf), data, **spec)
    pr.s.set_mixed_precision(bits)
    e = pr.s.internal_vector("e")
    nl = len(case["levels"])
    # the same (rounded) starting error on both sides
    for ln in range(nl):
        for p in range(len(case["levels"][ln])):
            v = ROUND[bits](data["u"][ln][p])
            e.put(ln, p, v)
            pr.o.put("u", ln, p, v)
    for ln in range(nl - 1, 0, -1):
        print("lp smooth", case["name"], spec.get("dvar"), cf, bits, ln, flush=True)
        pr.s.smooth(e, pr.f, ln, 2)
        pr.o.smooth("u", "f", ln, 2)
        scale = _interior_max(case, ln, lambda p: pr.o.get("u", ln, p))
        diff = 0.0
        for p, box in enumerate(case["levels"][ln]):
            shp = hc.cell_shape(box, 1, case["dim"])
            inner = (slice(1, -1),) * case["dim"]
            g = e.get(ln, p).reshape(shp, order="F")[inner]
            o = pr.o.get("u", ln, p).reshape(shp, order="F")[inner]
            diff = max(diff, float(np.abs(g - o).max()))
        print("   max |e_lp - e_64| / max|e| = %.2e" % (diff / scale), flush=True)
        assert diff <= SMOOTH_TOL[bits] * scale
        # carry on from identical values
        for p in range(len(case["levels"][ln])):
            v = ROUND[bits](pr.o.get("u", ln, p))
            e.put(ln, p, v)
            pr.o.put("u", ln, p, v)


def _c2like():
    return c2_like_case(3)


_c2like.__name__ = "c2like"


CONST = dict(dvar=False)
VAR_ROBIN = dict(dvar=True, cmode=2, robin=True)
# (the reduced C2 deck with the Robin data stalls near 1e-12 absolute in FP64
# too, short of its 1e-10 relative target, so it runs with its own Dirichlet
# data only)
SOLVES = [(hc.case_3d, CONST), (hc.case_3d, VAR_ROBIN), (hc.case_3d_boundary, CONST),
          (hc.case_3d_boundary, VAR_ROBIN), (hc.case_2d_c0, CONST), (hc.case_2d_c0, VAR_ROBIN),
          (_c2like, CONST)]


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("mk,spec", SOLVES,
                         ids=["%s-%s" % (m.__name__, "var" if s.get("dvar") else "const")
                              for m, s in SOLVES])
def test_mixed_solve_reaches_the_fp64_solution(gpu_lib, mk, spec, bits):
    case, data, spec = _case(mk, spec)
    pr = hc.Pair(case, PARAMS, data, **spec)
    pr.s.set_mixed_precision(bits)
    print("mixed solve", case["name"], spec.get("dvar"), bits, flush=True)
    conv = pr.s.solve_system(pr.u, pr.f)
    hist = list(pr.s.history())
    conv_o, it_o, _, hist_o = pr.o.solve()
    assert conv and conv_o
    assert hist[-1] <= 1e-10 * pr.o.rms_f() * (1 + 1e-12)
    scale = max(_interior_max(case, ln, lambda p: pr.o.get("u", ln, p))
                for ln in range(len(case["levels"])))
    diffs = []
    for ln, boxes in enumerate(case["levels"]):
        m = 0.0
        for p, box in enumerate(boxes):
            shp = hc.cell_shape(box, 1, case["dim"])
            inner = (slice(1, -1),) * case["dim"]
            g = pr.u.get(ln, p).reshape(shp, order="F")[inner]
            o = pr.o.get("u", ln, p).reshape(shp, order="F")[inner]
            m = max(m, float(np.abs(g - o).max()))
        diffs.append(m)
    print("  cycles %d (fp64 oracle %d), per-level max |u - u64| / max|u|:" %
          (pr.s.number_of_iterations(), it_o), ["%.2e" % (d / scale) for d in diffs], flush=True)
    assert max(diffs) <= 1e-8 * scale
    # the precision changes the arithmetic; 64 bits restores the FP64 bits
    pr64 = hc.Pair(case, PARAMS, data, **spec)
    pr64.s.solve_system(pr64.u, pr64.f)
    assert list(pr64.s.history()) != hist
    pr.s.set_mixed_precision(64)
    for ln, boxes in enumerate(case["levels"]):
        for p in range(len(boxes)):
            pr.u.put(ln, p, data["u"][ln][p])
    pr.s.solve_system(pr.u, pr.f)
    assert list(pr.s.history()) == list(pr64.s.history())


def test_mixed_precision_rejects_other_cycles_and_widths(gpu_lib):
    from samrai_amd import fac
    case = hc.case_3d()
    pr = hc.Pair(case, PARAMS, smooth_data(case, 3))
    with pytest.raises(fac.SolverError):
        pr.s.set_mixed_precision(8)
    pr.s.set_mixed_precision(32)
    with pytest.raises(fac.SolverError):
        pr.s.set_algorithm_choice("pernice")
    pr.s.set_mixed_precision(64)
    pr.s.set_algorithm_choice("pernice")
    with pytest.raises(fac.SolverError):
        pr.s.set_mixed_precision(16)


def _slope32(dm, dp, dxc):
    """The conservative refine's limited slope in FP32 (geom_m4cartopstuff.i:39-63)."""
    f = np.float32
    coef2 = f(0.5) * (dp + dm)
    bound = f(2.0) * min(abs(dp), abs(dm))
    if dm * dp > f(0.0):
        return f(np.copysign(min(abs(coef2), bound), coef2)) / dxc
    return f(0.0)


@pytest.mark.parametrize("bits", [32, 16])
def test_low_precision_prolongation_matches_fp32_emulation(gpu_lib, bits):
    """prolongErrorAndCorrect on the FP32 / BF16 error vector: every fine
    cell gains the conservative linear refine of its parent computed in FP32
    (geom_m4cartrefineops3d.i:118-167, in that expression order) and is
    stored back rounded, bit for bit against a float32 emulation from the
    coarse values (ghosts as the operation filled them) and the fine values
    before the call.  Ratio 2 and aligned boxes take the kernel's row-pair
    path; ghost cells of the fine level are untouched."""
    from samrai_amd import fac
    case = hc.case_3d()
    data = smooth_data(case, 4)
    pr = hc.Pair(case, dict(PARAMS, max_cycles=1), data)
    s = pr.s
    s.set_mixed_precision(bits)
    s.initialize_solver_state(pr.u, pr.f)
    e = s.internal_vector("e")
    rng = np.random.default_rng(40 + bits)
    levels = case["levels"]
    for ln in (0, 1):
        for p, box in enumerate(levels[ln]):
            e.put(ln, p, rng.standard_normal(e.size(ln, p)))
    before = [e.get(1, p) for p in range(len(levels[1]))]
    print("lp prolongation", bits, flush=True)
    s.prolong_and_correct(e, 1)
    s.synchronize()
    coarse = [e.get(0, q) for q in range(len(levels[0]))]
    rnd = (lambda x: np.float32(x)) if bits == 32 else \
        (lambda x: np.float32(_bf16(np.array([x], np.float64))[0]))
    f32 = np.float32
    cdx = [f32(1.0 / 8)] * 3
    fdx = [f32(1.0 / 16)] * 3

    def U(c):
        for q, (lo, hi) in enumerate(levels[0]):
            if all(lo[d] <= c[d] <= hi[d] for d in range(3)):
                break
        else:
            raise AssertionError(c)
        lo, hi = levels[0][q]
        shp = hc.cell_shape((lo, hi), 1, 3)
        arr = coarse[q].reshape(shp, order="F")
        return lambda a, b, cc: f32(arr[a - lo[0] + 1, b - lo[1] + 1, cc - lo[2] + 1])

    checked = 0
    for p, (lo, hi) in enumerate(levels[1]):
        shp = hc.cell_shape((lo, hi), 1, 3)
        got = e.get(1, p).reshape(shp, order="F")
        old = before[p].reshape(shp, order="F")
        exp = old.copy()
        for k in range(lo[2], hi[2] + 1):
            for j in range(lo[1], hi[1] + 1):
                for i in range(lo[0], hi[0] + 1):
                    c = (i // 2, j // 2, k // 2)
                    V = U(c)
                    uc = V(*c)
                    s0 = _slope32(uc - V(c[0] - 1, c[1], c[2]), V(c[0] + 1, c[1], c[2]) - uc, cdx[0])
                    s1 = _slope32(uc - V(c[0], c[1] - 1, c[2]), V(c[0], c[1] + 1, c[2]) - uc, cdx[1])
                    s2 = _slope32(uc - V(c[0], c[1], c[2] - 1), V(c[0], c[1], c[2] + 1) - uc, cdx[2])
                    d = [(f32(x - 2 * cx) + f32(0.5)) * fdx[a] - cdx[a] * f32(0.5)
                         for a, (x, cx) in enumerate(zip((i, j, k), c))]
                    v = uc + s0 * d[0] + s1 * d[1] + s2 * d[2]
                    ii, jj, kk = i - lo[0] + 1, j - lo[1] + 1, k - lo[2] + 1
                    exp[ii, jj, kk] = rnd(f32(old[ii, jj, kk]) + v)
                    checked += 1
        np.testing.assert_array_equal(got, exp)
    assert checked == sum(np.prod([h - l + 1 for l, h in zip(*b)]) for b in levels[1])
