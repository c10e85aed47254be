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


FIXED = dict(PARAMS, max_cycles=4, relative_residual_tol=1e-300, coarse_solver_max_iterations=60)


@pytest.mark.parametrize("spec", [dict(dvar=False), dict(dvar=True, cmode=2, robin=True)],
                         ids=["constD", "varD-varC-robin"])
@pytest.mark.parametrize("mk", hc.ALL + [hc.case_2d_c0, hc.case_2d_odd_split, hc.case_3d_wide0],
                         ids=lambda f: f.__name__)
def test_fused_coarsest_sweeps_are_bit_identical(gpu_lib, mk, spec):
    """The colour-fused coarsest-level sweeps (every tile shape), with and
    without whole-cycle HIP graphs, give the same bits as the patch-wise red
    pass / fill / black pass / fill launched eagerly.  A graph is replayed
    once per cycle, with the host's norm read-back between replays."""
    from samrai_amd import fac
    case = mk()
    spec = dict(spec)
    if spec.pop("robin", False):
        spec["bc"] = hc.bc_robin(case["dim"])
    data = hc.make_data(case, seed=7)
    shapes = fac.fused_tile_shapes(case["dim"])
    ref = None
    runs = [(False, 0, False), (False, 0, True), (True, 0, False)]
    runs += [(True, v, True) for v in range(len(shapes))]
    for on, v, graphs in runs:
        print("fused", case["name"], on, v, graphs, flush=True)
        pr = hc.Pair(case, FIXED, data, **spec)
        pr.s.set_fused_coarsest(on, v)
        pr.s.set_graphs(graphs)
        pr.s.solve_system(pr.u, pr.f)
        assert pr.s.fused_coarsest_active() == (on and case["name"] != "2doddsplit")
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


MG = dict(PARAMS, coarse_solver_choice="mg", coarse_solver_tolerance=1e-10,
          coarse_solver_max_iterations=20)


@pytest.mark.parametrize("spec", [dict(dvar=False), dict(dvar=True, cmode=2, robin=True)],
                         ids=["constD", "varD-varC-robin"])
@pytest.mark.parametrize("mk", hc.ALL + [hc.case_2d_c0], ids=lambda f: f.__name__)
def test_multigrid_coarse_solve_matches_oracle(gpu_lib, mk, spec):
    """coarse_solver_choice "mg": GPU and oracle agree cycle by cycle (the
    norms differ only in summation order), take the same number of V-cycles
    in the last coarse solve, and converge to the red-black path's solution
    within the solver tolerance."""
    case = mk()
    spec = dict(spec)
    if spec.pop("robin", False):
        spec["bc"] = hc.bc_robin(case["dim"])
    data = smooth_data(case, 3)
    pr = hc.Pair(case, MG, data, **spec)
    print("mg solve", case["name"], spec, flush=True)
    conv_g = pr.s.solve_system(pr.u, pr.f)
    hist_g = pr.s.history()
    conv_o, it_o, res_o, hist_o = pr.o.solve()
    assert conv_g and conv_o
    assert pr.s.number_of_iterations() == it_o
    np.testing.assert_allclose(hist_g, hist_o, rtol=1e-10, atol=0)
    assert pr.s.coarse_mg_stats()[0] == pr.o.mg_stats()[0]
    pr.compare(pr.u, "u", interior_only=True, rtol=1e-10)
    # same fixed point as the red-black coarse solver
    rb = hc.Pair(case, PARAMS, data, **spec)
    rb.s.solve_system(rb.u, rb.f)
    for ln in range(len(case["levels"])):
        for p in range(len(case["levels"][ln])):
            a, b = pr.u.get(ln, p), rb.u.get(ln, p)
            assert np.abs(a - b).max() <= 1e-7 * max(np.abs(b).max(), 1e-300)


def c2_like_case(seed):
    """A reduced C2 deck: base 32^3, three levels at ratio 2, patch sides
    from U{8..16}, D = exp(GRF), f = Gaussian sources (samrai_amd.decks)."""
    from samrai_amd import decks
    deck = decks.deck_3d_c2(seed, base=32, nlevels=3, smin=8, smax=16)
    case = dict(name="c2like-s%d" % seed, dim=3, levels=deck["levels"], ratio=2,
                dlo=deck["dlo"], dhi=deck["dhi"])
    modes = decks.grf_modes(seed)
    data = {"u": [], "f": [], "C": [], "D": []}
    for ln, boxes in enumerate(deck["levels"]):
        data["u"].append([np.zeros(hc.cell_shape(b, 1, 3)) for b in boxes])
        data["f"].append([decks.gaussian_sources(deck, decks.cell_points(deck, ln, b))
                          for b in boxes])
        data["C"].append([np.zeros(hc.cell_shape(b, 0, 3)) for b in boxes])
        data["D"].append([[np.exp(decks.grf_eval(decks.face_points(deck, ln, b, a), modes))
                           for a in range(3)] for b in boxes])
    return case, data


@pytest.mark.parametrize("choice", ["redblack", "mg"])
def test_c2_like_deck_matches_oracle(gpu_lib, choice):
    """Variable coefficient D = exp(GRF), uneven patches, 3 levels: a fixed
    number of FAC cycles agrees with the oracle cycle by cycle."""
    case, data = c2_like_case(1)
    P = dict(PARAMS, max_cycles=6, relative_residual_tol=1e-300,
             coarse_solver_choice=choice)
    if choice == "mg":
        P.update(coarse_solver_tolerance=1e-8, coarse_solver_max_iterations=10)
    pr = hc.Pair(case, P, data, dvar=True)
    print("c2like", choice, [len(l) for l in case["levels"]], flush=True)
    pr.s.solve_system(pr.u, pr.f)
    conv_o, it_o, res_o, hist_o = pr.o.solve()
    np.testing.assert_allclose(pr.s.history(), hist_o, rtol=1e-10, atol=0)
    pr.compare(pr.u, "u", interior_only=True, rtol=1e-10)


@pytest.mark.parametrize("method", ["LINEAR_REFINE", "CONSTANT_REFINE"])
@pytest.mark.parametrize("mk", [hc.case_2d, hc.case_3d], ids=lambda f: f.__name__)
def test_solve_with_other_prolongations_matches_oracle(gpu_lib, mk, method):
    """prolongation_method LINEAR_REFINE and CONSTANT_REFINE (the reference's
    default) through whole solves, cycle by cycle against the oracle."""
    case = mk()
    data = smooth_data(case, 3)
    pr = hc.Pair(case, dict(PARAMS, prolongation_method=method), data)
    print("solve", case["name"], method, flush=True)
    conv_g = pr.s.solve_system(pr.u, pr.f)
    conv_o, it_o, res_o, hist_o = pr.o.solve()
    assert conv_g == conv_o
    assert pr.s.number_of_iterations() == it_o
    np.testing.assert_allclose(pr.s.history(), hist_o, rtol=1e-12, atol=0)
    pr.compare(pr.u, "u", interior_only=True, rtol=1e-12)


@pytest.mark.parametrize("cf", ["CONSTANT_REFINE", "LINEAR_REFINE", "CONSERVATIVE_LINEAR_REFINE"])
@pytest.mark.parametrize("mk", hc.ALL, ids=lambda f: f.__name__)
def test_solve_with_other_cf_discretizations_matches_oracle(gpu_lib, mk, cf):
    """cf_discretization by a refine operator (C/F ghosts refined from the
    coarser level, no Ewing flux fix) through whole solves, cycle by cycle
    against the oracle, with variable D."""
    case = mk()
    data = smooth_data(case, 5)
    pr = hc.Pair(case, dict(PARAMS, cf_discretization=cf), data, dvar=True)
    print("solve", case["name"], cf, flush=True)
    conv_g = pr.s.solve_system(pr.u, pr.f)
    conv_o, it_o, res_o, hist_o = pr.o.solve()
    assert conv_g == conv_o
    assert pr.s.number_of_iterations() == it_o
    np.testing.assert_allclose(pr.s.history(), hist_o, rtol=1e-12, atol=0)
    pr.compare(pr.u, "u", interior_only=True, rtol=1e-12)


@pytest.mark.parametrize("graphs", [False, True], ids=["eager", "graphs"])
@pytest.mark.parametrize("algorithm", ["mccormick-s4.3", "pernice", "MCCORMICK-S4.3"])
@pytest.mark.parametrize("mk", hc.ALL + [hc.case_2d_c0], ids=lambda f: f.__name__)
def test_solve_with_other_cycles_matches_oracle(gpu_lib, mk, algorithm, graphs):
    """FACPreconditioner algorithm_choice "mccormick-s4.3" and "pernice"
    (case-insensitive) through whole solves, cycle by cycle against the
    oracle, eagerly and as replayed whole-cycle graphs."""
    case = mk()
    data = smooth_data(case, 3)
    pr = hc.Pair(case, PARAMS, data, algorithm=algorithm.lower())
    if algorithm != algorithm.lower():
        pr.s.set_algorithm_choice(algorithm)
    pr.s.set_graphs(graphs)
    print("solve", case["name"], algorithm, graphs, flush=True)
    conv_g = pr.s.solve_system(pr.u, pr.f)
    conv_o, it_o, res_o, hist_o = pr.o.solve()
    assert conv_g == conv_o
    assert pr.s.number_of_iterations() == it_o
    np.testing.assert_allclose(pr.s.history(), hist_o, rtol=1e-12, atol=0)
    pr.compare(pr.u, "u", interior_only=True, rtol=1e-12)
