"""Small static hierarchies and paired (GPU library, oracle) set-ups for the
level-operator and solve parity tests."""
import numpy as np

from oracle.oracle import OracleFAC


def case_2d():
    L0 = [((0, 0), (7, 15)), ((8, 0), (15, 15))]
    L1 = [((8, 8), (15, 23)), ((16, 8), (23, 15))]
    L2 = [((20, 20), (27, 35))]
    return dict(name="2d3lev", dim=2, levels=[L0, L1, L2], ratio=2, dlo=(0, 0), dhi=(15, 15))


def case_3d():
    L0 = [((0, 0, 0), (3, 7, 7)), ((4, 0, 0), (7, 7, 7))]
    L1 = [((4, 4, 4), (9, 11, 11)), ((10, 4, 4), (13, 7, 11))]
    L2 = [((10, 10, 10), (17, 19, 21))]
    return dict(name="3d3lev", dim=3, levels=[L0, L1, L2], ratio=2, dlo=(0, 0, 0), dhi=(7, 7, 7))


def case_3d_boundary():
    """Fine patches touching the physical boundary on several faces."""
    L0 = [((0, 0, 0), (7, 7, 3)), ((0, 0, 4), (7, 7, 7))]
    L1 = [((0, 0, 0), (5, 7, 15)), ((6, 0, 0), (9, 3, 7))]
    return dict(name="3dbdry", dim=3, levels=[L0, L1], ratio=2, dlo=(0, 0, 0), dhi=(7, 7, 7))


def case_2d_c0(n=32):
    """The 2-D deck: n^2 base, one level at ratio 2 over [0.25, 0.75]^2."""
    L0 = [((0, 0), (n - 1, n - 1))]
    L1 = [((n // 2, n // 2), (3 * n // 2 - 1, 3 * n // 2 - 1))]
    return dict(name="2dc0", dim=2, levels=[L0, L1], ratio=2, dlo=(0, 0), dhi=(n - 1, n - 1))


def case_3d_wide0():
    """Level 0 of 40 x 36 x 44 cells in four patches (several tiles and
    z-segments of every fused tile shape, ragged edges), one finer level."""
    L0 = [((0, 0, 0), (19, 35, 21)), ((20, 0, 0), (39, 35, 21)),
          ((0, 0, 22), (23, 35, 43)), ((24, 0, 22), (39, 35, 43))]
    L1 = [((20, 20, 20), (43, 39, 47)), ((44, 20, 20), (59, 39, 35))]
    return dict(name="3dwide0", dim=3, levels=[L0, L1], ratio=2, dlo=(0, 0, 0),
                dhi=(39, 35, 43))


def case_2d_odd_split():
    """Level 0 split at an odd x: the 2-D patch-relative colourings of the
    two patches differ, so the coarsest level cannot be swept as one box."""
    L0 = [((0, 0), (6, 15)), ((7, 0), (15, 15))]
    L1 = [((8, 8), (15, 23))]
    return dict(name="2doddsplit", dim=2, levels=[L0, L1], ratio=2, dlo=(0, 0), dhi=(15, 15))


def case_2d_ratio4():
    """Refinement ratio 4, a fine patch two coarse cells wide (8 fine cells)
    next to a wider one."""
    L0 = [((0, 0), (3, 7)), ((4, 0), (7, 7))]
    L1 = [((8, 8), (19, 23)), ((20, 8), (27, 15))]
    return dict(name="2dratio4", dim=2, levels=[L0, L1], ratio=4, dlo=(0, 0), dhi=(7, 7))


def case_3d_ratio4():
    """3-D at refinement ratio 4 over a patch split level 0."""
    L0 = [((0, 0, 0), (3, 7, 7)), ((4, 0, 0), (7, 7, 7))]
    L1 = [((8, 8, 8), (23, 19, 27))]
    return dict(name="3dratio4", dim=3, levels=[L0, L1], ratio=4, dlo=(0, 0, 0), dhi=(7, 7, 7))


def case_2d_thin():
    """A fine patch one coarse cell wide and a fine patch touching two
    physical boundaries (C/F and physical ghosts at the same corner)."""
    L0 = [((0, 0), (15, 15))]
    L1 = [((10, 8), (11, 23)), ((0, 0), (7, 5))]
    return dict(name="2dthin", dim=2, levels=[L0, L1], ratio=2, dlo=(0, 0), dhi=(15, 15))


ALL = [case_2d, case_3d, case_3d_boundary, case_2d_ratio4, case_3d_ratio4, case_2d_thin]


def cell_shape(box, g, dim, axis=-1):
    lo, hi = box
    return tuple(hi[d] - lo[d] + 1 + 2 * g + (1 if d == axis else 0) for d in range(dim))


def face_centres(box, dim, axis, dx):
    lo, hi = box
    axes = []
    for d in range(dim):
        if d == axis:
            axes.append(np.arange(lo[d], hi[d] + 2) * dx[d])
        else:
            axes.append((np.arange(lo[d], hi[d] + 1) + 0.5) * dx[d])
    return np.meshgrid(*axes, indexing="ij")


def level_dx(case, ln):
    n0 = np.array(case["dhi"]) - np.array(case["dlo"]) + 1
    return (1.0 / n0) / (case["ratio"] ** ln)


def make_data(case, seed=0, dvar=False, cmode=0):
    """Random u (with ghosts), f, C and a smooth positive D (face-centre
    function, so shared faces agree between patches)."""
    rng = np.random.default_rng(seed)
    dim = case["dim"]
    data = {"u": [], "f": [], "C": [], "D": []}
    for ln, boxes in enumerate(case["levels"]):
        dx = level_dx(case, ln)
        data["u"].append([rng.standard_normal(cell_shape(b, 1, dim)) for b in boxes])
        data["f"].append([rng.standard_normal(cell_shape(b, 0, dim)) for b in boxes])
        data["C"].append([-rng.uniform(0.1, 1.0, cell_shape(b, 0, dim)) for b in boxes])
        Ds = []
        for b in boxes:
            per_axis = []
            for a in range(dim):
                X = face_centres(b, dim, a, dx)
                per_axis.append(np.exp(0.5 * np.sin(3 * X[0] + 1) * np.cos(2 * X[1] - 0.5)
                                       + (0.3 * X[2] if dim == 3 else 0.0)))
            Ds.append(per_axis)
        data["D"].append(Ds)
    return data


BC_DIRICHLET = dict(a=[1.0] * 6, b=[0.0] * 6, g=[0.0] * 6)


def bc_robin(dim):
    a = [1.0, 0.5, 1.0, 2.0, 1.0, 0.7][: 2 * dim]
    b = [0.0, 0.25, 0.5, 0.0, 0.1, 0.0][: 2 * dim]
    g = [0.3, -0.2, 0.1, 0.4, -0.5, 0.2][: 2 * dim]
    return dict(a=a, b=b, g=g)


class Pair:
    """The GPU solver and the oracle on one hierarchy with identical data."""

    def __init__(self, case, params, data, dvar=False, dconst=1.0, cmode=0, cconst=0.0,
                 bc=None, algorithm=None):
        from samrai_amd import fac
        dim = case["dim"]
        self.case = case
        self.dim = dim
        bc = bc or BC_DIRICHLET
        # GPU
        h = fac.PatchHierarchy(dim, case["dlo"], case["dhi"], [0.0] * dim, [1.0] * dim)
        for ln, boxes in enumerate(case["levels"]):
            h.add_level(case["ratio"], boxes)
        h.finalize()
        self.h = h
        self.u = fac.PatchData(h, 1)
        self.f = fac.PatchData(h, 0)
        self.r = fac.PatchData(h, 0)
        self.s = fac.CellPoissonFACSolver(h, **params)
        if dvar:
            self.D = fac.PatchData(h, 0, "side")
            for ln, boxes in enumerate(case["levels"]):
                for p in range(len(boxes)):
                    for a in range(dim):
                        self.D.put(ln, p, data["D"][ln][p][a], axis=a)
            self.s.set_d_patch_data(self.D)
        else:
            self.s.set_d_constant(dconst)
        if cmode == 1:
            self.s.set_c_constant(cconst)
        elif cmode == 2:
            self.C = fac.PatchData(h, 0)
            for ln, boxes in enumerate(case["levels"]):
                for p in range(len(boxes)):
                    self.C.put(ln, p, data["C"][ln][p])
            self.s.set_c_patch_data(self.C)
        self.s.set_boundaries(bc["a"][: 2 * dim], bc["b"][: 2 * dim], bc["g"][: 2 * dim])
        for ln, boxes in enumerate(case["levels"]):
            for p in range(len(boxes)):
                self.u.put(ln, p, data["u"][ln][p])
                self.f.put(ln, p, data["f"][ln][p])
        if algorithm:
            self.s.set_algorithm_choice(algorithm)
        self.s.initialize_solver_state(self.u, self.f)
        # oracle
        o = OracleFAC(dim, case["levels"], case["ratio"], case["dlo"], case["dhi"], [0.0] * dim,
                      [1.0] * dim)
        o.set_poisson(dvar, dconst, cmode, cconst)
        o.set_bc(list(bc["a"]) + [0] * 6, list(bc["b"]) + [0] * 6, list(bc["g"]) + [0] * 6)
        o.set_params(max_cycles=params.get("max_cycles", 10),
                     residual_tol=params.get("residual_tol", 1e-6),
                     relative_residual_tol=params.get("relative_residual_tol", -1.0),
                     pre=params.get("num_pre_sweeps", 1), post=params.get("num_post_sweeps", 1),
                     coarse_max_iterations=params.get("coarse_solver_max_iterations", 10),
                     coarse_tol=params.get("coarse_solver_tolerance", 1e-8), ewing=True)
        if params.get("coarse_solver_choice") in ("mg", "hypre"):
            o.set_coarse_mg(True)
        # the GPU solver's default prolongation is the reference's (CONSTANT_REFINE)
        o.set_prolongation(params.get("prolongation_method", "CONSTANT_REFINE"))
        o.set_cf_discretization(params.get("cf_discretization", "Ewing"))
        if algorithm:
            o.set_algorithm(algorithm)
        for ln, boxes in enumerate(case["levels"]):
            for p in range(len(boxes)):
                o.put("u", ln, p, data["u"][ln][p])
                o.put("f", ln, p, data["f"][ln][p])
                if dvar:
                    for a in range(dim):
                        o.put("D", ln, p, data["D"][ln][p][a], axis=a)
                if cmode == 2:
                    o.put("C", ln, p, data["C"][ln][p])
        self.o = o

    def npatch(self, ln):
        return len(self.case["levels"][ln])

    def compare(self, gpu_data, ofield, levels=None, interior_only=False, rtol=0.0):
        """Max abs difference between a GPU vector and an oracle field (and
        assert it is within rtol of the field's magnitude)."""
        worst = 0.0
        levels = range(len(self.case["levels"])) if levels is None else levels
        for ln in levels:
            for p in range(self.npatch(ln)):
                g = gpu_data.get(ln, p)
                o = self.o.get(ofield, ln, p)
                assert g.shape == o.shape
                if interior_only and ofield in ("u", "e"):
                    box = self.case["levels"][ln][p]
                    shp = cell_shape(box, 1, self.dim)
                    sl = tuple(slice(1, -1) for _ in range(self.dim))
                    g = g.reshape(shp, order="F")[sl]
                    o = o.reshape(shp, order="F")[sl]
                d = float(np.max(np.abs(g - o))) if g.size else 0.0
                scale = float(np.max(np.abs(o))) if o.size else 0.0
                assert d <= rtol * scale, (ofield, ln, p, d, scale)
                worst = max(worst, d)
        return worst
