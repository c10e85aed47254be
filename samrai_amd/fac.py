"""Python mirror of the FAC Poisson solver interface over libsamrai_fac.so.

Names follow the reference (hier::PatchHierarchy, pdat::CellData/SideData,
solv::CellPoissonFACSolver and the FACOperatorStrategy operations of
solv::CellPoissonFACOps); every call goes through the C ABI declared in
include/samrai_fac.h.  There is no CPU fallback: if the HIP library is not
built, loading fails with an error.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsamrai_fac.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

_lib = None


class SolverError(RuntimeError):
    """A non-zero status from the library (the TBOX_ERROR analogue)."""


class _Params(C.Structure):
    _fields_ = [("max_cycles", C.c_int), ("residual_tol", C.c_double),
                ("relative_residual_tol", C.c_double), ("num_pre_sweeps", C.c_int),
                ("num_post_sweeps", C.c_int), ("coarse_solver_choice", C.c_char_p),
                ("coarse_solver_tolerance", C.c_double),
                ("coarse_solver_max_iterations", C.c_int), ("cf_discretization", C.c_char_p),
                ("prolongation_method", C.c_char_p), ("enable_logging", C.c_int)]


def lib():
    """Load libsamrai_fac.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsamrai_fac.so is not built: run python -m samrai_amd.build")
    L = C.CDLL(LIB_PATH)
    L.sfac_last_error.restype = C.c_char_p
    L.sfac_hierarchy_create.argtypes = [C.c_int, _ip, _ip, _dp, _dp, C.POINTER(_vp)]
    L.sfac_hierarchy_add_level.argtypes = [_vp, C.c_int, C.c_int, _ip]
    L.sfac_hierarchy_finalize.argtypes = [_vp]
    L.sfac_hierarchy_num_patches.argtypes = [_vp, C.c_int]
    L.sfac_hierarchy_destroy.argtypes = [_vp]
    L.sfac_cell_data_create.argtypes = [_vp, C.c_int, C.POINTER(_vp)]
    L.sfac_side_data_create.argtypes = [_vp, C.c_int, C.POINTER(_vp)]
    L.sfac_cell_data_wrap.argtypes = [_vp, C.c_int, C.POINTER(C.c_void_p), C.c_long,
                                      C.POINTER(_vp)]
    L.sfac_side_data_wrap.argtypes = [_vp, C.c_int, C.POINTER(C.c_void_p), C.c_long,
                                      C.POINTER(_vp)]
    L.sfac_patch_data_destroy.argtypes = [_vp]
    L.sfac_patch_data_pointer.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(_dp),
                                          C.POINTER(C.c_long)]
    L.sfac_patch_data_put.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp]
    L.sfac_patch_data_get.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp]
    L.sfac_patch_data_put_device.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp]
    L.sfac_patch_data_fill.argtypes = [_vp, C.c_double]
    L.sfac_solver_params_default.argtypes = [C.POINTER(_Params)]
    L.sfac_solver_params_default.restype = None
    L.sfac_solver_create.argtypes = [_vp, C.POINTER(_Params), C.POINTER(_vp)]
    L.sfac_solver_destroy.argtypes = [_vp]
    L.sfac_solver_set_d_constant.argtypes = [_vp, C.c_double]
    L.sfac_solver_set_d_patch_data.argtypes = [_vp, _vp]
    L.sfac_solver_set_c_constant.argtypes = [_vp, C.c_double]
    L.sfac_solver_set_c_patch_data.argtypes = [_vp, _vp]
    L.sfac_solver_set_boundaries.argtypes = [_vp, _dp, _dp, _dp]
    L.sfac_solver_initialize.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int]
    L.sfac_solver_deallocate.argtypes = [_vp]
    L.sfac_solver_solve.argtypes = [_vp, _vp, _vp, _ip]
    L.sfac_solver_get_iterations.argtypes = [_vp, _ip]
    L.sfac_solver_get_residual_norm.argtypes = [_vp, _dp]
    L.sfac_solver_get_convergence_factors.argtypes = [_vp, _dp, _dp]
    L.sfac_solver_get_history.argtypes = [_vp, _dp, C.c_int, _ip]
    L.sfac_solver_last_solve_seconds.argtypes = [_vp, _dp]
    L.sfac_solver_synchronize.argtypes = [_vp]
    L.sfac_solver_enable_kernel_timing.argtypes = [_vp, C.c_int]
    L.sfac_solver_kernel_timing.argtypes = [_vp, C.c_int, C.POINTER(C.c_long), _dp, _dp]
    L.sfac_solver_set_fused_coarsest.argtypes = [_vp, C.c_int, C.c_int]
    L.sfac_solver_fused_coarsest_active.argtypes = [_vp, _ip]
    L.sfac_solver_fused_tile_variant.argtypes = [_vp, _ip]
    L.sfac_solver_fused_tile_timings.argtypes = [_vp, _dp, C.c_int, _ip]
    L.sfac_solver_set_fused_levels.argtypes = [_vp, C.c_int, C.c_int]
    L.sfac_solver_fused_level_variant.argtypes = [_vp, C.c_int, _ip]
    L.sfac_solver_coarse_mg_stats.argtypes = [_vp, _ip, _dp]
    L.sfac_solver_set_graphs.argtypes = [_vp, C.c_int]
    L.sfac_solver_set_algorithm_choice.argtypes = [_vp, C.c_char_p]
    L.sfac_solver_set_coarse_precision.argtypes = [_vp, C.c_int]
    L.sfac_solver_set_mixed_precision.argtypes = [_vp, C.c_int]
    L.sfac_berger_rigoutsos.argtypes = [C.c_int, _ip, _ip, C.c_void_p, C.c_double, C.c_double,
                                        _ip, _ip, _ip, C.c_int, _ip]
    L.sfac_fused_tile_variants.argtypes = [C.c_int]
    L.sfac_fused_tile_shape.argtypes = [C.c_int, C.c_int, _ip]
    L.sfac_ops_internal_vector.argtypes = [_vp, C.c_int, C.POINTER(_vp)]
    L.sfac_ops_ghost_fill.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int]
    L.sfac_ops_smooth.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_double, _ip, _dp]
    L.sfac_ops_residual.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int]
    L.sfac_ops_restrict_solution.argtypes = [_vp, _vp, C.c_int]
    L.sfac_ops_restrict_residual.argtypes = [_vp, _vp, _vp, C.c_int]
    L.sfac_ops_prolong_and_correct.argtypes = [_vp, _vp, C.c_int]
    L.sfac_ops_rms_norm.argtypes = [_vp, _vp, _dp]
    _lib = L
    return L


def _check(status):
    if status != 0:
        raise SolverError(lib().sfac_last_error().decode())


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class PatchHierarchy:
    """hier::PatchHierarchy over a CartesianGridGeometry (static levels)."""

    def __init__(self, dim, domain_lo, domain_hi, xlo, xhi):
        self.dim = dim
        self.levels = []
        self.domain_lo = tuple(domain_lo)
        self.domain_hi = tuple(domain_hi)
        self.xlo = tuple(xlo)
        self.xhi = tuple(xhi)
        h = _vp()
        dlo = _arr(domain_lo, np.int32)
        dhi = _arr(domain_hi, np.int32)
        xl = _arr(xlo, np.float64)
        xh = _arr(xhi, np.float64)
        _check(lib().sfac_hierarchy_create(dim, dlo.ctypes.data_as(_ip), dhi.ctypes.data_as(_ip),
                                           xl.ctypes.data_as(_dp), xh.ctypes.data_as(_dp),
                                           C.byref(h)))
        self.h = h
        self.ratios = []

    def add_level(self, ratio, boxes):
        a = _arr([list(lo) + list(hi) for lo, hi in boxes], np.int32).ravel()
        _check(lib().sfac_hierarchy_add_level(self.h, ratio, len(boxes), a.ctypes.data_as(_ip)))
        self.levels.append([(tuple(lo), tuple(hi)) for lo, hi in boxes])
        self.ratios.append(ratio if self.levels[:-1] else 1)

    def finalize(self):
        _check(lib().sfac_hierarchy_finalize(self.h))

    def num_levels(self):
        return len(self.levels)

    def close(self):
        if self.h:
            lib().sfac_hierarchy_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PatchData:
    """pdat::CellData<double> (centring 'cell') or SideData<double> ('side')
    on every patch, in device memory."""

    def __init__(self, hierarchy, ghost, centring="cell", _handle=None):
        self.hierarchy = hierarchy
        self.ghost = ghost
        self.centring = centring
        self.owned = _handle is None
        if _handle is not None:
            self.d = _handle
            return
        d = _vp()
        fn = lib().sfac_cell_data_create if centring == "cell" else lib().sfac_side_data_create
        _check(fn(hierarchy.h, ghost, C.byref(d)))
        self.d = d

    @classmethod
    def wrap(cls, hierarchy, ghost, arrays, centring="cell"):
        """Patch data over device arrays the caller owns, bound without
        copying (sfac_cell_data_wrap / sfac_side_data_wrap): `arrays` holds
        one contiguous float64 device tensor per patch array (per patch, per
        axis for side data), levels in order, each the whole ghost box with i
        fastest.  The tensors are kept referenced by the returned object."""
        for t in arrays:
            if t.element_size() != 8 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("wrap needs contiguous float64 device tensors")
        ptrs = (C.c_void_p * max(1, len(arrays)))(*[t.data_ptr() for t in arrays])
        d = _vp()
        fn = lib().sfac_cell_data_wrap if centring == "cell" else lib().sfac_side_data_wrap
        _check(fn(hierarchy.h, ghost, ptrs, len(arrays), C.byref(d)))
        self = cls(hierarchy, ghost, centring, _handle=d)
        self.owned = True
        self.arrays = list(arrays)
        return self

    def pointer(self, ln, p, axis=0):
        ptr = _dp()
        n = C.c_long()
        _check(lib().sfac_patch_data_pointer(self.d, ln, p, axis, C.byref(ptr), C.byref(n)))
        return C.cast(ptr, C.c_void_p).value, n.value

    def size(self, ln, p, axis=0):
        return self.pointer(ln, p, axis)[1]

    def put(self, ln, p, arr, axis=0):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel(order="F"))
        if a.size != self.size(ln, p, axis):
            raise ValueError("array size %d != patch array size %d" % (a.size, self.size(ln, p, axis)))
        _check(lib().sfac_patch_data_put(self.d, ln, p, axis, a.ctypes.data_as(_dp)))

    def get(self, ln, p, axis=0):
        a = np.empty(self.size(ln, p, axis), dtype=np.float64)
        _check(lib().sfac_patch_data_get(self.d, ln, p, axis, a.ctypes.data_as(_dp)))
        return a

    def put_device(self, ln, p, tensor, axis=0):
        """Copy a device tensor holding the whole ghost box of patch p
        (i fastest in memory, float64) into the patch array."""
        t = tensor
        if t.element_size() != 8 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("put_device needs a contiguous float64 device tensor")
        if t.numel() != self.size(ln, p, axis):
            raise ValueError("tensor size %d != patch array size %d"
                             % (t.numel(), self.size(ln, p, axis)))
        _check(lib().sfac_patch_data_put_device(self.d, ln, p, axis, C.c_void_p(t.data_ptr())))

    def fill(self, value):
        _check(lib().sfac_patch_data_fill(self.d, float(value)))

    def close(self):
        if self.owned and self.d:
            lib().sfac_patch_data_destroy(self.d)
        self.d = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DEFAULTS = dict(max_cycles=10, residual_tol=1.0e-6, relative_residual_tol=-1.0,
                num_pre_sweeps=1, num_post_sweeps=1, coarse_solver_choice="redblack",
                coarse_solver_tolerance=1.0e-8, coarse_solver_max_iterations=10,
                cf_discretization="Ewing", prolongation_method="CONSTANT_REFINE",
                enable_logging=False)


class CellPoissonFACSolver:
    """solv::CellPoissonFACSolver: solves div(D grad u) + C u = f with FAC.

    Keyword arguments are the reference's input keys (fac_precond and fac_ops
    blocks); unspecified keys take the reference defaults.
    """

    def __init__(self, hierarchy, **params):
        unknown = set(params) - set(DEFAULTS)
        if unknown:
            raise ValueError("unknown input keys: %s" % sorted(unknown))
        p = dict(DEFAULTS, **params)
        self._keep = [s.encode() for s in (p["coarse_solver_choice"], p["cf_discretization"],
                                            p["prolongation_method"])]
        cp = _Params(p["max_cycles"], p["residual_tol"], p["relative_residual_tol"],
                     p["num_pre_sweeps"], p["num_post_sweeps"], self._keep[0],
                     p["coarse_solver_tolerance"], p["coarse_solver_max_iterations"],
                     self._keep[1], self._keep[2], int(p["enable_logging"]))
        s = _vp()
        _check(lib().sfac_solver_create(hierarchy.h, C.byref(cp), C.byref(s)))
        self.s = s
        self.hierarchy = hierarchy
        self.params = p

    # PoissonSpecifications / boundary conditions
    def set_d_constant(self, d):
        _check(lib().sfac_solver_set_d_constant(self.s, float(d)))

    def set_d_patch_data(self, side_data):
        _check(lib().sfac_solver_set_d_patch_data(self.s, side_data.d))

    def set_c_constant(self, c):
        _check(lib().sfac_solver_set_c_constant(self.s, float(c)))

    def set_c_patch_data(self, cell_data):
        _check(lib().sfac_solver_set_c_patch_data(self.s, cell_data.d))

    def set_boundaries(self, a, b, g):
        """LocationIndexRobinBcCoefs: a u + b du/dn = g per location index."""
        n = 2 * self.hierarchy.dim
        aa = _arr(list(a) + [0] * (6 - n), np.float64)
        bb = _arr(list(b) + [0] * (6 - n), np.float64)
        gg = _arr(list(g) + [0] * (6 - n), np.float64)
        _check(lib().sfac_solver_set_boundaries(self.s, aa.ctypes.data_as(_dp),
                                                bb.ctypes.data_as(_dp), gg.ctypes.data_as(_dp)))

    def set_dirichlet(self, value=0.0):
        n = 2 * self.hierarchy.dim
        self.set_boundaries([1.0] * n, [0.0] * n, [value] * n)

    def initialize_solver_state(self, u, f, coarsest=0, finest=-1):
        _check(lib().sfac_solver_initialize(self.s, u.d, f.d, coarsest, finest))

    def deallocate_solver_state(self):
        _check(lib().sfac_solver_deallocate(self.s))

    def solve_system(self, u, f):
        conv = C.c_int(0)
        _check(lib().sfac_solver_solve(self.s, u.d, f.d, C.byref(conv)))
        return bool(conv.value)

    def number_of_iterations(self):
        n = C.c_int()
        _check(lib().sfac_solver_get_iterations(self.s, C.byref(n)))
        return n.value

    def residual_norm(self):
        r = C.c_double()
        _check(lib().sfac_solver_get_residual_norm(self.s, C.byref(r)))
        return r.value

    def convergence_factors(self):
        a = C.c_double()
        f = C.c_double()
        _check(lib().sfac_solver_get_convergence_factors(self.s, C.byref(a), C.byref(f)))
        return a.value, f.value

    def history(self):
        n = C.c_int()
        _check(lib().sfac_solver_get_history(self.s, None, 0, C.byref(n)))
        h = np.zeros(max(n.value, 1))
        _check(lib().sfac_solver_get_history(self.s, h.ctypes.data_as(_dp), n.value, C.byref(n)))
        return h[: n.value]

    def last_solve_seconds(self):
        t = C.c_double()
        _check(lib().sfac_solver_last_solve_seconds(self.s, C.byref(t)))
        return t.value

    def synchronize(self):
        _check(lib().sfac_solver_synchronize(self.s))

    def enable_kernel_timing(self, on=True):
        """HIP-event timing of the red/black colour launches (solver stream)."""
        _check(lib().sfac_solver_enable_kernel_timing(self.s, int(on)))

    def set_fused_coarsest(self, on=True, tile_variant=-1):
        """Colour-fused sweeps on a coarsest level that tiles the domain
        (one launch per sweep, same results); off: patch-wise colour passes.
        tile_variant -1: autotuned tile shape."""
        _check(lib().sfac_solver_set_fused_coarsest(self.s, int(on), int(tile_variant)))

    def set_fused_levels(self, on=True, tile_variant=-1):
        """Colour-fused sweeps on the levels above the coarsest (3-D; one
        launch per sweep over all patches, same results); off: patch-wise
        colour passes.  tile_variant -1: autotuned tile shape."""
        _check(lib().sfac_solver_set_fused_levels(self.s, int(on), int(tile_variant)))

    def fused_level_variant(self, ln):
        """Tile shape level ln's fused sweeps use, -1 if it runs colour passes."""
        v = C.c_int()
        _check(lib().sfac_solver_fused_level_variant(self.s, int(ln), C.byref(v)))
        return v.value

    def set_algorithm_choice(self, choice):
        """FACPreconditioner::setAlgorithmChoice: "default", "mccormick-s4.3"
        or "pernice"."""
        _check(lib().sfac_solver_set_algorithm_choice(self.s, choice.encode()))

    def set_coarse_precision(self, bits):
        """64 (default) or 32: precision of the coarsest-level solve of the
        error equation inside the FP64 FAC iteration."""
        _check(lib().sfac_solver_set_coarse_precision(self.s, int(bits)))

    def set_mixed_precision(self, bits):
        """64 (default), 32 or 16: storage of the FAC iteration's error vector
        (FP64, FP32 or BF16); smoothing, the error's ghost fills and the
        prolongation then compute in FP32, residuals and u += e in FP64."""
        _check(lib().sfac_solver_set_mixed_precision(self.s, int(bits)))

    def set_graphs(self, on=True):
        """Capture each FAC cycle as a HIP graph and replay it (same results)."""
        _check(lib().sfac_solver_set_graphs(self.s, int(on)))

    def fused_tile_variant(self):
        v = C.c_int()
        _check(lib().sfac_solver_fused_tile_variant(self.s, C.byref(v)))
        return v.value

    def fused_tile_timings(self):
        """Autotuner milliseconds per sweep for each tile shape."""
        buf = (C.c_double * 64)()
        n = C.c_int()
        _check(lib().sfac_solver_fused_tile_timings(self.s, buf, 64, C.byref(n)))
        return [buf[i] for i in range(min(n.value, 64))]

    def coarse_mg_stats(self):
        """(V-cycles, relative residual) of the last multigrid coarse solve."""
        it = C.c_int()
        rel = C.c_double()
        _check(lib().sfac_solver_coarse_mg_stats(self.s, C.byref(it), C.byref(rel)))
        return it.value, rel.value

    def fused_coarsest_active(self):
        a = C.c_int()
        _check(lib().sfac_solver_fused_coarsest_active(self.s, C.byref(a)))
        return bool(a.value)

    def kernel_timing(self, level=-1):
        """(launches, total kernel ms, algorithmic bytes) of the timed colour
        launches on `level` (all levels when level < 0)."""
        n = C.c_long()
        ms = C.c_double()
        b = C.c_double()
        _check(lib().sfac_solver_kernel_timing(self.s, level, C.byref(n), C.byref(ms),
                                               C.byref(b)))
        return n.value, ms.value, b.value

    # FACOperatorStrategy operations (after initialize_solver_state)
    def smoother_kernel_name(self, ln):
        """Name of the kernel that smoothErrorByRedBlack launches on level ln
        (for reports)."""
        if ln > 0 and self.fused_level_variant(ln) >= 0:
            return "k_level_rbgs (colour-fused sweep)"
        return "k_rbgs (one colour pass)"

    def internal_vector(self, which):
        idx = {"e": 0, "r": 1, "tr": 2, "w": 3}[which]
        d = _vp()
        _check(lib().sfac_ops_internal_vector(self.s, idx, C.byref(d)))
        ghost = None
        return PatchData(self.hierarchy, ghost, "cell", _handle=d)

    def ghost_fill(self, x, ln, with_coarse, homogeneous):
        _check(lib().sfac_ops_ghost_fill(self.s, x.d, ln, int(with_coarse), int(homogeneous)))

    def smooth(self, e, r, ln, sweeps, tol=-1.0):
        n = C.c_int()
        mr = C.c_double()
        _check(lib().sfac_ops_smooth(self.s, e.d, r.d, ln, sweeps, tol, C.byref(n), C.byref(mr)))
        return n.value, mr.value

    def residual(self, r, u, f, ln, homogeneous):
        _check(lib().sfac_ops_residual(self.s, r.d, u.d, f.d, ln, int(homogeneous)))

    def restrict_solution(self, u, dest_ln):
        _check(lib().sfac_ops_restrict_solution(self.s, u.d, dest_ln))

    def restrict_residual(self, src, dst, dest_ln):
        _check(lib().sfac_ops_restrict_residual(self.s, src.d, dst.d, dest_ln))

    def prolong_and_correct(self, e, dest_ln):
        _check(lib().sfac_ops_prolong_and_correct(self.s, e.d, dest_ln))

    def rms_norm(self, x):
        o = C.c_double()
        _check(lib().sfac_ops_rms_norm(self.s, x.d, C.byref(o)))
        return o.value

    def close(self):
        if self.s:
            lib().sfac_solver_destroy(self.s)
            self.s = None
        self._learned_net = None  # the solver no longer holds the network

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fused_tile_shapes(dim):
    """Tile shapes (tx, ty, tz) of the colour-fused coarsest-level sweep."""
    L = lib()
    out = []
    for v in range(L.sfac_fused_tile_variants(dim)):
        a = (C.c_int * 3)()
        _check(L.sfac_fused_tile_shape(dim, v, a))
        out.append(tuple(a))
    return out


def berger_rigoutsos(tags, lo=None, efficiency_tolerance=0.8, combine_efficiency=0.8,
                     min_box=None, max_box=None):
    """mesh::BergerRigoutsos clustering (one process) of a boolean tag array
    indexed [i, j(, k)] whose cell (0, 0(, 0)) has index lo: boxes
    ((lo...), (hi...)) covering every tagged cell, none larger than max_box,
    none cut below min_box (host function)."""
    t = np.asfortranarray(np.asarray(tags, dtype=np.uint8))
    dim = t.ndim
    lo = _arr(lo if lo is not None else (0,) * dim, np.int32)
    hi = _arr([l + n - 1 for l, n in zip(lo, t.shape)], np.int32)
    mn = _arr(min_box if min_box is not None else (1,) * dim, np.int32)
    mx = _arr(max_box if max_box is not None else (1 << 30,) * dim, np.int32)
    n = C.c_int()
    args = (dim, lo.ctypes.data_as(_ip), hi.ctypes.data_as(_ip), t.ctypes.data_as(C.c_void_p),
            efficiency_tolerance, combine_efficiency, mn.ctypes.data_as(_ip), mx.ctypes.data_as(_ip))
    _check(lib().sfac_berger_rigoutsos(*args, None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1) * 2 * dim, dtype=np.int32)
    _check(lib().sfac_berger_rigoutsos(*args, out.ctypes.data_as(_ip), n.value, C.byref(n)))
    b = out[: n.value * 2 * dim].reshape(n.value, 2, dim)
    return [(tuple(int(v) for v in x[0]), tuple(int(v) for v in x[1])) for x in b]
