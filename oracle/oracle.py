"""ctypes bindings for the C oracle (liboracle_fac.so) and, when it was
built, the reference's own Fortran kernels (_ref/libsamrai_ref_kernels.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the samrai_amd package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_fac.so")
REF_SO = os.path.join(HERE, "_ref", "libsamrai_ref_kernels.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def dptr(a):
    if a is None:
        return C.cast(None, _dp)
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library missing: run make -C oracle")
        lib = C.CDLL(ORACLE_SO)
        lib.orc_fac_create.restype = C.c_void_p
        lib.orc_fac_create.argtypes = [C.c_int, C.c_int, _ip, C.POINTER(_ip), _ip, _ip,
                                       _ip, _dp, _dp]
        lib.orc_fac_destroy.argtypes = [C.c_void_p]
        lib.orc_fac_set_poisson.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int,
                                            C.c_double]
        lib.orc_fac_set_bc.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.orc_fac_set_params.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                           C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                           C.c_int]
        lib.orc_fac_array_size.restype = C.c_long
        lib.orc_fac_array_size.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.orc_fac_put.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        lib.orc_fac_get.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        lib.orc_fac_solve.restype = C.c_int
        lib.orc_fac_solve.argtypes = [C.c_void_p, _ip, _dp, _dp, C.c_int]
        lib.orc_fac_full_composite_residual.restype = C.c_double
        lib.orc_fac_full_composite_residual.argtypes = [C.c_void_p]
        lib.orc_fac_rms_norm_f.restype = C.c_double
        lib.orc_fac_rms_norm_f.argtypes = [C.c_void_p]
        lib.orc_fac_last_solve_seconds.restype = C.c_double
        lib.orc_fac_last_solve_seconds.argtypes = [C.c_void_p]
        lib.orc_fac_smooth_level.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _ip,
                                             _dp]
        lib.orc_fac_op_ghost_fill.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.orc_fac_op_smooth.restype = C.c_int
        lib.orc_fac_op_smooth.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_double, _dp]
        lib.orc_fac_op_residual.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int]
        lib.orc_fac_op_restrict.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        lib.orc_fac_op_restrict_solution.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.orc_fac_op_prolong.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.orc_fac_op_rms.restype = C.c_double
        lib.orc_fac_op_rms.argtypes = [C.c_void_p, C.c_int]
        lib.orc_fac_op_vcycle.argtypes = [C.c_void_p]
        lib.orc_fac_op_dot.restype = C.c_double
        lib.orc_fac_op_dot.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.orc_fac_op_sum_cv.restype = C.c_double
        lib.orc_fac_op_sum_cv.argtypes = [C.c_void_p]
        lib.orc_fac_op_get_flux.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp]
        lib.orc_fac_set_coarse_mg.argtypes = [C.c_void_p, C.c_int]
        lib.orc_fac_set_prolongation.argtypes = [C.c_void_p, C.c_int]
        lib.orc_fac_set_cf_discretization.argtypes = [C.c_void_p, C.c_int]
        lib.orc_fac_set_algorithm.argtypes = [C.c_void_p, C.c_int]
        lib.orc_fac_mg_iterations.argtypes = [C.c_void_p]
        lib.orc_fac_mg_iterations.restype = C.c_int
        lib.orc_fac_mg_relres.argtypes = [C.c_void_p]
        lib.orc_fac_mg_relres.restype = C.c_double
        _oracle = lib
    return _oracle


_ref = None


def ref_lib():
    """The flang-compiled reference kernels, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = C.CDLL(REF_SO)
    return _ref


FIELDS = {"u": 0, "f": 1, "C": 2, "D": 3, "e": 4, "r": 5, "tr": 6, "te": 7}


class OracleFAC:
    """The oracle FAC solver over a static hierarchy.

    levels: list of lists of boxes (lo tuple, hi tuple) in each level's index
    space; ratio: refinement ratio between consecutive levels.
    """

    def __init__(self, dim, levels, ratio, domain_lo, domain_hi, xlo, xhi):
        self.lib = oracle_lib()
        self.dim = dim
        self.levels = levels
        npat = np.array([len(l) for l in levels], dtype=np.int32)
        self._boxarrs = []
        ptrs = (_ip * len(levels))()
        for ln, boxes in enumerate(levels):
            a = np.array([list(lo) + list(hi) for lo, hi in boxes], dtype=np.int32).ravel()
            self._boxarrs.append(a)
            ptrs[ln] = a.ctypes.data_as(_ip)
        rat = np.array([1] + [ratio] * (len(levels) - 1), dtype=np.int32)
        dlo = np.array(domain_lo, dtype=np.int32)
        dhi = np.array(domain_hi, dtype=np.int32)
        xl = np.array(xlo, dtype=np.float64)
        xh = np.array(xhi, dtype=np.float64)
        self.h = self.lib.orc_fac_create(dim, len(levels), npat.ctypes.data_as(_ip), ptrs,
                                         rat.ctypes.data_as(_ip), dlo.ctypes.data_as(_ip),
                                         dhi.ctypes.data_as(_ip), dptr(xl), dptr(xh))
        if not self.h:
            raise RuntimeError("orc_fac_create failed")

    def close(self):
        if self.h:
            self.lib.orc_fac_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_poisson(self, dvar=False, dconst=1.0, cmode=0, cconst=0.0):
        self.lib.orc_fac_set_poisson(self.h, int(dvar), dconst, cmode, cconst)

    def set_bc(self, a, b, g):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        g = np.asarray(g, np.float64)
        self.lib.orc_fac_set_bc(self.h, dptr(a), dptr(b), dptr(g))

    def set_params(self, max_cycles=10, residual_tol=1e-6, relative_residual_tol=-1.0,
                   pre=1, post=1, coarse_max_iterations=10, coarse_tol=1e-8, ewing=True,
                   nthreads=1):
        self.lib.orc_fac_set_params(self.h, max_cycles, residual_tol, relative_residual_tol,
                                    pre, post, coarse_max_iterations, coarse_tol, int(ewing),
                                    nthreads)

    def set_coarse_mg(self, on=True):
        """Coarsest level by geometric multigrid (cmax V-cycles, relative tol ctol)."""
        self.lib.orc_fac_set_coarse_mg(self.h, int(on))

    def set_prolongation(self, method):
        """prolongation_method name of CellPoissonFACOps."""
        m = {"CONSERVATIVE_LINEAR_REFINE": 0, "CONSTANT_REFINE": 1, "LINEAR_REFINE": 2}[method]
        self.lib.orc_fac_set_prolongation(self.h, m)

    def set_cf_discretization(self, name):
        """cf_discretization of CellPoissonFACOps."""
        m = {"Ewing": 0, "CONSTANT_REFINE": 1, "LINEAR_REFINE": 2,
             "CONSERVATIVE_LINEAR_REFINE": 3}[name]
        self.lib.orc_fac_set_cf_discretization(self.h, m)

    def set_algorithm(self, name):
        """algorithm_choice of FACPreconditioner."""
        a = {"default": 0, "mccormick-s4.3": 1, "pernice": 2}[name]
        self.lib.orc_fac_set_algorithm(self.h, a)

    def mg_stats(self):
        return self.lib.orc_fac_mg_iterations(self.h), self.lib.orc_fac_mg_relres(self.h)

    def size(self, field, ln, p, axis=0):
        return self.lib.orc_fac_array_size(self.h, FIELDS[field], ln, p, axis)

    def put(self, field, ln, p, arr, axis=0):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel(order="F")
        assert a.size == self.size(field, ln, p, axis), (field, a.size)
        self.lib.orc_fac_put(self.h, FIELDS[field], ln, p, axis, dptr(a))

    def get(self, field, ln, p, axis=0):
        a = np.empty(self.size(field, ln, p, axis), dtype=np.float64)
        self.lib.orc_fac_get(self.h, FIELDS[field], ln, p, axis, dptr(a))
        return a

    def solve(self, max_hist=10000):
        it = C.c_int(0)
        res = C.c_double(0)
        hist = np.zeros(max_hist)
        conv = self.lib.orc_fac_solve(self.h, C.byref(it), C.byref(res), dptr(hist), max_hist)
        return bool(conv), it.value, res.value, hist[: min(it.value, max_hist)].copy()

    def last_solve_seconds(self):
        return self.lib.orc_fac_last_solve_seconds(self.h)

    # level operators
    def ghost_fill(self, field, ln, with_coarse, homogeneous):
        self.lib.orc_fac_op_ghost_fill(self.h, FIELDS[field], ln, int(with_coarse),
                                       int(homogeneous))

    def smooth(self, ef, rf, ln, sweeps, tol=-1.0):
        mr = C.c_double(0)
        n = self.lib.orc_fac_op_smooth(self.h, FIELDS[ef], FIELDS[rf], ln, sweeps, tol,
                                       C.byref(mr))
        return n, mr.value

    def residual(self, rf, uf, ff, ln, homogeneous):
        self.lib.orc_fac_op_residual(self.h, FIELDS[rf], FIELDS[uf], FIELDS[ff], ln,
                                     int(homogeneous))

    def restrict(self, dstf, srcf, dest_ln):
        self.lib.orc_fac_op_restrict(self.h, FIELDS[dstf], FIELDS[srcf], dest_ln)

    def restrict_solution(self, field, dest_ln):
        self.lib.orc_fac_op_restrict_solution(self.h, FIELDS[field], dest_ln)

    def prolong(self, ef, ln):
        self.lib.orc_fac_op_prolong(self.h, FIELDS[ef], ln)

    def rms(self, field):
        return self.lib.orc_fac_op_rms(self.h, FIELDS[field])

    def dot(self, f1, f2):
        return self.lib.orc_fac_op_dot(self.h, FIELDS[f1], FIELDS[f2])

    def sum_cv(self):
        return self.lib.orc_fac_op_sum_cv(self.h)

    def vcycle(self):
        self.lib.orc_fac_op_vcycle(self.h)

    def full_composite_residual(self):
        return self.lib.orc_fac_full_composite_residual(self.h)

    def rms_f(self):
        return self.lib.orc_fac_rms_norm_f(self.h)
