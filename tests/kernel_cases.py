"""Seeded per-patch kernel cases shared by the golden-vector generator, the
oracle tests and the GPU per-patch parity tests.

A case is a dict: {"kernel", "dim", "ints": {...}, "dbl": {...},
"arrays": {...}, "outs": [names of arrays the kernel writes]}.  Arrays are
flat float64 buffers in column-major (Fortran / pdat::ArrayData) order.

Runners:
  run_ref(lib, case)    -- the flang-built reference Fortran (by-reference ABI)
  run_oracle(lib, case) -- the C oracle (oracle/kernels.c)
Both return {name: array} for the output arrays.
"""
import ctypes as C

import numpy as np

_dp = C.POINTER(C.c_double)


def _hp(a):
    return a.ctypes.data_as(_dp) if a is not None else C.cast(None, _dp)


_p = _hp


def _ri(v):
    return C.byref(C.c_int(int(v)))


def _rd(v):
    return C.byref(C.c_double(float(v)))


def _iv(v):
    a = (C.c_int * len(v))(*[int(x) for x in v])
    return a


def _dv(v):
    a = (C.c_double * len(v))(*[float(x) for x in v])
    return a


def cell_n(lo, hi, g):
    return int(np.prod([h - l + 1 + 2 * gg for l, h, gg in zip(lo, hi, g)]))


def side_n(lo, hi, g, axis):
    return int(np.prod([h - l + 1 + 2 * gg + (1 if d == axis else 0)
                        for d, (l, h, gg) in enumerate(zip(lo, hi, g))]))


def _box(rng, dim, lo_rng=(-4, 4), size_rng=(2, 7)):
    lo = [int(rng.integers(*lo_rng)) for _ in range(dim)]
    hi = [l + int(rng.integers(*size_rng)) - 1 for l in lo]
    return lo, hi


def make_cases(seed=20260101):
    rng = np.random.default_rng(seed)
    cases = []
    for dim in (3, 2):
        for rep in range(2):
            lo, hi = _box(rng, dim)
            fg = [int(rng.integers(0, 2))] * dim
            dg = [int(rng.integers(0, 2))] * dim
            sg = [int(rng.integers(1, 3))] * dim
            rg = [int(rng.integers(0, 2))] * dim
            sfg = [int(rng.integers(0, 2))] * dim
            dx = list(rng.uniform(0.05, 0.5, dim))
            soln = rng.standard_normal(cell_n(lo, hi, sg))
            dcs = [rng.uniform(0.5, 2.0, side_n(lo, hi, dg, a)) for a in range(dim)]
            fl = [rng.standard_normal(side_n(lo, hi, fg, a)) for a in range(dim)]
            rhs = rng.standard_normal(cell_n(lo, hi, rg))
            sf = rng.uniform(-1.0, 0.0, cell_n(lo, hi, sfg))
            base = {"dim": dim, "lo": lo, "hi": hi, "fg": fg, "dg": dg, "sg": sg,
                    "rg": rg, "sfg": sfg, "dx": dx}
            # flux
            for var in (True, False):
                c = dict(base, kernel="compflux", var=var, dc=float(rng.uniform(0.5, 2)))
                c["arrays"] = {"soln": soln.copy()}
                for a in range(dim):
                    c["arrays"]["flux%d" % a] = fl[a].copy()
                    if var:
                        c["arrays"]["dc%d" % a] = dcs[a].copy()
                c["outs"] = ["flux%d" % a for a in range(dim)]
                cases.append(c)
            # red-black GS, 4 variants x 2 colours
            for var in (True, False):
                for varsf in (True, False):
                    for offset in (0, 1):
                        c = dict(base, kernel="rbgs", var=var, varsf=varsf, offset=offset,
                                 dc=float(rng.uniform(0.5, 2)),
                                 sfc=float(rng.uniform(-1, 0)))
                        c["arrays"] = {"soln": soln.copy(), "rhs": rhs.copy()}
                        for a in range(dim):
                            c["arrays"]["flux%d" % a] = fl[a].copy()
                            if var:
                                c["arrays"]["dc%d" % a] = dcs[a].copy()
                        if varsf:
                            c["arrays"]["sf"] = sf.copy()
                        c["outs"] = ["soln", "maxres"]
                        c["arrays"]["maxres"] = np.zeros(1)
                        cases.append(c)
            # residual
            for varsf in (True, False):
                c = dict(base, kernel="compres", varsf=varsf, sfc=float(rng.uniform(-1, 0)))
                c["arrays"] = {"soln": soln.copy(), "rhs": rhs.copy(),
                               "residual": rng.standard_normal(cell_n(lo, hi, rg))}
                for a in range(dim):
                    c["arrays"]["flux%d" % a] = fl[a].copy()
                if varsf:
                    c["arrays"]["sf"] = sf.copy()
                c["outs"] = ["residual"]
                cases.append(c)
        # Ewing fix: aligned patches, every location, ratio 2 and 4
        for ratio in (2, 4):
            for loc in range(2 * dim):
                for var in (True, False):
                    r = ratio
                    lo = [r * int(rng.integers(-2, 3)) for _ in range(dim)]
                    hi = [l + r * int(rng.integers(1, 4)) - 1 for l in lo]
                    fg = [int(rng.integers(0, 2))] * dim
                    dg = [int(rng.integers(0, 2))] * dim
                    sg = [1] * dim
                    dx = list(rng.uniform(0.05, 0.5, dim))
                    n = loc // 2
                    upper = loc % 2
                    blo = [0] * dim
                    bhi = [0] * dim
                    for d in range(dim):
                        if d == n:
                            g = hi[d] + 1 if upper else lo[d] - 1
                            blo[d] = bhi[d] = g
                        else:
                            # either the full face, a sub-range, or one that
                            # spills past the patch (clipped by the kernel)
                            mode = int(rng.integers(0, 3))
                            if mode == 0:
                                blo[d], bhi[d] = lo[d], hi[d]
                            elif mode == 1:
                                blo[d], bhi[d] = lo[d], lo[d] + r - 1
                            else:
                                blo[d], bhi[d] = lo[d] - r, hi[d] + r
                    c = {"kernel": "ewing", "dim": dim, "lo": lo, "hi": hi, "fg": fg,
                         "dg": dg, "sg": sg, "dx": dx, "var": var, "loc": loc,
                         "ratio": [r] * dim, "blo": blo, "bhi": bhi,
                         "dc": float(rng.uniform(0.5, 2))}
                    c["arrays"] = {"soln": rng.standard_normal(cell_n(lo, hi, sg))}
                    for a in range(dim):
                        c["arrays"]["flux%d" % a] = rng.standard_normal(side_n(lo, hi, fg, a))
                        if var:
                            c["arrays"]["dc%d" % a] = rng.uniform(0.5, 2, side_n(lo, hi, dg, a))
                    c["outs"] = ["flux%d" % a for a in range(dim)]
                    cases.append(c)
        # Robin type-1 ghosts (faces)
        for loc in range(2 * dim):
            for zerog in (0, 1):
                dlo = [-3] * dim
                dhi = [6] * dim
                n = loc // 2
                upper = loc % 2
                ghos = dhi[n] if upper else dlo[n]
                inte = ghos - 1 if upper else ghos + 1
                face = ghos if upper else ghos + 1
                beg = [dlo[d] + 1 for d in range(dim)]
                end = [dhi[d] - 1 for d in range(dim)]
                beg[n] = end[n] = ghos
                clo = list(beg)
                chi = list(end)
                clo[n] = chi[n] = face
                ncoef = int(np.prod([h - l + 1 for l, h in zip(clo, chi)]))
                c = {"kernel": "settype1", "dim": dim, "dlo": dlo, "dhi": dhi, "clo": clo,
                     "chi": chi, "beg": beg, "end": end, "face": face, "ghos": ghos,
                     "inte": inte, "loc": loc, "h": float(rng.uniform(0.05, 0.5)),
                     "zerog": zerog}
                c["arrays"] = {
                    "data": rng.standard_normal(cell_n(dlo, dhi, [0] * dim)),
                    "a": rng.uniform(0.2, 1.5, ncoef),
                    "b": rng.uniform(0.2, 1.5, ncoef) * (loc % 3 != 0),
                    "g": rng.standard_normal(ncoef)}
                c["outs"] = ["data"]
                cases.append(c)
        # Robin type-2 (edges in 3-D, nodes in 2-D) and type-3 (3-D nodes)
        nloc2 = 12 if dim == 3 else 4
        for loc in range(nloc2):
            dlo = [-3] * dim
            dhi = [6] * dim
            lower = [0] * dim
            upper = [0] * dim
            if dim == 3:
                along = 2 - loc // 4
                pq = [d for d in range(3) if d != along]
                m = loc % 4
                for d in range(3):
                    if d == along:
                        lower[d], upper[d] = -2, 5
                lower[pq[0]] = upper[pq[0]] = (5 if m & 1 else -2)
                lower[pq[1]] = upper[pq[1]] = (5 if m & 2 else -2)
            else:
                lower[0] = upper[0] = 5 if loc & 1 else -2
                lower[1] = upper[1] = 5 if loc & 2 else -2
            c = {"kernel": "settype2", "dim": dim, "dlo": dlo, "dhi": dhi, "lower": lower,
                 "upper": upper, "loc": loc}
            c["arrays"] = {"data": rng.standard_normal(cell_n(dlo, dhi, [0] * dim))}
            c["outs"] = ["data"]
            cases.append(c)
        if dim == 3:
            for loc in range(8):
                dlo = [-3] * 3
                dhi = [6] * 3
                lower = [-2] * 3
                upper = [5] * 3
                c = {"kernel": "settype3", "dim": 3, "dlo": dlo, "dhi": dhi, "lower": lower,
                     "upper": upper, "loc": loc}
                c["arrays"] = {"data": rng.standard_normal(cell_n(dlo, dhi, [0] * 3))}
                c["outs"] = ["data"]
                cases.append(c)
    return cases


def _copy_arrays(case):
    return {k: np.array(v, dtype=np.float64, copy=True) for k, v in case["arrays"].items()}


def _boxargs_ref(lo, hi):
    out = []
    for l, h in zip(lo, hi):
        out += [_ri(l), _ri(h)]
    return out


# --------------------------------------------------------------------- ref

def run_ref(lib, case, torch=None):
    """The by-reference (SAMRAI_F77_FUNC) calling convention: the
    flang-built reference with host arrays, or, with torch, the library's
    link-compatible twins (include/samrai_f77.h) with the arrays on the
    device and scalars and maxres on the host."""
    A = _copy_arrays(case)
    dev = {}
    if torch is not None:
        for k, v in A.items():
            dev[k] = torch.from_numpy(v).cuda()

    def _p(a):
        if torch is None or a is None:
            return _hp(a)
        key = next(k for k, v in A.items() if v is a)
        return C.cast(C.c_void_p(dev[key].data_ptr()), _dp)

    dim = case["dim"]
    s = "3d" if dim == 3 else "2d"
    k = case["kernel"]
    if k in ("compflux", "rbgs", "compres", "ewing"):
        lo, hi = case["lo"], case["hi"]
        fl = [_p(A["flux%d" % a]) for a in range(dim)]
        fg = [_ri(case["fg"][d]) for d in range(dim)]
        sg = [_ri(case["sg"][d]) for d in range(dim)]
        box = _boxargs_ref(lo, hi)
        dx = _dv(case["dx"])
        if case.get("var", False):
            dc = [_p(A["dc%d" % a]) for a in range(dim)]
            dg = [_ri(case["dg"][d]) for d in range(dim)]
        if k == "compflux":
            if case["var"]:
                getattr(lib, "compfluxvardc%s_" % s)(*fl, *fg, *dc, *dg, _p(A["soln"]), *sg,
                                                     *box, dx)
            else:
                getattr(lib, "compfluxcondc%s_" % s)(*fl, *fg, _rd(case["dc"]), _p(A["soln"]),
                                                     *sg, *box, dx)
        elif k == "rbgs":
            rg = [_ri(case["rg"][d]) for d in range(dim)]
            name = "rbgswithfluxmax%s%s%s_" % ("vardc" if case["var"] else "condc",
                                               "varsf" if case["varsf"] else "consf", s)
            dpart = [*dc, *dg] if case["var"] else [_rd(case["dc"])]
            if case["varsf"]:
                sfg = [_ri(case["sfg"][d]) for d in range(dim)]
                spart = [_p(A["sf"]), *sfg]
            else:
                spart = [_rd(case["sfc"])]
            mr = C.c_double(0.0)
            getattr(lib, name)(*fl, *fg, *dpart, _p(A["rhs"]), *rg, *spart, _p(A["soln"]), *sg,
                               *box, dx, _ri(case["offset"]), C.byref(mr))
            A["maxres"][0] = mr.value
        elif k == "compres":
            rg = [_ri(case["rg"][d]) for d in range(dim)]
            name = "compres%ssca%s_" % ("var" if case["varsf"] else "con", s)
            if case["varsf"]:
                sfg = [_ri(case["sfg"][d]) for d in range(dim)]
                spart = [_p(A["sf"]), *sfg]
            else:
                spart = [_rd(case["sfc"])]
            getattr(lib, name)(*fl, *fg, _p(A["rhs"]), *rg, _p(A["residual"]), *rg, *spart,
                               _p(A["soln"]), *sg, *box, dx)
        elif k == "ewing":
            name = "ewingfixflux%sdc%s_" % ("var" if case["var"] else "con", s)
            dpart = [*dc, *dg] if case["var"] else [_rd(case["dc"])]
            getattr(lib, name)(*fl, *fg, *dpart, _p(A["soln"]), *sg, *box, _ri(case["loc"]),
                               _iv(case["ratio"]), _iv(case["blo"]), _iv(case["bhi"]), dx)
    elif k == "settype1":
        dbox = _boxargs_ref(case["dlo"], case["dhi"])
        cbox = _boxargs_ref(case["clo"], case["chi"])
        lbox = _boxargs_ref(case["beg"], case["end"])
        getattr(lib, "settype1cells%s_" % s)(_p(A["data"]), *dbox, _p(A["a"]), _p(A["b"]),
                                             _p(A["g"]), *cbox, *lbox, _ri(case["face"]),
                                             _ri(case["ghos"]), _ri(case["inte"]),
                                             _ri(case["loc"]), _rd(case["h"]),
                                             _ri(case["zerog"]))
    elif k in ("settype2", "settype3"):
        dbox = _boxargs_ref(case["dlo"], case["dhi"])
        name = ("settype2cells%s_" if k == "settype2" else "settype3cells%s_") % s
        getattr(lib, name)(_p(A["data"]), *dbox, _iv(case["lower"]), _iv(case["upper"]),
                           _ri(case["loc"]))
    else:
        raise ValueError(k)
    if torch is not None:
        torch.cuda.synchronize()
        for o in case["outs"]:
            if o != "maxres":
                A[o] = dev[o].cpu().numpy()
    return {o: A[o] for o in case["outs"]}


# ------------------------------------------------------------------ oracle

def _bv(lo, hi):
    out = []
    for l, h in zip(lo, hi):
        out += [C.c_int(l), C.c_int(h)]
    return out


def _ci(v):
    return C.c_int(int(v))


def _cd(v):
    return C.c_double(float(v))


def run_oracle(lib, case):
    return run_c(lib, case, "orc_", None)


def run_gpu(lib, case, torch):
    """Run a case through the device entry points sfac_* (device arrays,
    null stream) and copy the outputs back."""
    return run_c(lib, case, "sfac_", torch)


def run_c(lib, case, prefix, torch):
    A = _copy_arrays(case)
    dev = {}
    if torch is not None:
        for k, v in A.items():
            dev[k] = torch.from_numpy(v).cuda()

    def P(name):
        if name is None:
            return _p(None)
        if torch is not None:
            return C.cast(C.c_void_p(dev[name].data_ptr()), _dp)
        return _p(A[name])

    tail = [C.c_void_p(None)] if torch is not None else []
    dim = case["dim"]
    s = "3d" if dim == 3 else "2d"
    k = case["kernel"]

    def fn(name):
        return getattr(lib, prefix + name)

    if k in ("compflux", "rbgs", "compres", "ewing"):
        lo, hi = case["lo"], case["hi"]
        fl = [P("flux%d" % a) for a in range(dim)]
        fg = [_ci(case["fg"][d]) for d in range(dim)]
        sg = [_ci(case["sg"][d]) for d in range(dim)]
        box = _bv(lo, hi)
        dx = _dv(case["dx"])
        var = case.get("var", False)
        dc = [P("dc%d" % a) if var else P(None) for a in range(dim)]
        dg = [_ci(case["dg"][d]) if var else _ci(0) for d in range(dim)]
        dcv = _cd(case.get("dc", 0.0))
        if k == "compflux":
            if var:
                st = fn("compfluxvardc%s" % s)(*fl, *fg, *dc, *dg, P("soln"), *sg, *box, dx, *tail)
            else:
                st = fn("compfluxcondc%s" % s)(*fl, *fg, dcv, P("soln"), *sg, *box, dx, *tail)
        elif k in ("rbgs", "compres"):
            rg = [_ci(case["rg"][d]) for d in range(dim)]
            varsf = case["varsf"]
            sfp = P("sf") if varsf else P(None)
            sfg = [_ci(case["sfg"][d]) if varsf else _ci(0) for d in range(dim)]
            sfc = _cd(case["sfc"])
            if k == "rbgs":
                if torch is not None:
                    mrp = P("maxres")
                    st = fn("rbgswithfluxmax%s" % s)(*fl, *fg, *dc, *dg, dcv, P("rhs"), *rg, sfp,
                                                     *sfg, sfc, P("soln"), *sg, *box, dx,
                                                     _ci(case["offset"]), mrp, *tail)
                else:
                    mr = C.c_double(0.0)
                    st = fn("rbgswithfluxmax%s" % s)(*fl, *fg, *dc, *dg, dcv, P("rhs"), *rg, sfp,
                                                     *sfg, sfc, P("soln"), *sg, *box, dx,
                                                     _ci(case["offset"]), C.byref(mr))
                    A["maxres"][0] = mr.value
            else:
                st = fn("compres%s" % s)(*fl, *fg, P("rhs"), *rg, P("residual"), sfp, *sfg, sfc,
                                         P("soln"), *sg, *box, dx, *tail)
        elif k == "ewing":
            st = fn("ewingfixflux%s" % s)(*fl, *fg, *dc, *dg, dcv, P("soln"), *sg, *box,
                                          _ci(case["loc"]), _iv(case["ratio"]), _iv(case["blo"]),
                                          _iv(case["bhi"]), dx, *tail)
    elif k == "settype1":
        dbox = _bv(case["dlo"], case["dhi"])
        cbox = _bv(case["clo"], case["chi"])
        lbox = _bv(case["beg"], case["end"])
        st = fn("settype1cells%s" % s)(P("data"), *dbox, P("a"), P("b"), P("g"), *cbox, *lbox,
                                       _ci(case["face"]), _ci(case["ghos"]), _ci(case["inte"]),
                                       _ci(case["loc"]), _cd(case["h"]), _ci(case["zerog"]),
                                       *tail)
    elif k in ("settype2", "settype3"):
        dbox = _bv(case["dlo"], case["dhi"])
        name = ("settype2cells%s" if k == "settype2" else "settype3cells%s") % s
        st = fn(name)(P("data"), *dbox, _iv(case["lower"]), _iv(case["upper"]), _ci(case["loc"]),
                      *tail)
    else:
        raise ValueError(k)
    if torch is not None:
        if st != 0:
            raise RuntimeError("device kernel returned %d" % st)
        torch.cuda.synchronize()
        for o in case["outs"]:
            A[o] = dev[o].cpu().numpy()
    return {o: A[o] for o in case["outs"]}


# -------------------------------------------------------- fixture (de)serialisation

_META_KEYS = ("kernel", "dim", "lo", "hi", "fg", "dg", "sg", "rg", "sfg", "dx", "var", "varsf",
              "offset", "dc", "sfc", "loc", "ratio", "blo", "bhi", "dlo", "dhi", "clo", "chi",
              "beg", "end", "face", "ghos", "inte", "h", "zerog", "lower", "upper", "outs")


def case_to_json(case):
    return {k: case[k] for k in _META_KEYS if k in case}
