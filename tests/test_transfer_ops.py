"""Transfer operators: closed-form pins of the oracle, and device parity.

The reference's transfer kernels (geom_m4cartcoarsenops3d.i,
geom_m4cartrefineops3d.i, pdat_m4conrefineops3d.i) are m4 templates that
cannot be expanded here, so the oracle's restatement is pinned by what the
mathematics fixes:
  * CartesianCellDoubleConservativeLinearRefine reproduces linear data
    exactly (the reference's own check: T/communication/CellDataTest.C:155-200
    with Acoef..Dcoef = 2.1, 3.2, 4.3, 5.4 from
    T/communication/test_inputs/cell_refine_a.3d.input.conservative_refine:79-82),
    conserves the coarse value (mean of the children), and its limiter
    (geom_m4cartopstuff.i:39-63) gives hand-computed slopes on a 1-D profile;
  * CartesianCellDoubleWeightedAverage is the volume-weighted mean of the
    children (conservation), on anisotropic ratios and negative indices;
  * CartesianOutersideDoubleWeightedAverage is the area-weighted mean of the
    fine faces on each coarse face;
  * CellConstantRefine copies the parent value, parent = floor(fine / ratio).
Each test plants nothing in the oracle; a transposed child index, a floor
taken as truncation, a dropped limiter branch or a wrong volume fails one of
them.  The GPU tests run the device entry points on the same inputs and
require bit-identical results.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def iv(v):
    a = (C.c_int * 3)(*(list(v) + [0] * (3 - len(v))))
    return a


def dv(v):
    return (C.c_double * 3)(*(list(v) + [0.0] * (3 - len(v))))


def shape(lo, hi):
    """numpy shape (k, j, i) of a Fortran-ordered box array."""
    return tuple(h - l + 1 for l, h in zip(lo, hi))[::-1]


class Runner:
    """Calls orc_* on host arrays, or sfac_* on device copies."""

    def __init__(self, torch=None, lib=None):
        self.torch = torch
        self.lib = lib if lib is not None else O.oracle_lib()
        self.prefix = "sfac_" if torch is not None else "orc_"

    def __call__(self, name, ints, arrays, out):
        """ints: leading scalar/vector args; arrays: list of numpy arrays in
        call order (the last one named by `out` is the output)."""
        fn = getattr(self.lib, self.prefix + name)
        if self.torch is None:
            ptrs = [a.ctypes.data_as(_dp) for a in arrays]
            fn(*ints, *ptrs)
            return arrays[out]
        dev = [self.torch.from_numpy(a).cuda() for a in arrays]
        ptrs = [C.cast(C.c_void_p(d.data_ptr()), _dp) for d in dev]
        st = fn(*ints, *ptrs, C.c_void_p(None))
        assert st == 0, self.lib.sfac_last_error()
        self.torch.cuda.synchronize()
        return dev[out].cpu().numpy()


# ----------------------------------------------------------------- operators

def clinref(run, dim, cbox, fbox, cg, fg, ratio, dxc, dxf, uc):
    (clo, chi), (flo, fhi) = cbox, fbox
    cilo, cihi = cg
    filo, fihi = fg
    uf = np.full(shape(filo, fihi), np.nan)
    return run("cartclinrefcelldoub",
               [C.c_int(dim), iv(clo), iv(chi), iv(flo), iv(fhi), iv(cilo), iv(cihi), iv(filo),
                iv(fihi), iv(ratio), dv(dxc), dv(dxf)],
               [np.ascontiguousarray(uc, np.float64), uf], 1)


def linref(run, dim, cbox, fbox, cg, fg, ratio, dxc, dxf, uc):
    (clo, chi), (flo, fhi) = cbox, fbox
    uf = np.full(shape(*fg), np.nan)
    return run("cartlinrefcelldoub",
               [C.c_int(dim), iv(clo), iv(chi), iv(flo), iv(fhi), iv(cg[0]), iv(cg[1]),
                iv(fg[0]), iv(fg[1]), iv(ratio), dv(dxc), dv(dxf)],
               [np.ascontiguousarray(uc, np.float64), uf], 1)


def wgtavg(run, dim, cbox, fg, cg, ratio, dxf, dxc, uf):
    clo, chi = cbox
    uc = np.full(shape(*cg), np.nan)
    return run("cartwgtavgcelldoub",
               [C.c_int(dim), iv(clo), iv(chi), iv(fg[0]), iv(fg[1]), iv(cg[0]), iv(cg[1]),
                iv(ratio), dv(dxf), dv(dxc)],
               [np.ascontiguousarray(uf, np.float64), uc], 1)


def conref(run, dim, fbox, cg, fg, ratio, uc):
    flo, fhi = fbox
    uf = np.full(shape(*fg), np.nan)
    return run("conrefcelldoub",
               [C.c_int(dim), iv(flo), iv(fhi), iv(cg[0]), iv(cg[1]), iv(fg[0]), iv(fg[1]),
                iv(ratio)],
               [np.ascontiguousarray(uc, np.float64), uf], 1)


def outside(run, dim, axis, cbox, fg, cg, ratio, dxf, dxc, ff, cshape):
    clo, chi = cbox
    fc = np.full(cshape, np.nan)
    return run("cartwgtavgoutsidedoub",
               [C.c_int(dim), C.c_int(axis), iv(clo), iv(chi), iv(fg[0]), iv(fg[1]), iv(cg[0]),
                iv(cg[1]), iv(ratio), dv(dxf), dv(dxc)],
               [np.ascontiguousarray(ff, np.float64), fc], 1)


def centres(lo, hi, dx, xlo=0.0):
    """Cell-centre coordinate arrays over a box (numpy order k, j, i)."""
    axes = [xlo + dx[d] * (np.arange(lo[d], hi[d] + 1) + 0.5) for d in range(len(lo))]
    grids = np.meshgrid(*axes[::-1], indexing="ij")
    return grids[::-1]  # x, y, (z)


def linear(lo, hi, dx):
    """D + A x + B y + C z, the CellDataTest profile."""
    A, B, Cc, D = 2.1, 3.2, 4.3, 5.4
    X = centres(lo, hi, dx)
    v = D + A * X[0] + B * X[1]
    if len(lo) == 3:
        v = v + Cc * X[2]
    return v


# --------------------------------------------------------------- CPU pins

@pytest.mark.parametrize("dim,ratio", [(3, (2, 2, 2)), (3, (2, 3, 4)), (2, (2, 2)), (2, (4, 3))])
def test_clinref_reproduces_linear_data(dim, ratio):
    run = Runner()
    clo, chi = (-3, 1, 2)[:dim], (2, 4, 5)[:dim]
    dxc = (0.2, 0.125, 0.25)[:dim]
    dxf = tuple(dxc[d] / ratio[d] for d in range(dim))
    cg = (tuple(l - 1 for l in clo), tuple(h + 1 for h in chi))
    flo = tuple(clo[d] * ratio[d] for d in range(dim))
    fhi = tuple((chi[d] + 1) * ratio[d] - 1 for d in range(dim))
    uc = linear(cg[0], cg[1], dxc)
    uf = clinref(run, dim, (clo, chi), (flo, fhi), cg, (flo, fhi), ratio, dxc, dxf, uc)
    want = linear(flo, fhi, dxf)
    np.testing.assert_allclose(uf, want, rtol=1e-13, atol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_clinref_conserves_coarse_values(dim):
    run = Runner()
    rng = np.random.default_rng(11)
    ratio = (2,) * dim
    clo, chi = (0,) * dim, (5, 4, 3)[:dim]
    dxc = (1 / 16,) * dim
    dxf = tuple(d / 2 for d in dxc)
    cg = (tuple(l - 1 for l in clo), tuple(h + 1 for h in chi))
    uc = rng.standard_normal(shape(*cg))
    flo = tuple(2 * c for c in clo)
    fhi = tuple(2 * c + 1 for c in chi)
    uf = clinref(run, dim, (clo, chi), (flo, fhi), cg, (flo, fhi), ratio, dxc, dxf, uc)
    kids = uf.reshape(sum(((s // 2, 2) for s in uf.shape), ()))
    mean = kids.mean(axis=tuple(range(1, 2 * dim, 2)))
    inner = uc[(slice(1, -1),) * dim]
    np.testing.assert_allclose(mean, inner, rtol=0, atol=1e-14 * np.abs(uc).max())


def test_clinref_limiter_hand_values():
    """1-D profile 0, 1, 3, 10, 10, 2 along x (dx = 1, ratio 2, children
    at -1/4 and +1/4):
      cell 1: diffs (1, 2)   -> slope min(1.5, 2)  = 1.5 -> 0.625, 1.375
      cell 2: diffs (2, 7)   -> slope min(4.5, 4)  = 4   -> 2, 4
      cell 3: diffs (7, 0)   -> local extremum, 0  -> 10, 10
      cell 4: diffs (0, -8)  -> 0                  -> 10, 10"""
    run = Runner()
    prof = np.array([0.0, 1.0, 3.0, 10.0, 10.0, 2.0])
    cg = ((0, 0, 0), (5, 2, 2))
    uc = np.broadcast_to(prof, shape(*cg)).copy()
    clo, chi = (1, 1, 1), (4, 1, 1)
    flo, fhi = (2, 2, 2), (9, 3, 3)
    uf = clinref(run, 3, (clo, chi), (flo, fhi), cg, (flo, fhi), (2, 2, 2), (1.0,) * 3,
                 (0.5,) * 3, uc)
    want = np.array([0.625, 1.375, 2.0, 4.0, 10.0, 10.0, 10.0, 10.0])
    for k in range(2):
        for j in range(2):
            np.testing.assert_array_equal(uf[k, j], want)


@pytest.mark.parametrize("dim,ratio", [(3, (2, 2, 2)), (3, (2, 3, 4)), (2, (2, 2)), (2, (4, 3))])
def test_linear_refine_reproduces_linear_data(dim, ratio):
    """CartesianCellDoubleLinearRefine: multilinear interpolation of the
    coarse cell centres is exact for linear data (with one ghost cell of
    coarse data for the fine cells nearest the coarse box edges)."""
    run = Runner()
    clo, chi = (-3, 1, 2)[:dim], (2, 4, 5)[:dim]
    dxc = (0.2, 0.125, 0.25)[:dim]
    dxf = tuple(dxc[d] / ratio[d] for d in range(dim))
    cg = (tuple(l - 1 for l in clo), tuple(h + 1 for h in chi))
    flo = tuple(clo[d] * ratio[d] for d in range(dim))
    fhi = tuple((chi[d] + 1) * ratio[d] - 1 for d in range(dim))
    uc = linear(cg[0], cg[1], dxc)
    uf = linref(run, dim, (clo, chi), (flo, fhi), cg, (flo, fhi), ratio, dxc, dxf, uc)
    np.testing.assert_allclose(uf, linear(flo, fhi, dxf), rtol=1e-13, atol=0)


def test_linear_refine_hand_values():
    """1-D profile 0, 1, 3, 10 along x (dx = 1, ratio 2): the children of
    coarse cell 1 sit at x = -1/4, +1/4 from its centre, so they interpolate
    between cells (0, 1) at 3/4 and (1, 2) at 1/4:
      cell 1: 0 + 1 * 0.75 = 0.75,  1 + 2 * 0.25 = 1.5
      cell 2: 1 + 2 * 0.75 = 2.5,   3 + 7 * 0.25 = 4.75"""
    run = Runner()
    prof = np.array([0.0, 1.0, 3.0, 10.0])
    cg = ((0, 0, 0), (3, 2, 2))
    uc = np.broadcast_to(prof, shape(*cg)).copy()
    uf = linref(run, 3, ((1, 1, 1), (2, 1, 1)), ((2, 2, 2), (5, 3, 3)), cg,
                ((2, 2, 2), (5, 3, 3)), (2, 2, 2), (1.0,) * 3, (0.5,) * 3, uc)
    for k in range(2):
        for j in range(2):
            np.testing.assert_array_equal(uf[k, j], [0.75, 1.5, 2.5, 4.75])


@pytest.mark.parametrize("dim,ratio", [(3, (2, 2, 2)), (3, (3, 2, 4)), (2, (2, 2)), (2, (2, 3))])
def test_weighted_average_is_volume_mean(dim, ratio):
    run = Runner()
    rng = np.random.default_rng(5)
    clo, chi = (-2, 0, 3)[:dim], (1, 2, 5)[:dim]
    dxc = (0.1, 0.2, 0.05)[:dim]
    dxf = tuple(dxc[d] / ratio[d] for d in range(dim))
    flo = tuple(clo[d] * ratio[d] for d in range(dim))
    fhi = tuple((chi[d] + 1) * ratio[d] - 1 for d in range(dim))
    fg = (tuple(l - 1 for l in flo), tuple(h + 1 for h in fhi))  # ghost 1 on the fine array
    uf = rng.standard_normal(shape(*fg))
    uc = wgtavg(run, dim, (clo, chi), fg, (clo, chi), ratio, dxf, dxc, uf)
    inner = uf[(slice(1, -1),) * dim]
    rr = ratio[::-1]
    split = sum(((inner.shape[a] // rr[a], rr[a]) for a in range(dim)), ())
    mean = inner.reshape(split).mean(axis=tuple(range(1, 2 * dim, 2)))
    np.testing.assert_allclose(uc, mean, rtol=0, atol=1e-14 * np.abs(uf).max())
    # conservation with the actual volumes
    vf = np.prod(dxf)
    vc = np.prod(dxc)
    np.testing.assert_allclose(uc.sum() * vc, inner.sum() * vf, rtol=1e-13)


@pytest.mark.parametrize("dim,ratio", [(3, (2, 3, 4)), (2, (4, 2))])
def test_constant_refine_copies_parent(dim, ratio):
    run = Runner()
    rng = np.random.default_rng(3)
    cg = ((-4, -3, -2)[:dim], (3, 2, 1)[:dim])
    uc = rng.standard_normal(shape(*cg))
    # every fine cell's parent lies inside the coarse array
    fbox = (((-7, -8, -6), (6, 7, 7)) if dim == 3 else ((-13, -5), (14, 5)))
    uf = conref(run, dim, fbox, cg, fbox, ratio, uc)
    idx = np.meshgrid(*[np.arange(fbox[0][d], fbox[1][d] + 1) for d in range(dim)][::-1],
                      indexing="ij")[::-1]
    parent = [np.floor_divide(idx[d], ratio[d]) - cg[0][d] for d in range(dim)]
    want = uc[tuple(parent[::-1])]
    np.testing.assert_array_equal(uf, want)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_outerside_average_3d(axis):
    run = Runner()
    rng = np.random.default_rng(7 + axis)
    ratio = (2, 3, 2)
    tang = [d for d in range(3) if d != axis]
    clo, chi = [0, 0, 0], [0, 0, 0]
    for d, (l, h) in zip(tang, [(-1, 3), (2, 4)]):
        clo[d], chi[d] = l, h
    dxc = (0.1, 0.3, 0.2)
    dxf = tuple(dxc[d] / ratio[d] for d in range(3))
    flo = [clo[d] * ratio[d] for d in range(3)]
    fhi = [(chi[d] + 1) * ratio[d] - 1 for d in range(3)]
    i, o = tang
    ff = rng.standard_normal((fhi[o] - flo[o] + 1, fhi[i] - flo[i] + 1))
    cshape = (chi[o] - clo[o] + 1, chi[i] - clo[i] + 1)
    fc = outside(run, 3, axis, (clo, chi), (flo, fhi), (clo, chi), ratio, dxf, dxc, ff, cshape)
    mean = ff.reshape(cshape[0], ratio[o], cshape[1], ratio[i]).mean(axis=(1, 3))
    np.testing.assert_allclose(fc, mean, rtol=0, atol=1e-14 * np.abs(ff).max())


@pytest.mark.parametrize("axis", [0, 1])
def test_outerside_average_2d(axis):
    run = Runner()
    rng = np.random.default_rng(2)
    ratio = (2, 4)
    t = 1 - axis
    clo, chi = [0, 0], [0, 0]
    clo[t], chi[t] = -2, 5
    dxc = (0.25, 0.5)
    dxf = (dxc[0] / ratio[0], dxc[1] / ratio[1])
    flo = [clo[d] * ratio[d] for d in range(2)]
    fhi = [(chi[d] + 1) * ratio[d] - 1 for d in range(2)]
    ff = rng.standard_normal(fhi[t] - flo[t] + 1)
    fc = outside(run, 2, axis, (clo, chi), (flo, fhi), (clo, chi), ratio, dxf, dxc, ff,
                 (chi[t] - clo[t] + 1,))
    mean = ff.reshape(-1, ratio[t]).mean(axis=1)
    np.testing.assert_allclose(fc, mean, rtol=0, atol=1e-14 * np.abs(ff).max())


# ------------------------------------------------------- device parity

def _random_cases():
    rng = np.random.default_rng(2024)
    out = []
    for dim, ratio in [(3, (2, 2, 2)), (3, (2, 3, 4)), (2, (2, 2)), (2, (4, 3))]:
        clo = tuple(int(x) for x in rng.integers(-5, 5, dim))
        chi = tuple(clo[d] + int(rng.integers(0, 6)) for d in range(dim))
        out.append((dim, ratio, clo, chi))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dim,ratio,clo,chi", _random_cases())
def test_device_transfers_match_oracle(gpu_lib, torch_gpu, dim, ratio, clo, chi):
    cpu, gpu = Runner(), Runner(torch_gpu, gpu_lib)
    rng = np.random.default_rng(abs(hash((dim, ratio, clo))) % 2**32)
    dxc = tuple(float(x) for x in rng.uniform(0.05, 0.3, dim))
    dxf = tuple(dxc[d] / ratio[d] for d in range(dim))
    cg = (tuple(l - 1 for l in clo), tuple(h + 1 for h in chi))
    flo = tuple(clo[d] * ratio[d] for d in range(dim))
    fhi = tuple((chi[d] + 1) * ratio[d] - 1 for d in range(dim))
    fg = (tuple(l - 1 for l in flo), tuple(h + 1 for h in fhi))
    uc = rng.standard_normal(shape(*cg))
    uf = rng.standard_normal(shape(*fg))
    for run_args in [
        lambda r: clinref(r, dim, (clo, chi), (flo, fhi), cg, fg, ratio, dxc, dxf, uc),
        lambda r: linref(r, dim, (clo, chi), (flo, fhi), cg, fg, ratio, dxc, dxf, uc),
        lambda r: wgtavg(r, dim, (clo, chi), fg, cg, ratio, dxf, dxc, uf),
        lambda r: conref(r, dim, fg, cg, fg, ratio, uc),
    ]:
        a, b = run_args(cpu), run_args(gpu)
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(np.nan_to_num(a), np.nan_to_num(b))
    for axis in range(dim):
        tang = [d for d in range(dim) if d != axis]
        if dim == 3:
            i, o = tang
            ff = rng.standard_normal((fhi[o] - flo[o] + 1, fhi[i] - flo[i] + 1))
            cshape = (chi[o] - clo[o] + 1, chi[i] - clo[i] + 1)
        else:
            ff = rng.standard_normal(fhi[tang[0]] - flo[tang[0]] + 1)
            cshape = (chi[tang[0]] - clo[tang[0]] + 1,)
        a = outside(cpu, dim, axis, (clo, chi), (flo, fhi), (clo, chi), ratio, dxf, dxc, ff,
                    cshape)
        b = outside(gpu, dim, axis, (clo, chi), (flo, fhi), (clo, chi), ratio, dxf, dxc, ff,
                    cshape)
        np.testing.assert_array_equal(a, b)


def _by_axis(dim, *vecs):
    """Box corners as the Fortran calls pass them: one by-address int per axis."""
    out = []
    for v in vecs:
        out += [C.byref(C.c_int(int(v[d]))) for d in range(dim)]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dim,ratio,clo,chi", _random_cases()[:4])
def test_fortran_link_twins_of_transfer_ops(gpu_lib, torch_gpu, dim, ratio, clo, chi):
    """The link-compatible twins cartclinrefcelldoub{2,3}d_,
    cartlinrefcelldoub{2,3}d_, cartwgtavgcelldoub{2,3}d_, conrefcelldoub{2,3}d_
    and cartwgtavgoutsidedoub{2d0,2d1,3d0,3d1,3d2}_ (include/samrai_f77.h: the
    reference's call order,
    box corners one by-address int per axis, device arrays) give the sfac_
    entries' results bit for bit."""
    torch = torch_gpu
    gpu = Runner(torch, gpu_lib)
    rng = np.random.default_rng(abs(hash((dim, ratio, clo, 77))) % 2**32)
    dxc = tuple(float(x) for x in rng.uniform(0.05, 0.3, dim))
    dxf = tuple(dxc[d] / ratio[d] for d in range(dim))
    cg = (tuple(l - 1 for l in clo), tuple(h + 1 for h in chi))
    flo = tuple(clo[d] * ratio[d] for d in range(dim))
    fhi = tuple((chi[d] + 1) * ratio[d] - 1 for d in range(dim))
    fg = (tuple(l - 1 for l in flo), tuple(h + 1 for h in fhi))
    uc = rng.standard_normal(shape(*cg))
    uf = rng.standard_normal(shape(*fg))
    s = "%dd_" % dim

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()

    def ptr(t):
        return C.cast(C.c_void_p(t.data_ptr()), _dp)

    def twin(name, ints, tail, inp, out_shape):
        x, y = dev(inp), dev(np.full(out_shape, np.nan))
        getattr(gpu_lib, name + s)(*ints, *tail, ptr(x), ptr(y),
                                   *([C.c_void_p(None)] * (2 * dim if name.startswith("cartclin")
                                                           else 0)))
        torch.cuda.synchronize()
        return y.cpu().numpy()

    refine_box = _by_axis(dim, clo, chi, flo, fhi, cg[0], cg[1], fg[0], fg[1])
    cases = [
        (twin("cartclinrefcelldoub", refine_box, [iv(ratio), dv(dxc), dv(dxf)], uc, shape(*fg)),
         clinref(gpu, dim, (clo, chi), (flo, fhi), cg, fg, ratio, dxc, dxf, uc)),
        (twin("cartlinrefcelldoub", refine_box, [iv(ratio), dv(dxc), dv(dxf)], uc, shape(*fg)),
         linref(gpu, dim, (clo, chi), (flo, fhi), cg, fg, ratio, dxc, dxf, uc)),
        (twin("cartwgtavgcelldoub", _by_axis(dim, clo, chi, fg[0], fg[1], cg[0], cg[1]),
              [iv(ratio), dv(dxf), dv(dxc)], uf, shape(*cg)),
         wgtavg(gpu, dim, (clo, chi), fg, cg, ratio, dxf, dxc, uf)),
        (twin("conrefcelldoub", _by_axis(dim, cg[0], cg[1], fg[0], fg[1], cg[0], cg[1], fg[0],
                                         fg[1]), [iv(ratio)], uc, shape(*fg)),
         conref(gpu, dim, fg, cg, fg, ratio, uc)),
    ]
    # outerside flux average, one routine per face axis (cartwgtavgoutsidedoub<dim>d<axis>_)
    for axis in range(dim):
        tang = [d for d in range(dim) if d != axis]
        if dim == 3:
            i, o = tang
            ff = rng.standard_normal((fhi[o] - flo[o] + 1, fhi[i] - flo[i] + 1))
            cshape = (chi[o] - clo[o] + 1, chi[i] - clo[i] + 1)
        else:
            ff = rng.standard_normal(fhi[tang[0]] - flo[tang[0]] + 1)
            cshape = (chi[tang[0]] - clo[tang[0]] + 1,)
        x, y = dev(ff), dev(np.full(cshape, np.nan))
        getattr(gpu_lib, "cartwgtavgoutsidedoub%dd%d_" % (dim, axis))(
            *_by_axis(dim, clo, chi, flo, fhi, clo, chi), iv(ratio), dv(dxf), dv(dxc), ptr(x),
            ptr(y))
        torch.cuda.synchronize()
        cases.append((y.cpu().numpy(),
                      outside(gpu, dim, axis, (clo, chi), (flo, fhi), (clo, chi), ratio, dxf, dxc,
                              ff, cshape)))
    for a, b in cases:
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_array_equal(np.nan_to_num(a), np.nan_to_num(b))
        assert not np.isnan(b).all()
