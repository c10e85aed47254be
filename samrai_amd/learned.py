"""Learned block smoother over libsamrai_fac.so (include/samrai_learned.h).

A 3x3x3 convolution network (2 -> 16 -> 16 -> 1 channels, ReLU, BF16
weights on the matrix cores, FP32 accumulation, FP32 master weights and
Adam state) maps the residual of a 16^3 block (scaled by h^2 and the block
RMS, plus log D) to an error correction.  It is trained on the GPU by the
library's own forward/backward kernels on (residual, block error) pairs
that the FAC solver generates: residuals of the solver's iterates on seeded
variable-coefficient problems, and as target the solution of the block's
error equation div(D grad e) = r with homogeneous Dirichlet data on the
block faces (computed by FP64 red-black SOR on the device).

There is no counterpart in the reference: SAMRAI's CellPoissonFACOps smooths
with red-black Gauss-Seidel only (solv/CellPoissonFACOps.C:1338-1514); that
smoother stays the fallback and the FP64 residual stays the convergence test.
Device memory comes from torch tensors; every computation runs in the HIP
library (no CPU fallback).
"""
import ctypes as C

import numpy as np

from . import fac

NB = 16
NCELL = NB ** 3
CH = 16
NTAP = 27
PSZ = NTAP * CH * CH + CH
GSZ = (NTAP + 1) * CH * CH

_vp = C.c_void_p
_bound = False


def lib():
    global _bound
    L = fac.lib()
    if _bound:
        return L
    i, d, f, u64 = C.c_int, C.c_double, C.c_float, C.c_ulonglong
    L.sfac_conv3x3x3_bf16.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, i, i, _vp]
    L.sfac_conv3x3x3_wgrad_bf16.argtypes = [_vp, _vp, _vp, _vp, i, _vp]
    L.sfac_conv3x3x3_wgrad_partial_size.argtypes = [i]
    L.sfac_conv3x3x3_wgrad_partial_size.restype = C.c_size_t
    L.sfac_block_sor.argtypes = [_vp, _vp, _vp, _vp, d, i, d, i, _vp, i, _vp]
    L.sfac_block_residual.argtypes = [_vp, _vp, _vp, _vp, d, _vp, _vp, i, _vp]
    L.sfac_learned_create.argtypes = [i, u64, C.POINTER(_vp)]
    L.sfac_learned_destroy.argtypes = [_vp]
    L.sfac_learned_features.argtypes = [_vp, _vp, _vp, d, i, _vp]
    L.sfac_learned_forward.argtypes = [_vp, i, _vp]
    L.sfac_learned_correction.argtypes = [_vp, d, _vp, i, _vp]
    L.sfac_learned_backward.argtypes = [_vp, _vp, i, _vp]
    L.sfac_learned_adam.argtypes = [_vp, f, _vp]
    L.sfac_learned_train_step.argtypes = [_vp, _vp, _vp, _vp, d, i, f, _vp]
    L.sfac_learned_get_params.argtypes = [_vp, _vp, i]
    L.sfac_learned_set_params.argtypes = [_vp, _vp]
    L.sfac_learned_loss.argtypes = [_vp, _vp]
    L.sfac_learned_buffers.argtypes = [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]
    L.sfac_ops_learned_correct.argtypes = [_vp, _vp, _vp, _vp, _vp, i, d]
    L.sfac_solver_set_learned_smoother.argtypes = [_vp, _vp, d]
    L.sfac_solver_learned_active.argtypes = [_vp]
    L.sfac_solver_learned_rollbacks.argtypes = [_vp]
    L.sfac_solver_set_learned_phases.argtypes = [_vp, i]
    _bound = True
    return L


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _blocks_check(*ts):
    n = None
    for t in ts:
        if t is None:
            continue
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("block arrays must be contiguous device tensors")
        if n is None:
            n = t.shape[0]
        elif t.shape[0] != n:
            raise ValueError("block counts differ")
    return n


def conv3x3x3(torch, x, w, bias=None, mask=None, mode=0):
    """Device convolution (modes of sfac_conv3x3x3_bf16); x, w, mask int16 bit patterns."""
    nb = x.shape[0]
    out = pred = None
    if mode in (0, 2):
        out = torch.empty_like(x)
    if mode == 1:
        pred = torch.empty((nb, NB, NB, NB), dtype=torch.float32, device=x.device)
    fac._check(lib().sfac_conv3x3x3_bf16(_p(x), _p(w), _p(bias), _p(mask), _p(out), _p(pred),
                                         nb, mode, _stream(torch)))
    return out if mode != 1 else pred


def conv3x3x3_wgrad(torch, x, gz):
    nb = x.shape[0]
    part = torch.empty(max(1, lib().sfac_conv3x3x3_wgrad_partial_size(nb)), dtype=torch.float32,
                       device=x.device)
    grad = torch.empty(GSZ, dtype=torch.float32, device=x.device)
    fac._check(lib().sfac_conv3x3x3_wgrad_bf16(_p(x), _p(gz), _p(grad), _p(part), nb,
                                               _stream(torch)))
    return grad.view(NTAP + 1, CH, CH)


def block_sor(torch, r, dxf, dyf, dzf, h, sweeps, omega=1.0, e0=None):
    nb = _blocks_check(r, dxf, dyf, dzf, e0)
    e = torch.zeros_like(r) if e0 is None else e0.clone()
    fac._check(lib().sfac_block_sor(_p(r), _p(dxf), _p(dyf), _p(dzf), float(h), int(sweeps),
                                    float(omega), int(e0 is not None), _p(e), nb,
                                    _stream(torch)))
    return e


def block_residual(torch, r, e, dxf, dyf, dzf, h):
    nb = _blocks_check(r, e, dxf, dyf, dzf)
    out = torch.empty_like(r)
    fac._check(lib().sfac_block_residual(_p(r), _p(dxf), _p(dyf), _p(dzf), float(h), _p(e),
                                         _p(out), nb, _stream(torch)))
    return out


class LearnedSmoother:
    """The network and its optimizer state on the device."""

    def __init__(self, torch, max_blocks, seed=0):
        self.torch = torch
        self.max_blocks = int(max_blocks)
        h = _vp()
        fac._check(lib().sfac_learned_create(self.max_blocks, C.c_ulonglong(seed), C.byref(h)))
        self.h = h

    def close(self):
        """Free the network (an error while a solver still holds it)."""
        if self.h:
            fac._check(lib().sfac_learned_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(3 * PSZ, np.float32)
        fac._check(lib().sfac_learned_get_params(self.h, out.ctypes.data_as(_vp), 0))
        return [out[l * PSZ:(l + 1) * PSZ].copy() for l in range(3)]

    def grads(self):
        out = np.empty(3 * GSZ, np.float32)
        fac._check(lib().sfac_learned_get_params(self.h, out.ctypes.data_as(_vp), 1))
        return [out[l * GSZ:(l + 1) * GSZ].reshape(NTAP + 1, CH, CH).copy() for l in range(3)]

    def set_params(self, layers):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32) for p in layers]))
        assert flat.size == 3 * PSZ
        fac._check(lib().sfac_learned_set_params(self.h, flat.ctypes.data_as(_vp)))

    def loss(self):
        v = C.c_float()
        fac._check(lib().sfac_learned_loss(self.h, C.byref(v)))
        return v.value

    def buffers(self):
        f, p, s = _vp(), _vp(), _vp()
        fac._check(lib().sfac_learned_buffers(self.h, C.byref(f), C.byref(p), C.byref(s)))
        return f.value, p.value, s.value

    def _n(self, r):
        n = _blocks_check(r)
        if n > self.max_blocks:
            raise ValueError("%d blocks > max_blocks %d" % (n, self.max_blocks))
        return n

    def features(self, r, logd, h):
        n = self._n(r)
        fac._check(lib().sfac_learned_features(self.h, _p(r), _p(logd), float(h), n,
                                               _stream(self.torch)))
        return n

    def forward(self, r, logd, h):
        n = self.features(r, logd, h)
        fac._check(lib().sfac_learned_forward(self.h, n, _stream(self.torch)))
        return n

    def correction(self, r, logd, h, omega=1.0):
        """e = omega * s * net(r h^2 / s, log D), FP64 per cell."""
        n = self.forward(r, logd, h)
        e = self.torch.empty_like(r)
        fac._check(lib().sfac_learned_correction(self.h, float(omega), _p(e), n,
                                                 _stream(self.torch)))
        return e

    def backward(self, target):
        n = self._n(target)
        fac._check(lib().sfac_learned_backward(self.h, _p(target), n, _stream(self.torch)))

    def train_step(self, r, logd, target, h, lr):
        n = self._n(r)
        fac._check(lib().sfac_learned_train_step(self.h, _p(r), _p(logd), _p(target), float(h),
                                                 n, float(lr), _stream(self.torch)))

    def train(self, r, logd, target, h, steps, batch, lr=2e-3, seed=0, log_every=0):
        """Adam over random mini-batches of the pairs (device tensors); returns
        the loss of every log_every-th step (0: none)."""
        torch = self.torch
        n = r.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for k in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), generator=g).to(r.device)
            rb, tb = r.index_select(0, idx), target.index_select(0, idx)
            lb = None if logd is None else logd.index_select(0, idx)
            self.train_step(rb, lb, tb, h, lr)
            if log_every and (k % log_every == 0 or k == steps - 1):
                losses.append((k, self.loss()))
        return losses


def learned_correct(solver, net, u, r, D, ln, omega=1.0):
    """u += omega * s * net(r h^2 / s, log D) over the 16^3 windows covering
    the patches of level ln
    (sfac_ops_learned_correct); r: cell data with ghost 0, D: side data or None."""
    fac._check(lib().sfac_ops_learned_correct(solver.s, net.h, u.d, r.d,
                                              None if D is None else D.d, int(ln), float(omega)))


def set_learned_smoother(solver, net, omega=1.0):
    """Attach the network to a solver's smoothError (None detaches).  The
    solver keeps a reference to the network while attached: the library
    refuses to destroy a network a solver holds (its captured cycle graphs
    use the network's buffers)."""
    fac._check(lib().sfac_solver_set_learned_smoother(solver.s, None if net is None else net.h,
                                                      float(omega)))
    solver._learned_net = net


def set_learned_phases(solver, mask):
    """Bit 0: apply in pre-smoothing, bit 1: in post-smoothing (default 3)."""
    fac._check(lib().sfac_solver_set_learned_phases(solver.s, int(mask)))


def learned_active(solver):
    return lib().sfac_solver_learned_active(solver.s) == 1


def learned_rollbacks(solver):
    """Cycles of the last solve that the fallback undid."""
    return int(lib().sfac_solver_learned_rollbacks(solver.s))


class BlockLevelProblem:
    """A seeded hierarchy tiled by 16^3 patches on the unit cube (one n^3 level,
    or with two_level an (n/2)^3 level under a full n^3 level):
    D = exp(g) on faces (g a Gaussian random field, correlation length 0.1,
    unit variance), f = 8 Gaussian sources of width 0.02, homogeneous
    Dirichlet.  Holds u (ghost 1), f and a residual r (ghost 0), D, and the
    block face arrays and cell log D of every patch (patch p = block p)."""

    def __init__(self, torch, seed, n=64, two_level=False):
        from . import decks
        assert n % NB == 0 and (not two_level or n % (2 * NB) == 0)
        rng = np.random.default_rng(5000 + seed)
        n0 = n // 2 if two_level else n
        deck = dict(name="block-level-s%d" % seed, dim=3, ratio=2, dlo=(0, 0, 0),
                    dhi=(n0 - 1,) * 3, centres=rng.uniform(0.1, 0.9, (8, 3)),
                    amps=rng.uniform(-1.0, 1.0, 8), levels=[])

        def tiling(nn):
            m = nn // NB
            return [((a * NB, b * NB, c * NB),
                     (a * NB + NB - 1, b * NB + NB - 1, c * NB + NB - 1))
                    for c in range(m) for b in range(m) for a in range(m)]

        deck["levels"] = [tiling(n0)] + ([tiling(n)] if two_level else [])
        self.fine = len(deck["levels"]) - 1
        self.boxes = deck["levels"][self.fine]
        self.torch = torch
        self.h = 1.0 / n
        self.hier = fac.PatchHierarchy(3, deck["dlo"], deck["dhi"], [0.0] * 3, [1.0] * 3)
        for boxes in deck["levels"]:
            self.hier.add_level(2, boxes)
        self.hier.finalize()
        self.u = fac.PatchData(self.hier, 1)
        self.f = fac.PatchData(self.hier, 0)
        # r over device tensors: the fine level's patches are the blocks of one
        # [patches, 16, 16, 16] tensor, so its residual needs no gather
        kw = dict(dtype=torch.float64, device="cuda")
        self.rblk = torch.zeros((len(self.boxes), NB, NB, NB), **kw)
        arrays = []
        for ln, boxes in enumerate(deck["levels"]):
            if ln == self.fine:
                arrays += [self.rblk[p] for p in range(len(boxes))]
            else:
                arrays += [torch.zeros((NB, NB, NB), **kw) for _ in boxes]
        self.r = fac.PatchData.wrap(self.hier, 0, arrays)
        self.D = fac.PatchData(self.hier, 0, "side")
        modes = decks.grf_modes(seed)
        faces = [[], [], []]
        for ln, boxes in enumerate(deck["levels"]):
            for p, box in enumerate(boxes):
                fv = decks.gaussian_sources(deck, decks.cell_points(deck, ln, box, xp=torch, **kw),
                                            xp=torch)
                self.f.put_device(ln, p, fv.permute(2, 1, 0).contiguous())
                for a in range(3):
                    g = decks.grf_eval(decks.face_points(deck, ln, box, a, xp=torch, **kw), modes,
                                       xp=torch)
                    dv = torch.exp(g).permute(2, 1, 0).contiguous()
                    self.D.put_device(ln, p, dv, axis=a)
                    if ln == self.fine:
                        faces[a].append(dv)
        self.dxf, self.dyf, self.dzf = (torch.stack(fl).contiguous() for fl in faces)
        self.logd = (torch.log(self.dxf[..., :-1]) + torch.log(self.dxf[..., 1:]) +
                     torch.log(self.dyf[:, :, :-1]) + torch.log(self.dyf[:, :, 1:]) +
                     torch.log(self.dzf[:, :-1]) + torch.log(self.dzf[:, 1:])) / 6.0

    def solver(self, **params):
        s = fac.CellPoissonFACSolver(self.hier, **params)
        s.set_d_patch_data(self.D)
        s.set_dirichlet(0.0)
        s.initialize_solver_state(self.u, self.f)
        return s

    def residual_blocks(self):
        """The fine level's residual as a [patches, 16, 16, 16] device tensor
        (a copy of the tensor the residual patch data is bound to)."""
        return self.rblk.clone()

    def close(self):
        for x in (self.u, self.f, self.r, self.D):
            x.close()
        self.hier.close()


def fac_training_pairs(torch, seed, cycles=(1, 2, 3), n=64, coarse_sweeps=2, sor_sweeps=400,
                       sor_omega=1.8, faces=True):
    """(residual, log D, block error, D faces) for every 16^3 block of a
    one-level n^3 hierarchy (patches of 16^3) after each FAC cycle count in
    `cycles` (increasing), on a seeded problem: D = exp(g), g a Gaussian
    random field of correlation length 0.1 and unit variance, f = 8 Gaussian
    sources of width 0.02, homogeneous Dirichlet.  The FAC solver (red-black
    coarse solve with `coarse_sweeps` sweeps per cycle) produces the iterates,
    one cycle at a time from u = 0; the residual is its composite residual;
    the target is the FP64 solution of the block problem div(D grad e) = r
    with zero Dirichlet data on the block faces.  Everything stays on the
    device; faces=False leaves the D face blocks out of the result (they are
    needed only to judge a correction by its residual).  n = 128 with cycles
    1..8 gives 4096 pairs per seed."""
    from . import decks
    assert n % NB == 0
    cycles = [int(c) for c in cycles]
    assert all(c > 0 for c in cycles) and cycles == sorted(set(cycles)), "cycles must increase"
    m = n // NB
    rng = np.random.default_rng(5000 + seed)
    deck = dict(name="learned-pairs-s%d" % seed, dim=3, ratio=2, dlo=(0, 0, 0),
                dhi=(n - 1,) * 3, centres=rng.uniform(0.1, 0.9, (8, 3)),
                amps=rng.uniform(-1.0, 1.0, 8), levels=[])
    boxes = [((a * NB, b * NB, c * NB), (a * NB + NB - 1, b * NB + NB - 1, c * NB + NB - 1))
             for c in range(m) for b in range(m) for a in range(m)]
    deck["levels"] = [boxes]
    hh = fac.PatchHierarchy(3, deck["dlo"], deck["dhi"], [0.0] * 3, [1.0] * 3)
    hh.add_level(2, boxes)
    hh.finalize()
    u = fac.PatchData(hh, 1)
    f = fac.PatchData(hh, 0)
    # the residual is written straight into the block tensor (patch p = block p)
    rblk = torch.empty((len(boxes), NB, NB, NB), dtype=torch.float64, device="cuda")
    rv = fac.PatchData.wrap(hh, 0, [rblk[p] for p in range(len(boxes))])
    D = fac.PatchData(hh, 0, "side")
    modes = decks.grf_modes(seed)
    kw = dict(dtype=torch.float64, device="cuda")
    faces = [[], [], []]
    for p, box in enumerate(boxes):
        fv = decks.gaussian_sources(deck, decks.cell_points(deck, 0, box, xp=torch, **kw), xp=torch)
        f.put_device(0, p, fv.permute(2, 1, 0).contiguous())
        for a in range(3):
            g = decks.grf_eval(decks.face_points(deck, 0, box, a, xp=torch, **kw), modes, xp=torch)
            dv = torch.exp(g).permute(2, 1, 0).contiguous()
            D.put_device(0, p, dv, axis=a)
            faces[a].append(dv)
    dxf, dyf, dzf = (torch.stack(fl).contiguous() for fl in faces)
    logd = (torch.log(dxf[..., :-1]) + torch.log(dxf[..., 1:]) + torch.log(dyf[:, :, :-1]) +
            torch.log(dyf[:, :, 1:]) + torch.log(dzf[:, :-1]) + torch.log(dzf[:, 1:])) / 6.0
    h = 1.0 / n
    rs = []
    # one-cycle solves continue from the current iterate, so the residuals
    # after 1, 2, ... cycles come from one pass
    s = fac.CellPoissonFACSolver(hh, max_cycles=1, relative_residual_tol=1e-30,
                                 coarse_solver_choice="redblack", coarse_solver_tolerance=1e-30,
                                 coarse_solver_max_iterations=int(coarse_sweeps))
    s.set_d_patch_data(D)
    s.set_dirichlet(0.0)
    u.fill(0.0)
    s.initialize_solver_state(u, f)
    done = 0
    for c in cycles:
        while done < c:
            s.solve_system(u, f)
            done += 1
        s.residual(rv, u, f, 0, False)
        rs.append(rblk.clone())
    s.synchronize()
    s.close()
    r = torch.cat(rs).contiguous()
    k = len(cycles)
    dxf, dyf, dzf, logd = (t.repeat(k, 1, 1, 1).contiguous() for t in (dxf, dyf, dzf, logd))
    e = block_sor(torch, r, dxf, dyf, dzf, h, sor_sweeps, sor_omega)
    for x in (u, f, rv, D):
        x.close()
    hh.close()
    if not faces:
        return dict(r=r, logd=logd, e=e, h=h)
    return dict(r=r, logd=logd, e=e, dxf=dxf, dyf=dyf, dzf=dzf, h=h)
