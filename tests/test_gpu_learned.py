"""The learned block smoother on the device against the numpy oracle:
matrix-core convolutions (forward, backward data, weight gradient), the
network's loss and gradients, the FP64 block solver, and training on pairs
the FAC solver generates, with what the trained network gains."""
import sys

import numpy as np
import pytest

from oracle import learned as L

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bits(torch, a):
    return torch.from_numpy(L.bf16_bits(a).view(np.int16)).cuda()


def _f32(t):
    return L.from_bits(t.cpu().numpy().view(np.uint16))


def _rand_layers(rng, scale=0.3):
    layers = []
    for l in range(3):
        w = np.zeros((L.NTAP, L.CH, L.CH), np.float32)
        w[:, :L.COUT[l], :L.CIN[l]] = rng.uniform(-scale, scale, (L.NTAP, L.COUT[l], L.CIN[l]))
        b = np.zeros(L.CH, np.float32)
        b[:L.COUT[l]] = rng.uniform(-0.1, 0.1, L.COUT[l])
        layers.append(np.concatenate([w.ravel(), b]).astype(np.float32))
    return layers


def _close_bf16(got, exp):
    """Device outputs are FP32 sums rounded to BF16: within one BF16 ulp of
    the float64 sum rounded the same way."""
    tol = 2.0 ** -7 * np.abs(exp) + 1e-6 * np.abs(exp).max()
    bad = np.abs(got - exp) > tol
    assert not bad.any(), (int(bad.sum()), float(np.abs(got - exp).max()))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_conv_modes_match_oracle(torch_gpu, mode):
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("conv mode %d, 3 blocks" % mode)
    rng = np.random.default_rng(10 + mode)
    x = L.bf16(rng.standard_normal((3, 16, 16, 16, 16)).astype(np.float32))
    w = L.bf16(rng.uniform(-0.3, 0.3, (27, 16, 16)).astype(np.float32))
    bias = rng.uniform(-0.5, 0.5, 16).astype(np.float32)
    mask = L.bf16(np.maximum(rng.standard_normal(x.shape), 0).astype(np.float32))
    out = SL.conv3x3x3(torch, _bits(torch, x), _bits(torch, w),
                       torch.from_numpy(bias).cuda(), _bits(torch, mask), mode)
    torch.cuda.synchronize()
    z = L.conv(x, w)
    if mode == 0:
        _close_bf16(_f32(out).reshape(z.shape), L.bf16(np.maximum(z + bias, 0).astype(np.float32)))
    elif mode == 1:
        np.testing.assert_allclose(out.cpu().numpy(), z[..., 0] + bias[0], rtol=1e-5,
                                   atol=1e-5 * np.abs(z[..., 0]).max())
    else:
        # backward data: the tap-flipped transposed weights give the adjoint
        wb = np.ascontiguousarray(w[::-1].transpose(0, 2, 1))
        out = SL.conv3x3x3(torch, _bits(torch, x), _bits(torch, wb), None, _bits(torch, mask), 2)
        torch.cuda.synchronize()
        exp = L.conv_bwd(x, w) * (mask > 0)
        _close_bf16(_f32(out).reshape(exp.shape), L.bf16(exp.astype(np.float32)))


def test_wgrad_matches_oracle(torch_gpu):
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("weight gradient, 3 blocks")
    rng = np.random.default_rng(20)
    x = L.bf16(rng.standard_normal((3, 16, 16, 16, 16)).astype(np.float32))
    g = L.bf16(rng.standard_normal((3, 16, 16, 16, 16)).astype(np.float32))
    got = SL.conv3x3x3_wgrad(torch, _bits(torch, x), _bits(torch, g)).cpu().numpy()
    exp = L.wgrad(x, g)
    # FP32 sums over 12288 cells of O(1) products
    np.testing.assert_allclose(got[:27], exp[:27], rtol=0, atol=2e-6 * np.sqrt(12288) * 4)
    np.testing.assert_allclose(got[27, :, 0], exp[27, :, 0], rtol=0, atol=1e-3)
    assert np.all(got[27, :, 1:] == 0)


def _problem(torch, rng, nb):
    r = rng.standard_normal((nb, 16, 16, 16))
    logd = rng.uniform(-1.5, 1.5, (nb, 16, 16, 16))
    target = rng.standard_normal((nb, 16, 16, 16)) * 1e-4
    return (torch.from_numpy(r).cuda(), torch.from_numpy(logd).cuda(),
            torch.from_numpy(target).cuda())


def test_network_loss_and_gradients_match_oracle(torch_gpu):
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("network loss and gradients, 3 blocks (the last with a zero residual)")
    rng = np.random.default_rng(30)
    net = SL.LearnedSmoother(torch, 4, seed=1)
    layers = _rand_layers(rng)
    net.set_params(layers)
    np.testing.assert_array_equal(np.concatenate(net.params()), np.concatenate(layers))
    r, logd, target = _problem(torch, rng, 3)
    r[2] = 0.0  # scale 0: no features, no correction, no loss term
    h = 1.0 / 64
    net.forward(r, logd, h)
    net.backward(target)
    torch.cuda.synchronize()
    feat_ptr, _, scale_ptr = net.buffers()
    # features as the device made them (bits), and its block scales
    feat = _wrap(torch, feat_ptr, 3 * 4096 * 16, torch.int16).view(3, 16, 16, 16, 16)
    scale = _wrap(torch, scale_ptr, 3, torch.float64).cpu().numpy()
    assert scale[2] == 0.0 and scale[0] > 0.0
    assert not bool(feat[2, ..., 0].any())
    x = _f32(feat)
    xo, so = L.features(r.cpu().numpy(), logd.cpu().numpy(), h)
    np.testing.assert_allclose(scale, so, rtol=1e-13)
    _close_bf16(x, L.bf16(xo))
    loss_o, grads_o = L.loss_and_grads(layers, x, target.cpu().numpy(), scale, emulate=True)
    loss = net.loss()
    assert abs(loss - loss_o) <= 1e-4 * loss_o, (loss, loss_o)
    grads = net.grads()
    for l in range(3):
        ref = grads_o[l]
        err = np.linalg.norm(grads[l] - ref) / np.linalg.norm(ref)
        _say("layer %d: relative gradient error %.2e" % (l, err))
        assert err < 1e-4, (l, err)
    e = net.correction(r, logd, h)
    assert not bool(e[2].any()) and bool(e[0].any())
    net.close()


def _wrap(torch, ptr, n, dtype):
    """A device tensor viewing n elements at ptr (copied out immediately)."""
    import ctypes
    esize = torch.tensor([], dtype=dtype).element_size()
    out = torch.empty(n, dtype=dtype, device="cuda")
    from samrai_amd import fac
    hip = ctypes.CDLL(fac.LIB_PATH)  # hipMemcpy through the HIP runtime the library uses
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), n * esize,
                         3) == 0  # hipMemcpyDeviceToDevice
    return out


def test_block_sor_and_residual(torch_gpu):
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("block SOR / residual, 4 blocks")
    rng = np.random.default_rng(40)
    h = 1.0 / 64
    nb = 4
    dxf = np.exp(rng.standard_normal((nb, 16, 16, 17)))
    dyf = np.exp(rng.standard_normal((nb, 16, 17, 16)))
    dzf = np.exp(rng.standard_normal((nb, 17, 16, 16)))
    r = rng.standard_normal((nb, 16, 16, 16))
    t = [torch.from_numpy(a).cuda() for a in (r, dxf, dyf, dzf)]
    for sweeps, omega in ((1, 1.0), (7, 1.6)):
        e = SL.block_sor(torch, *t, h, sweeps, omega).cpu().numpy()
        eo = L.block_sor(r, dxf, dyf, dzf, h, sweeps, omega)
        np.testing.assert_allclose(e, eo, rtol=1e-12, atol=1e-12 * np.abs(eo).max())
    res = SL.block_residual(torch, t[0], torch.from_numpy(eo).cuda(), *t[1:], h).cpu().numpy()
    reso = L.block_residual(r, eo, dxf, dyf, dzf, h)
    np.testing.assert_allclose(res, reso, rtol=0, atol=1e-10 * np.abs(r).max())
    # closed form: the discrete eigenfunction is reproduced
    v, lam = L.block_eigenfunction(2, 5, 1, nblocks=1)
    one = np.ones((1, 16, 16, 17))
    f = [torch.from_numpy(a).cuda() for a in
         (lam / h ** 2 * v, one, one.transpose(0, 1, 3, 2).copy(), one.transpose(0, 3, 1, 2).copy())]
    e = SL.block_sor(torch, *f, h, 400, 1.8).cpu().numpy()
    assert np.abs(e - v).max() < 1e-10


def test_training_on_fac_pairs_gains_over_no_network(torch_gpu):
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("FAC-generated pairs: seeds 0-3 train, seed 9 test")
    train = [SL.fac_training_pairs(torch, s, cycles=(1, 2)) for s in range(4)]
    test = SL.fac_training_pairs(torch, 9, cycles=(1, 2))
    cat = {k: torch.cat([p[k] for p in train]).contiguous() for k in ("r", "logd", "e")}
    h = train[0]["h"]
    _say("pairs: %d train blocks, %d test blocks" % (cat["r"].shape[0], test["r"].shape[0]))
    net = SL.LearnedSmoother(torch, 128, seed=3)

    def quality(n):
        e = n.correction(test["r"], test["logd"], h)
        rr = SL.block_residual(torch, test["r"], e, test["dxf"], test["dyf"], test["dzf"], h)
        err = (torch.linalg.vector_norm(e - test["e"]) / torch.linalg.vector_norm(test["e"])).item()
        res = (torch.linalg.vector_norm(rr) / torch.linalg.vector_norm(test["r"])).item()
        return err, res, e

    err0, res0, _ = quality(net)
    losses = net.train(cat["r"], cat["logd"], cat["e"], h, steps=600, batch=32, lr=2e-3,
                       log_every=100)
    _say("losses: %s" % losses)
    err1, res1, e_net = quality(net)
    # red-black Gauss-Seidel with the same data, from zero and after the network
    gs1 = SL.block_sor(torch, test["r"], test["dxf"], test["dyf"], test["dzf"], h, 1, 1.0)
    err_gs = (torch.linalg.vector_norm(gs1 - test["e"]) / torch.linalg.vector_norm(test["e"])).item()
    both = SL.block_sor(torch, test["r"], test["dxf"], test["dyf"], test["dzf"], h, 1, 1.0,
                        e0=e_net)
    err_both = (torch.linalg.vector_norm(both - test["e"]) /
                torch.linalg.vector_norm(test["e"])).item()
    _say("unseen seed: block-error ratio untrained %.3f trained %.3f (1 RBGS sweep %.3f, "
         "network + 1 sweep %.3f); residual ratio untrained %.3f trained %.3f"
         % (err0, err1, err_gs, err_both, res0, res1))
    assert losses[-1][1] < 0.5 * losses[0][1]
    assert err1 < 0.5 * err0
    assert err1 < err_gs        # the trained network beats one red-black sweep on the block error
    assert err_both < err_gs    # and still helps when the sweep follows it
    net.close()


def _interior_blocks(torch, P, x):
    """Interior cells of every patch of cell data x (ghost 1) as [p, 16, 16, 16]."""
    return torch.stack([torch.as_tensor(x.get(0, p), device="cuda").view(18, 18, 18)[1:-1, 1:-1, 1:-1]
                        for p in range(len(P.boxes))]).contiguous()


def test_level_learned_correction_and_smoothing(torch_gpu):
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("learned correction of a level of 16^3 patches (sfac_ops_learned_correct)")
    train = [SL.fac_training_pairs(torch, s, cycles=(1, 2)) for s in range(4)]
    cat = {k: torch.cat([p[k] for p in train]).contiguous() for k in ("r", "logd", "e")}
    net = SL.LearnedSmoother(torch, 128, seed=3)
    net.train(cat["r"], cat["logd"], cat["e"], train[0]["h"], steps=600, batch=32, lr=2e-3)
    P = SL.BlockLevelProblem(torch, 11)
    s = P.solver()
    # the entry point applies exactly the block path: gather r, network, add into u
    P.u.fill(0.0)
    s.residual(P.r, P.u, P.f, 0, False)
    rb = P.residual_blocks()
    e_ref = net.correction(rb, None, P.h)
    SL.learned_correct(s, net, P.u, P.r, None, 0)
    s.synchronize()
    assert torch.equal(_interior_blocks(torch, P, P.u), e_ref)
    P.u.fill(0.0)
    e_ref = net.correction(rb, P.logd, P.h)
    SL.learned_correct(s, net, P.u, P.r, P.D, 0)
    s.synchronize()
    got = _interior_blocks(torch, P, P.u)
    assert (torch.linalg.vector_norm(got - e_ref) / torch.linalg.vector_norm(e_ref)).item() < 1e-3
    # smoothing the level: red-black (2 sweeps) per iteration, with and without
    # the learned correction first; the FP64 residual judges both
    hist = {}
    for learned in (False, True):
        P.u.fill(0.0)
        h = []
        for it in range(6):
            if learned:
                s.residual(P.r, P.u, P.f, 0, False)
                SL.learned_correct(s, net, P.u, P.r, P.D, 0)
            s.smooth(P.u, P.f, 0, 2)
            s.residual(P.r, P.u, P.f, 0, False)
            h.append(s.rms_norm(P.r))
        hist[learned] = h
    _say("RMS residual, red-black only: %s" % ["%.3e" % v for v in hist[False]])
    _say("RMS residual, learned + red-black: %s" % ["%.3e" % v for v in hist[True]])
    assert hist[True][-1] < hist[False][-1]
    s.close()
    P.close()
    net.close()


def test_fac_solve_with_learned_smoother(torch_gpu):
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("FAC solve (2 levels, 32^3 + 64^3 in 16^3 patches) with and without the learned smoother")
    train = [SL.fac_training_pairs(torch, s, cycles=(1, 2)) for s in range(4)]
    cat = {k: torch.cat([p[k] for p in train]).contiguous() for k in ("r", "logd", "e")}
    net = SL.LearnedSmoother(torch, 128, seed=3)
    net.train(cat["r"], cat["logd"], cat["e"], train[0]["h"], steps=600, batch=32, lr=2e-3)
    P = SL.BlockLevelProblem(torch, 12, n=64, two_level=True)
    s = P.solver(max_cycles=60, residual_tol=1e-300, relative_residual_tol=1e-10,
                 num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_choice="mg",
                 coarse_solver_tolerance=1e-8, coarse_solver_max_iterations=10,
                 prolongation_method="CONSERVATIVE_LINEAR_REFINE")
    out = {}
    for learned in (False, True):
        SL.set_learned_smoother(s, net if learned else None)
        P.u.fill(0.0)
        ok = s.solve_system(P.u, P.f)
        s.synchronize()
        u = torch.stack([torch.as_tensor(P.u.get(1, p)) for p in range(len(P.boxes))])
        out[learned] = (ok, s.history(), u, SL.learned_active(s))
        _say("learned=%s: converged %s in %d cycles, factors %s, still active %s"
             % (learned, ok, len(out[learned][1]),
                ["%.3f" % (b / a) for a, b in zip(out[learned][1][:-1], out[learned][1][1:])][:8],
                out[learned][3]))
    assert out[False][0] and out[True][0]
    u0, u1 = out[False][2], out[True][2]
    # both meet the FP64 relative residual 1e-10: the solutions agree to that level
    assert (u1 - u0).abs().max().item() <= 1e-7 * u0.abs().max().item()
    assert not out[False][3]
    SL.set_learned_smoother(s, None)
    s.close()
    P.close()
    net.close()


def _window_level(torch, boxes, ext, seed):
    """A one-level hierarchy of `boxes` on a domain of ext cells (cubic cells
    of width 1/16), seeded r (ghost 0), positive D (ghost 0) given cell by
    cell from global fields, u = 0."""
    from samrai_amd import fac
    h = fac.PatchHierarchy(3, (0, 0, 0), tuple(e - 1 for e in ext), [0.0] * 3,
                           [e / 16.0 for e in ext])
    h.add_level(2, boxes)
    h.finalize()
    rng = np.random.default_rng(seed)
    rg = rng.standard_normal(tuple(ext))                     # [i, j, k]
    dg = [np.exp(0.5 * rng.standard_normal(tuple(ext[d] + (a == d) for d in range(3))))
          for a in range(3)]
    u, r, D = fac.PatchData(h, 1), fac.PatchData(h, 0), fac.PatchData(h, 0, "side")
    for p, (lo, hi) in enumerate(boxes):
        sl = tuple(slice(lo[d], hi[d] + 1) for d in range(3))
        r.put(0, p, rg[sl])
        for a in range(3):
            sa = tuple(slice(lo[d], hi[d] + 1 + (d == a)) for d in range(3))
            D.put(0, p, dg[a][sa], axis=a)
    u.fill(0.0)
    return h, u, r, D, rg, dg


def test_learned_correction_on_patches_of_any_size(torch_gpu):
    """The correction covers each patch with 16^3 windows from its lower
    corner: a level of 32x16x16 patches gets exactly the correction of the
    same cells in 16^3 patches, and a window reaching past a 24-wide patch
    equals the network on that block with zeros beyond the patch."""
    torch = torch_gpu
    from samrai_amd import fac
    from samrai_amd import learned as SL
    _say("learned correction over 16^3 windows of non-16^3 patches")
    net = SL.LearnedSmoother(torch, 64, seed=4)
    n = 32
    cubes = [((a, b, c), (a + 15, b + 15, c + 15)) for c in (0, 16) for b in (0, 16) for a in (0, 16)]
    slabs = [((0, b, c), (31, b + 15, c + 15)) for c in (0, 16) for b in (0, 16)]
    got = {}
    for name, boxes in (("cubes", cubes), ("slabs", slabs)):
        h, u, r, D, rg, dg = _window_level(torch, boxes, (n, n, n), 21)
        s = fac.CellPoissonFACSolver(h, max_cycles=1)
        s.set_d_patch_data(D)
        s.set_dirichlet(0.0)
        f = fac.PatchData(h, 0)
        s.initialize_solver_state(u, f)
        SL.learned_correct(s, net, u, r, D, 0)
        s.synchronize()
        full = np.zeros((n, n, n))
        for p, (lo, hi) in enumerate(boxes):
            shp = tuple(hi[d] - lo[d] + 3 for d in range(3))
            v = u.get(0, p).reshape(shp, order="F")[1:-1, 1:-1, 1:-1]
            full[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = v
        got[name] = full
        s.close()
    assert np.abs(got["cubes"]).max() > 0.0
    assert np.array_equal(got["cubes"], got["slabs"])
    # a 24 x 16 x 16 patch: the window at x = 16 holds 8 patch cells and 8 zeros
    boxes = [((0, 0, 0), (23, 15, 15))]
    h, u, r, D, rg, dg = _window_level(torch, boxes, (24, 16, 16), 22)
    s = fac.CellPoissonFACSolver(h, max_cycles=1)
    s.set_d_patch_data(D)
    s.set_dirichlet(0.0)
    f = fac.PatchData(h, 0)
    s.initialize_solver_state(u, f)
    rb = np.zeros((2, 16, 16, 16))        # [block, z, y, x]
    lb = np.zeros((2, 16, 16, 16))
    for w, x0 in enumerate((0, 16)):
        m = 16 if w == 0 else 8
        rb[w, :, :, :m] = rg[x0:x0 + m, 0:16, 0:16].transpose(2, 1, 0)
        ld = (np.log(dg[0][x0:x0 + m, 0:16, 0:16]) + np.log(dg[0][x0 + 1:x0 + m + 1, 0:16, 0:16]) +
              np.log(dg[1][x0:x0 + m, 0:16, 0:16]) + np.log(dg[1][x0:x0 + m, 1:17, 0:16]) +
              np.log(dg[2][x0:x0 + m, 0:16, 0:16]) + np.log(dg[2][x0:x0 + m, 0:16, 1:17])) / 6.0
        lb[w, :, :, :m] = ld.transpose(2, 1, 0)
    for with_d in (False, True):
        u.fill(0.0)
        SL.learned_correct(s, net, u, r, D if with_d else None, 0)
        s.synchronize()
        v = u.get(0, 0).reshape((26, 18, 18), order="F")[1:-1, 1:-1, 1:-1]
        e = net.correction(torch.from_numpy(rb).cuda(),
                           torch.from_numpy(lb).cuda() if with_d else None, 1.0 / 16)
        e = e.cpu().numpy()
        want = np.concatenate([e[0].transpose(2, 1, 0), e[1].transpose(2, 1, 0)[:8]], axis=0)
        assert np.abs(v).max() > 0.0
        if with_d:
            # log D formed on the device and by numpy may differ in the last bit
            assert np.linalg.norm(v - want) <= 1e-3 * np.linalg.norm(want)
        else:
            assert np.array_equal(v, want)
    s.close()
    net.close()


def test_fac_solve_with_learned_smoother_on_uneven_patches(torch_gpu):
    """Whole solves on a reduced variable-coefficient deck of the batched
    family (3 levels, patch sides 8..16, so the finer levels are covered by
    partial 16^3 windows) with the learned correction in every smoothing
    call and in post-smoothing only: each converges to 1e-10 and agrees with
    the red-black solve to that level, with at most one cycle undone by the
    fallback; a solve that repeats the red-black history undid its first
    learned cycle and ends on the red-black solution bit for bit."""
    torch = torch_gpu
    from samrai_amd import decks, fac
    from samrai_amd import learned as SL
    deck = decks.deck_3d_c2(seed=2, base=32, nlevels=3, smin=8, smax=16)
    sides = {hi[d] - lo[d] + 1 for lv in deck["levels"][1:] for lo, hi in lv for d in range(3)}
    assert sides != {16}
    nwin = max(sum(int(np.prod([-(-(hh - ll + 1) // 16) for ll, hh in zip(*b)])) for b in lv)
               for lv in deck["levels"][1:])
    _say("learned FAC solve on a reduced uneven deck: %s patches, %d windows at most per level"
         % ([len(l) for l in deck["levels"]], nwin))
    train = [SL.fac_training_pairs(torch, s, cycles=(1, 2)) for s in range(4)]
    cat = {k: torch.cat([p[k] for p in train]).contiguous() for k in ("r", "logd", "e")}
    net = SL.LearnedSmoother(torch, max(128, nwin), seed=3)
    net.train(cat["r"], cat["logd"], cat["e"], train[0]["h"], steps=600, batch=32, lr=2e-3)
    params = dict(max_cycles=60, residual_tol=1e-300, relative_residual_tol=1e-10,
                  num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_choice="mg",
                  coarse_solver_tolerance=1e-8, coarse_solver_max_iterations=10,
                  prolongation_method="CONSERVATIVE_LINEAR_REFINE")
    _, h, u, f, D, s = decks.build_c2_problem(fac, torch, params, deck=deck)

    def solution():
        return [np.asarray(u.get(ln, p)) for ln, lv in enumerate(deck["levels"])
                for p in range(len(lv))]

    out = {}
    for name, on, phases in (("rbgs", False, 3), ("both", True, 3), ("post", True, 2)):
        SL.set_learned_smoother(s, net if on else None)
        SL.set_learned_phases(s, phases)
        u.fill(0.0)
        ok = s.solve_system(u, f)
        s.synchronize()
        out[name] = (ok, list(s.history()), solution(), SL.learned_rollbacks(s))
        _say("%s: converged %s in %d cycles, rollbacks %d" % (name, ok, len(out[name][1]),
                                                            out[name][3]))
    ok0, hist0, u0, rb0 = out["rbgs"]
    assert ok0 and rb0 == 0
    scale = max(np.abs(a).max() for a in u0)
    for name in ("both", "post"):
        ok, hist, u1, rb = out[name]
        assert ok and rb in (0, 1)
        assert max(np.abs(a - b).max() for a, b in zip(u1, u0)) <= 1e-7 * scale
        if hist == hist0:
            # the fallback undid the first learned cycle: red-black from there
            assert rb == 1
            assert all(np.array_equal(a, b) for a, b in zip(u1, u0))
    SL.set_learned_smoother(s, None)
    s.close()
    for x in (u, f, D):
        x.close()
    h.close()
    net.close()


def test_learned_phases_select_the_smoothing_calls(torch_gpu):
    """sfac_solver_set_learned_phases: mask 0 applies the correction nowhere
    (the red-black history, bit for bit); masks 1 (pre-smoothing), 2
    (post-smoothing) and 3 (both) each change the history, and differ from
    one another, unless the fallback undid the first learned cycle (then
    the history is red-black's)."""
    torch = torch_gpu
    from samrai_amd import learned as SL
    _say("learned phases 0..3 on a 2-level problem of 16^3 patches")
    train = [SL.fac_training_pairs(torch, s, cycles=(1, 2)) for s in range(4)]
    cat = {k: torch.cat([p[k] for p in train]).contiguous() for k in ("r", "logd", "e")}
    net = SL.LearnedSmoother(torch, 128, seed=3)
    net.train(cat["r"], cat["logd"], cat["e"], train[0]["h"], steps=600, batch=32, lr=2e-3)
    P = SL.BlockLevelProblem(torch, 13, n=64, two_level=True)
    s = P.solver(max_cycles=3, residual_tol=1e-300, relative_residual_tol=1e-300,
                 num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_choice="mg",
                 coarse_solver_tolerance=1e-8, coarse_solver_max_iterations=10,
                 prolongation_method="CONSERVATIVE_LINEAR_REFINE")
    hist = {}
    for mask in (None, 0, 1, 2, 3):
        SL.set_learned_smoother(s, None if mask is None else net)
        SL.set_learned_phases(s, 3 if mask is None else mask)
        P.u.fill(0.0)
        s.solve_system(P.u, P.f)
        s.synchronize()
        hist[mask] = (list(s.history()), SL.learned_rollbacks(s))
        _say("mask %s: history %s, rollbacks %d" % (mask, ["%.4e" % v for v in hist[mask][0]],
                                                    hist[mask][1]))
    rb = hist[None][0]
    assert hist[0][0] == rb
    applied = []
    for mask in (1, 2, 3):
        h, undone = hist[mask]
        if h == rb:
            assert undone == 1
        else:
            applied.append(h)
    for i in range(len(applied)):
        for j in range(i + 1, len(applied)):
            assert applied[i] != applied[j]
    SL.set_learned_smoother(s, None)
    s.close()
    P.close()
    net.close()
