"""The learned-smoother oracle pinned without restating it: finite
differences of the float64 loss against the backward pass, the adjoint
identity of the backward-data convolution, a brute-force convolution loop,
torch's own bfloat16 rounding, and the discrete eigenfunctions of the block
operator for the block solver and residual."""
import numpy as np
import pytest

from oracle import learned as L


def _rand_layers(rng, scale=0.3):
    layers = []
    for l in range(3):
        w = np.zeros((L.NTAP, L.CH, L.CH))
        w[:, :L.COUT[l], :L.CIN[l]] = rng.uniform(-scale, scale, (L.NTAP, L.COUT[l], L.CIN[l]))
        b = np.zeros(L.CH)
        b[:L.COUT[l]] = rng.uniform(-0.1, 0.1, L.COUT[l])
        layers.append(np.concatenate([w.ravel(), b]))
    return layers


def _problem(rng, nb=1):
    r = rng.standard_normal((nb, 16, 16, 16))
    logd = rng.uniform(-1, 1, (nb, 16, 16, 16))
    x, s = L.features(r, logd, 1.0 / 64)
    target = rng.standard_normal((nb, 16, 16, 16)) * s[:, None, None, None]
    return x, s, target


def test_conv_matches_brute_force_loop():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 16, 16, 16, 16))
    w = rng.standard_normal((27, 16, 16))
    out = L.conv(x, w)
    for (b, z, y, xx) in [(0, 0, 0, 0), (1, 15, 3, 7), (0, 8, 15, 15), (1, 5, 6, 0)]:
        s = np.zeros(16)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    t = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)
                    zz, yy, xc = z + dz, y + dy, xx + dx
                    if 0 <= zz < 16 and 0 <= yy < 16 and 0 <= xc < 16:
                        s += w[t] @ x[b, zz, yy, xc]
        np.testing.assert_allclose(out[b, z, y, xx], s, rtol=1e-12, atol=1e-12)


def test_backward_data_is_the_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 16, 16, 16, 16))
    g = rng.standard_normal((1, 16, 16, 16, 16))
    w = rng.standard_normal((27, 16, 16))
    lhs = np.sum(L.conv(x, w) * g)
    rhs = np.sum(x * L.conv_bwd(g, w))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # a backward pass without the tap flip is not the adjoint
    bad = sum(np.einsum("bzyxo,oc->bzyxc", L._shift(g, dz, dy, dx), w[t])
              for t, (dz, dy, dx) in enumerate(L.taps()))
    assert abs(np.sum(x * bad) - lhs) > 1e-3 * abs(lhs)


@pytest.mark.parametrize("seed", [3, 4])
def test_gradients_match_finite_differences(seed):
    rng = np.random.default_rng(seed)
    layers = _rand_layers(rng)
    x, s, target = _problem(rng)
    loss, grads = L.loss_and_grads(layers, x, target, s, emulate=False)
    eps = 1e-6
    picks = []
    for l in range(3):
        for _ in range(4):
            t = rng.integers(0, 27)
            co = rng.integers(0, L.COUT[l])
            ci = rng.integers(0, L.CIN[l])
            picks.append((l, t * 256 + co * 16 + ci, grads[l][t, co, ci]))
        co = rng.integers(0, L.COUT[l])
        picks.append((l, 27 * 256 + co, grads[l][27, co, 0]))
    for l, idx, g in picks:
        lp = [p.copy() for p in layers]
        lm = [p.copy() for p in layers]
        lp[l][idx] += eps
        lm[l][idx] -= eps
        fp = L.loss_and_grads(lp, x, target, s, emulate=False)[0]
        fm = L.loss_and_grads(lm, x, target, s, emulate=False)[0]
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - g) <= 1e-5 * max(abs(g), 1e-3 * loss), (l, idx, fd, g)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(
        -20, 20, 10000), np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0])])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(L.bf16(a), ref)
    assert L.bf16(np.float32([1 + 2 ** -8]))[0] == 1.0            # tie to even
    assert L.bf16(np.float32([1 + 3 * 2 ** -8]))[0] == 1 + 2 ** -6  # tie to even (up)


def _unit_faces(nb=1, d=1.0):
    one = np.full((nb, 16, 16, 17), d)
    return one, one.transpose(0, 1, 3, 2).copy(), one.transpose(0, 3, 1, 2).copy()


@pytest.mark.parametrize("k", [(1, 1, 1), (1, 2, 3), (15, 4, 9)])
def test_block_operator_eigenfunctions(k):
    h = 1.0 / 64
    v, lam = L.block_eigenfunction(*k)
    dxf, dyf, dzf = _unit_faces(d=2.5)
    Av = L.block_apply(v, dxf, dyf, dzf, h)
    np.testing.assert_allclose(Av, 2.5 * lam / h ** 2 * v, atol=1e-10 * abs(lam) / h ** 2)
    # the block solve inverts it
    e = L.block_sor(2.5 * lam / h ** 2 * v, dxf, dyf, dzf, h, 400, 1.8)
    assert np.abs(e - v).max() < 1e-10
    assert np.abs(L.block_residual(2.5 * lam / h ** 2 * v, e, dxf, dyf, dzf, h)).max() < \
        1e-8 * abs(lam) / h ** 2


def test_block_sor_one_sweep_is_red_black_gauss_seidel():
    """One sweep at omega 1 from zero equals a scalar loop of red then black updates."""
    rng = np.random.default_rng(6)
    h = 1.0 / 32
    r = rng.standard_normal((1, 16, 16, 16))
    dxf = np.exp(rng.standard_normal((1, 16, 16, 17)))
    dyf = np.exp(rng.standard_normal((1, 16, 17, 16)))
    dzf = np.exp(rng.standard_normal((1, 17, 16, 16)))
    e = L.block_sor(r, dxf, dyf, dzf, h, 1, 1.0)
    ref = np.zeros((16, 16, 16))
    for colour in (0, 1):
        for z in range(16):
            for y in range(16):
                for x in range(16):
                    if (x + y + z) % 2 != colour:
                        continue
                    num, den = -h * h * r[0, z, y, x], 0.0
                    for c, nb in ((dxf[0, z, y, x], ref[z, y, x - 1] if x else None),
                                  (dxf[0, z, y, x + 1], ref[z, y, x + 1] if x < 15 else None),
                                  (dyf[0, z, y, x], ref[z, y - 1, x] if y else None),
                                  (dyf[0, z, y + 1, x], ref[z, y + 1, x] if y < 15 else None),
                                  (dzf[0, z, y, x], ref[z - 1, y, x] if z else None),
                                  (dzf[0, z + 1, y, x], ref[z + 1, y, x] if z < 15 else None)):
                        if nb is None:
                            den += 2 * c   # ghost = -interior
                        else:
                            num += c * nb
                            den += c
                    ref[z, y, x] = num / den
    np.testing.assert_allclose(e[0], ref, rtol=1e-12, atol=1e-14)


def test_zero_residual_block_has_no_features_and_no_loss():
    """A block whose residual is zero gets scale 0, a zero first channel and
    no loss term: the loss over [live, zero] blocks is half the loss of the
    live block alone (the mean runs over every cell), with the same
    gradients halved."""
    rng = np.random.default_rng(7)
    layers = _rand_layers(rng)
    r = rng.standard_normal((2, 16, 16, 16))
    r[1] = 0.0
    logd = rng.uniform(-1, 1, (2, 16, 16, 16))
    x, s = L.features(r, logd, 1.0 / 64)
    assert s[1] == 0.0 and s[0] > 0.0
    assert not x[1, ..., 0].any() and x[1, ..., 1].any()
    target = rng.standard_normal((2, 16, 16, 16)) * s[0]
    loss2, g2 = L.loss_and_grads(layers, x, target, s, emulate=False)
    loss1, g1 = L.loss_and_grads(layers, x[:1], target[:1], s[:1], emulate=False)
    assert abs(loss2 - 0.5 * loss1) <= 1e-12 * loss1
    for a, b in zip(g2, g1):
        np.testing.assert_allclose(a, 0.5 * b, rtol=1e-10, atol=1e-14)
