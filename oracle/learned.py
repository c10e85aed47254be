"""Numpy restatement of the learned block smoother and of the block
problems it is trained on.

Test infrastructure only: imported by tests/ and bench.py's cpu_baseline
leg, never by the samrai_amd package.

The learned smoother has no counterpart in the reference (SAMRAI smooths with
red-black Gauss-Seidel only, solv/CellPoissonFACOps.C:1338-1514); this file
is the CPU definition the device kernels are held to.  Pins (tests/
test_oracle_learned.py): finite differences of the float64 loss against the
backward pass, the adjoint identity <conv(x), g> = <x, conv_bwd(g)>, a
brute-force loop for the convolution, and the discrete eigenfunctions of the
block operator for the block solver and residual.

Layout as in the library: blocks of 16^3 cells as arrays [b, z, y, x],
channels last; weights per layer w[t, co, ci] with
t = 9 (dz+1) + 3 (dy+1) + (dx+1), then bias[co].
"""
import numpy as np

NB = 16
CH = 16
NTAP = 27
CIN = (2, CH, CH)
COUT = (CH, CH, 1)
PSZ = NTAP * CH * CH + CH
GSZ = (NTAP + 1) * CH * CH


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(NTAP)]


def bf16(a):
    """Round to bfloat16 (nearest even) and return float32."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def bf16_bits(a):
    return (bf16(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def _shift(x, dz, dy, dx):
    """y[b, z, y, x] = x[b, z + dz, y + dy, x + dx] (0 outside the block)."""
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    return xp[:, 1 + dz:1 + dz + NB, 1 + dy:1 + dy + NB, 1 + dx:1 + dx + NB]


def conv(x, w):
    """out[p, co] = sum_t sum_ci x[p + off(t), ci] w[t, co, ci] (float64)."""
    x = x.astype(np.float64)
    out = np.zeros(x.shape[:4] + (w.shape[1],))
    for t, (dz, dy, dx) in enumerate(taps()):
        out += np.einsum("bzyxc,oc->bzyxo", _shift(x, dz, dy, dx), w[t].astype(np.float64))
    return out


def conv_bwd(g, w):
    """dx[p, ci] = sum_t sum_co g[p - off(t), co] w[t, co, ci] (adjoint of conv)."""
    g = g.astype(np.float64)
    out = np.zeros(g.shape[:4] + (w.shape[2],))
    for t, (dz, dy, dx) in enumerate(taps()):
        out += np.einsum("bzyxo,oc->bzyxc", _shift(g, -dz, -dy, -dx), w[t].astype(np.float64))
    return out


def wgrad(x, g):
    """grad[t, co, ci] = sum_p g[p, co] x[p + off(t), ci]; grad[27, co, 0] = sum_p g[p, co]."""
    x = x.astype(np.float64)
    g = g.astype(np.float64)
    out = np.zeros((NTAP + 1, g.shape[-1], x.shape[-1]))
    for t, (dz, dy, dx) in enumerate(taps()):
        out[t] = np.einsum("bzyxo,bzyxc->oc", g, _shift(x, dz, dy, dx))
    out[NTAP, :, 0] = g.sum(axis=(0, 1, 2, 3))
    return out


def split(params):
    """Flat parameters of one layer -> (w[27, 16, 16], bias[16])."""
    p = np.asarray(params)
    return p[:NTAP * CH * CH].reshape(NTAP, CH, CH), p[NTAP * CH * CH:]


def features(r, logd, h):
    """X[b, z, y, x, :] = (r h^2 / s, log D, 0, ...), s = RMS over the block of r h^2.
    A block whose residual is zero has s = 0: its first channel is 0, and
    the correction s * pred and its loss term are 0."""
    rh = r * h * h
    s = np.sqrt(np.mean(rh * rh, axis=(1, 2, 3)))
    live = s > 0
    x = np.zeros(r.shape + (CH,), np.float32)
    x[..., 0] = np.where(live[:, None, None, None],
                         rh / np.where(live, s, 1.0)[:, None, None, None], 0.0).astype(np.float32)
    s = np.where(live, s, 0.0)
    if logd is not None:
        x[..., 1] = logd.astype(np.float32)
    return x, s


def forward(layers, x, emulate=True):
    """Activations of the network.  emulate: round to BF16 where the device
    stores BF16 (weights, activations) and to FP32 where it stores FP32."""
    rnd = bf16 if emulate else (lambda a: a)
    acts = [rnd(x)]
    for l in range(3):
        w, b = split(layers[l])
        z = conv(acts[-1], rnd(w))
        if l < 2:
            a = np.maximum(z + b.astype(np.float64), 0.0)
            acts.append(rnd(a.astype(np.float32)) if emulate else a)
        else:
            pred = z[..., 0] + np.float64(b[0])
            acts.append(pred.astype(np.float32) if emulate else pred)
    return acts


def loss_and_grads(layers, x, target, scale, emulate=True):
    """Loss mean((pred - target/s)^2) and per-layer gradients [28, 16, 16]
    (the layout of the device gradient slabs)."""
    rnd = bf16 if emulate else (lambda a: a)
    acts = forward(layers, x, emulate)
    pred = acts[3]
    n = pred.size
    live = (scale > 0)[:, None, None, None]
    tgt = target / np.where(scale > 0, scale, 1.0)[:, None, None, None]
    if emulate:
        d = np.where(live, pred - tgt.astype(np.float32), np.float32(0.0))
    else:
        d = np.where(live, pred - tgt, 0.0)
    loss = float(np.sum(np.asarray(d, np.float64) ** 2) / n)
    g = np.zeros(pred.shape + (CH,), np.float32 if emulate else np.float64)
    g[..., 0] = 2.0 * d / n
    g = rnd(g)
    grads = [None] * 3
    for l in (2, 1, 0):
        w, _ = split(layers[l])
        grads[l] = wgrad(acts[l], g)
        if l > 0:
            gx = conv_bwd(g, rnd(w)) * (acts[l] > 0)
            g = rnd(gx.astype(np.float32)) if emulate else gx
    for l in range(3):  # entries outside the layer's real channel counts carry no gradient
        grads[l][:NTAP, COUT[l]:, :] = 0.0
        grads[l][:NTAP, :, CIN[l]:] = 0.0
    return loss, grads


# ---------------------------------------------------------------------------
# Block problems div(D grad e) = r, homogeneous Dirichlet on the block faces
# (ghost = -interior).  D on faces: dxf[b, z, y, 0..16], dyf[b, z, 0..16, x],
# dzf[b, 0..16, y, x].

def block_apply(e, dxf, dyf, dzf, h):
    """div(D grad e) with ghost = -interior."""
    out = np.zeros_like(e)
    # x faces
    lo = np.concatenate([-e[..., :1], e[..., :-1]], axis=3)
    hi = np.concatenate([e[..., 1:], -e[..., -1:]], axis=3)
    out += dxf[..., :-1] * (lo - e) + dxf[..., 1:] * (hi - e)
    lo = np.concatenate([-e[:, :, :1], e[:, :, :-1]], axis=2)
    hi = np.concatenate([e[:, :, 1:], -e[:, :, -1:]], axis=2)
    out += dyf[:, :, :-1] * (lo - e) + dyf[:, :, 1:] * (hi - e)
    lo = np.concatenate([-e[:, :1], e[:, :-1]], axis=1)
    hi = np.concatenate([e[:, 1:], -e[:, -1:]], axis=1)
    out += dzf[:, :-1] * (lo - e) + dzf[:, 1:] * (hi - e)
    return out / (h * h)


def block_residual(r, e, dxf, dyf, dzf, h):
    return r - block_apply(e, dxf, dyf, dzf, h)


def block_sor(r, dxf, dyf, dzf, h, sweeps, omega=1.0, e0=None):
    """Red-black SOR (red: x + y + z even), vectorised per colour."""
    e = np.zeros_like(r) if e0 is None else e0.copy()
    z, y, x = np.meshgrid(np.arange(NB), np.arange(NB), np.arange(NB), indexing="ij")
    diag = (dxf[..., :-1] + dxf[..., 1:] + dyf[:, :, :-1] + dyf[:, :, 1:] +
            dzf[:, :-1] + dzf[:, 1:])
    bdiag = (dxf[..., :-1] * (x == 0) + dxf[..., 1:] * (x == NB - 1) +
             dyf[:, :, :-1] * (y == 0) + dyf[:, :, 1:] * (y == NB - 1) +
             dzf[:, :-1] * (z == 0) + dzf[:, 1:] * (z == NB - 1))
    for _ in range(sweeps):
        for colour in (0, 1):
            m = ((x + y + z) & 1) == colour
            nb_int = block_apply(e, dxf, dyf, dzf, 1.0) + (diag + bdiag) * e
            gs = (nb_int - h * h * r) / (diag + bdiag)
            e = np.where(m, e + omega * (gs - e), e)
    return e


def block_eigenfunction(kx, ky, kz, nblocks=1):
    """v = prod sin(pi k (i + 1/2) / 16): div grad v = -(4/h^2) sum sin^2(pi k / 32) v
    for D = 1 under the ghost = -interior condition."""
    i = np.arange(NB) + 0.5
    vz, vy, vx = (np.sin(np.pi * k * i / NB) for k in (kz, ky, kx))
    v = vz[:, None, None] * vy[None, :, None] * vx[None, None, :]
    lam = -4.0 * sum(np.sin(np.pi * k / (2 * NB)) ** 2 for k in (kx, ky, kz))
    return np.broadcast_to(v, (nblocks, NB, NB, NB)).copy(), lam
