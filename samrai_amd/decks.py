"""Deterministic hierarchies and data for the benchmark decks.

The reference builds hierarchies with BergerRigoutsos / TreeLoadBalancer
(source/SAMRAI/mesh), which is outside this path.  These generators produce
properly nested, ratio-aligned static hierarchies of the shapes the decks
describe, deterministically by seed:

  deck_2d_c0(n)   2-D unit square, n^2 base + one level at ratio 2 over
                  [0.25, 0.75]^2 (the FAC_staticrefinement-style 2-D case);
  deck_3d_c1()    3-D unit cube, 128^3 base (64^3 patches), three levels at
                  ratio 2, cells tagged where |grad u| exceeds the 70th
                  percentile of the level, clustered into boxes of 16^3..64^3;
  deck_3d_c2(s)   3-D, four levels at ratio 2, base 128^3, patches whose
                  per-dimension sizes are drawn from U{8..48}, refined around
                  eight Gaussian sources; D = exp(g) with g a Gaussian random
                  field (correlation length 0.1, sigma 1; grf_modes / grf_eval)
                  and f the sum of the eight Gaussians of width 0.02.

Boxes are ((lo...), (hi...)) in each level's index space.
"""
import numpy as np

PI = np.pi


def _split(n, size):
    return [(s, min(s + size, n) - 1) for s in range(0, n, size)]


def deck_2d_c0(n=64):
    L0 = [((0, 0), (n - 1, n - 1))]
    L1 = [((n // 2, n // 2), (3 * n // 2 - 1, 3 * n // 2 - 1))]
    return dict(name="c0-2d-%d" % n, dim=2, levels=[L0, L1], ratio=2, dlo=(0, 0),
                dhi=(n - 1, n - 1))


def _grad_sine_mag(n, lo=None, hi=None):
    """|grad u| for u = prod sin(pi x_d) at the cell centres of an n^3 grid."""
    x = (np.arange(n) + 0.5) / n
    s, c = np.sin(PI * x), np.cos(PI * x)
    sx, sy, sz = s[:, None, None], s[None, :, None], s[None, None, :]
    cx, cy, cz = c[:, None, None], c[None, :, None], c[None, None, :]
    g2 = (cx * sy * sz) ** 2 + (sx * cy * sz) ** 2 + (sx * sy * cz) ** 2
    return PI * np.sqrt(g2)


def _erode1(mask):
    """Cells whose 3x3x3 neighbourhood is covered (outside the domain counts
    as covered): the proper-nesting buffer of one coarse cell."""
    m = np.pad(mask, 1, constant_values=True)
    out = np.ones_like(mask)
    n0, n1, n2 = mask.shape
    for a in (0, 1, 2):
        for b in (0, 1, 2):
            for c in (0, 1, 2):
                out &= m[a:a + n0, b:b + n1, c:c + n2]
    return out


def _cover_mask(boxes, n):
    m = np.zeros((n, n, n), dtype=bool)
    for lo, hi in boxes:
        m[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = True
    return m


def _greedy_boxes(blocks, maxb):
    """Greedy rectangular clustering of a boolean block grid into boxes of
    at most maxb blocks per side (x, then y, then z growth)."""
    nb = blocks.shape
    taken = np.zeros_like(blocks)
    out = []
    for k in range(nb[2]):
        for j in range(nb[1]):
            for i in range(nb[0]):
                if not blocks[i, j, k] or taken[i, j, k]:
                    continue
                ok = lambda a, b, c, d, e, f: bool(np.all(blocks[a:b, c:d, e:f])) and not bool(
                    np.any(taken[a:b, c:d, e:f]))
                i1 = i + 1
                while i1 - i < maxb and i1 < nb[0] and ok(i1, i1 + 1, j, j + 1, k, k + 1):
                    i1 += 1
                j1 = j + 1
                while j1 - j < maxb and j1 < nb[1] and ok(i, i1, j1, j1 + 1, k, k + 1):
                    j1 += 1
                k1 = k + 1
                while k1 - k < maxb and k1 < nb[2] and ok(i, i1, j, j1, k1, k1 + 1):
                    k1 += 1
                taken[i:i1, j:j1, k:k1] = True
                out.append(((i, j, k), (i1 - 1, j1 - 1, k1 - 1)))
    return out


def _carve(box, allowed, tags):
    """Pieces of box lying in the allowed mask that hold every tag of box:
    the box itself when it is allowed throughout, else its two halves
    (longest side) carved in turn; pieces without tags are dropped."""
    lo, hi = box
    sl = tuple(slice(l, h + 1) for l, h in zip(lo, hi))
    if not tags[sl].any():
        return []
    if allowed[sl].all():
        return [box]
    ext = [h - l + 1 for l, h in zip(lo, hi)]
    d = int(np.argmax(ext))
    mid = lo[d] + ext[d] // 2
    hi1 = list(hi)
    hi1[d] = mid - 1
    lo2 = list(lo)
    lo2[d] = mid
    return _carve((lo, tuple(hi1)), allowed, tags) + _carve((tuple(lo2), hi), allowed, tags)


def deck_3d_c1(base=128, nlevels=3, percentile=70.0, min_size=16, max_size=64,
               base_patch=64, clustering="berger_rigoutsos"):
    """The 3-D plain-FP64 deck (sine solution, D = 1).

    Level ln+1 is built on level ln's index space: cells where |grad u|
    exceeds the percentile (over the cells level ln covers) are tagged inside
    the proper-nesting region (level ln's cover eroded by one cell), the tags
    are clustered by Berger-Rigoutsos (efficiency and combine tolerances 0.8,
    boxes of min_size / r .. max_size / r coarse cells; the library's
    sfac_berger_rigoutsos), boxes leaving the nesting region are carved to
    it, and the boxes are refined by r.  clustering="greedy" instead tiles
    tagged min_size blocks into boxes (no library call)."""
    r = 2
    levels = [[((lo0, lo1, lo2), (hi0, hi1, hi2))
               for (lo2, hi2) in _split(base, base_patch)
               for (lo1, hi1) in _split(base, base_patch)
               for (lo0, hi0) in _split(base, base_patch)]]
    cover = np.ones((base, base, base), dtype=bool)
    n = base
    blk = min_size // r  # block edge in coarse cells
    maxb = max_size // min_size
    for ln in range(1, nlevels):
        g = _grad_sine_mag(n)
        thr = np.percentile(g[cover], percentile)
        tags = (g > thr) & cover
        allowed = _erode1(cover)
        if clustering == "berger_rigoutsos":
            from samrai_amd import fac
            tags &= allowed
            bx = fac.berger_rigoutsos(tags, (0, 0, 0), 0.8, 0.8, (blk,) * 3,
                                      (max_size // r,) * 3)
            coarse = [p for b in bx for p in _carve(b, allowed, tags)]
            fine = [(tuple(c * r for c in lo), tuple((c + 1) * r - 1 for c in hi))
                    for lo, hi in coarse]
        else:
            nb = n // blk
            tb = tags.reshape(nb, blk, nb, blk, nb, blk).any(axis=(1, 3, 5))
            eb = allowed.reshape(nb, blk, nb, blk, nb, blk).all(axis=(1, 3, 5))
            bx = _greedy_boxes(tb & eb, maxb)
            fine = [(tuple(c * min_size for c in lo), tuple((c + 1) * min_size - 1 for c in hi))
                    for lo, hi in bx]
        levels.append(fine)
        n *= r
        cover = _cover_mask(fine, n)
    return dict(name="c1-3d", dim=3, levels=levels, ratio=r, dlo=(0, 0, 0),
                dhi=(base - 1,) * 3)


def _random_partition(rng, lo, hi, smin, smax, align=2):
    """Segments of [lo, hi] with lengths drawn from U{smin..smax} (even)."""
    segs = []
    s = lo
    while s <= hi:
        ln = int(rng.integers(smin // align, smax // align + 1)) * align
        e = min(s + ln, hi + 1) - 1
        segs.append((s, e))
        s = e + 1
    return segs


def deck_3d_c2(seed=0, base=128, nlevels=4, nsources=8, smin=8, smax=48):
    """The batched variable-coefficient deck: refinement around eight
    Gaussian sources, patch sides drawn from U{8..48}."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.1, 0.9, (nsources, 3))
    amps = rng.uniform(-1.0, 1.0, nsources)
    r = 2
    half = [None, 0.22, 0.11, 0.05][:nlevels]
    levels = []
    # level 0: random partition of the whole domain
    segs = [_random_partition(rng, 0, base - 1, smin, smax) for _ in range(3)]
    levels.append([((a[0], b[0], c[0]), (a[1], b[1], c[1]))
                   for c in segs[2] for b in segs[1] for a in segs[0]])
    n = base
    for ln in range(1, nlevels):
        n *= r
        h = half[ln]
        segs = [_random_partition(rng, 0, n - 1, smin, smax) for _ in range(3)]
        # which partition cells meet a source cube of half-width h
        cubes = [(np.floor((c - h) * n).astype(int), np.ceil((c + h) * n).astype(int) - 1)
                 for c in centres]
        boxes = []
        for c in segs[2]:
            for b in segs[1]:
                for a in segs[0]:
                    lo = np.array([a[0], b[0], c[0]])
                    hi = np.array([a[1], b[1], c[1]])
                    for clo, chi in cubes:
                        if np.all(lo <= chi) and np.all(hi >= clo):
                            boxes.append((tuple(int(v) for v in lo), tuple(int(v) for v in hi)))
                            break
        levels.append(boxes)
    d = dict(name="c2-3d-s%d" % seed, dim=3, levels=levels, ratio=r, dlo=(0, 0, 0),
             dhi=(base - 1,) * 3, centres=centres, amps=amps, seed=seed)
    return d


def count_cells(deck):
    return sum(int(np.prod([h - l + 1 for l, h in zip(lo, hi)]))
               for boxes in deck["levels"] for lo, hi in boxes)


def level_dx(deck, ln):
    n0 = np.array(deck["dhi"]) - np.array(deck["dlo"]) + 1
    return (1.0 / n0) / (deck["ratio"] ** ln)


def cell_centres(box, dx):
    lo, hi = box
    return np.meshgrid(*[(np.arange(lo[d], hi[d] + 1) + 0.5) * dx[d] for d in range(len(lo))],
                       indexing="ij")


def sine_rhs(deck, ln, box):
    """f = -dim pi^2 prod sin(pi x_d) (SAMRAI sign: div grad u = f)."""
    X = cell_centres(box, level_dx(deck, ln))
    dim = deck["dim"]
    return -dim * PI ** 2 * np.prod([np.sin(PI * x) for x in X], axis=0)


def sine_exact(deck, ln, box):
    X = cell_centres(box, level_dx(deck, ln))
    return np.prod([np.sin(PI * x) for x in X], axis=0)


def gaussian_rhs(deck, ln, box, width=0.02):
    X = cell_centres(box, level_dx(deck, ln))
    f = np.zeros(X[0].shape)
    for c, a in zip(deck["centres"], deck["amps"]):
        r2 = sum((X[d] - c[d]) ** 2 for d in range(3))
        f += a * np.exp(-r2 / (2 * width ** 2))
    return f


def grf_modes(seed, corr=0.1, nmodes=256):  # nmodes: a power of two
    """Random Fourier features of a stationary Gaussian random field with
    covariance exp(-|r|^2 / (2 corr^2)) and unit variance:
    g(x) = sqrt(2 / M) sum_m cos(k_m . x + phi_m), k_m ~ N(0, I / corr^2),
    phi_m ~ U[0, 2 pi)."""
    rng = np.random.default_rng(1000 + seed)
    k = rng.standard_normal((nmodes, 3)) / corr
    phi = rng.uniform(0.0, 2 * PI, nmodes)
    return k, phi


def grf_eval(points, modes, xp=np):
    """g at points (..., 3) with the array module xp (numpy or torch)."""
    k, phi = modes
    if xp is not np:
        k = xp.as_tensor(k, dtype=points.dtype, device=points.device)
        phi = xp.as_tensor(phi, dtype=points.dtype, device=points.device)
    # elementwise only (no matrix product, no library reduction, a fixed
    # pairwise sum over the modes): each point's value is the same bits
    # whatever array it is evaluated in, so patches sharing a face agree
    arg = (points[..., 0:1] * k[:, 0] + points[..., 1:2] * k[:, 1] +
           points[..., 2:3] * k[:, 2] + phi)
    c = xp.cos(arg)
    while c.shape[-1] > 1:
        h = c.shape[-1] // 2
        c = c[..., :h] + c[..., h:2 * h]
    return (2.0 / k.shape[0]) ** 0.5 * c[..., 0]


def face_points(deck, ln, box, axis, xp=np, **kw):
    """Face-centre coordinates (n0', n1', n2', 3) of the side data of `axis`
    on `box` (i fastest when flattened in Fortran order)."""
    lo, hi = box
    dx = level_dx(deck, ln)
    axes = []
    for d in range(3):
        if d == axis:
            axes.append(xp.arange(lo[d], hi[d] + 2, **kw) * dx[d])
        else:
            axes.append((xp.arange(lo[d], hi[d] + 1, **kw) + 0.5) * dx[d])
    g = xp.meshgrid(*axes, indexing="ij")
    return xp.stack(g, -1)


def cell_points(deck, ln, box, xp=np, **kw):
    lo, hi = box
    dx = level_dx(deck, ln)
    axes = [(xp.arange(lo[d], hi[d] + 1, **kw) + 0.5) * dx[d] for d in range(3)]
    return xp.stack(xp.meshgrid(*axes, indexing="ij"), -1)


def gaussian_sources(deck, points, width=0.02, xp=np):
    """f = sum_s a_s exp(-|x - c_s|^2 / (2 width^2)) at points (..., 3)."""
    f = 0.0 * points[..., 0]
    for c, a in zip(deck["centres"], deck["amps"]):
        r2 = ((points[..., 0] - float(c[0])) ** 2 + (points[..., 1] - float(c[1])) ** 2 +
              (points[..., 2] - float(c[2])) ** 2)
        f = f + float(a) * xp.exp(-r2 / (2 * width ** 2))
    return f


def build_sine_problem(fac, deck, params, dconst=1.0):
    """Hierarchy, u (ghost 1, zero), f = the sine RHS (ghost 0) and a solver
    with D constant, Dirichlet 0 and the given input keys, initialized."""
    h = fac.PatchHierarchy(deck["dim"], deck["dlo"], deck["dhi"], [0.0] * 3, [1.0] * 3)
    for boxes in deck["levels"]:
        h.add_level(deck["ratio"], boxes)
    h.finalize()
    u = fac.PatchData(h, 1)
    f = fac.PatchData(h, 0)
    for ln, boxes in enumerate(deck["levels"]):
        for p, box in enumerate(boxes):
            f.put(ln, p, sine_rhs(deck, ln, box))
    s = fac.CellPoissonFACSolver(h, **params)
    s.set_d_constant(dconst)
    s.set_dirichlet(0.0)
    s.initialize_solver_state(u, f)
    return h, u, f, s


def build_c2_problem(fac, torch, params, seed=0, deck=None):
    """The variable-coefficient deck on the GPU: D = exp(GRF) on the faces
    and f = eight Gaussian sources, evaluated with torch on the device and
    copied device to device into the patch data; u = 0; Dirichlet 0.
    Returns (deck, hierarchy, u, f, D, solver)."""
    deck = deck or deck_3d_c2(seed)
    h = fac.PatchHierarchy(3, deck["dlo"], deck["dhi"], [0.0] * 3, [1.0] * 3)
    for boxes in deck["levels"]:
        h.add_level(deck["ratio"], boxes)
    h.finalize()
    u = fac.PatchData(h, 1)
    f = fac.PatchData(h, 0)
    D = fac.PatchData(h, 0, "side")
    modes = grf_modes(deck["seed"])
    kw = dict(dtype=torch.float64, device="cuda")
    for ln, boxes in enumerate(deck["levels"]):
        for p, box in enumerate(boxes):
            fv = gaussian_sources(deck, cell_points(deck, ln, box, xp=torch, **kw), xp=torch)
            f.put_device(ln, p, fv.permute(2, 1, 0).contiguous())
            for a in range(3):
                g = grf_eval(face_points(deck, ln, box, a, xp=torch, **kw), modes, xp=torch)
                D.put_device(ln, p, torch.exp(g).permute(2, 1, 0).contiguous(), axis=a)
    s = fac.CellPoissonFACSolver(h, **params)
    s.set_d_patch_data(D)
    s.set_dirichlet(0.0)
    s.initialize_solver_state(u, f)
    return deck, h, u, f, D, s
