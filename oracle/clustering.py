"""Berger-Rigoutsos clustering restated in Python (test infrastructure only).

Oracle for the library's host clustering (sfac_berger_rigoutsos).  Follows
mesh::BergerRigoutsosNode for one process, with the reference defaults
DEV_max_inflection_cut_from_center = 1, DEV_inflection_cut_threshold_ar = 0,
DEV_min_box_size_from_cutting = 0:
  computeMinimalBoundingBoxForTags   BergerRigoutsosNode.C:1251-1325
  acceptOrSplitBox                   BergerRigoutsosNode.C:1326-1630
  findZeroCutSwath                   BergerRigoutsosNode.C:1636-1716
  cutAtInflection                    BergerRigoutsosNode.C:1718-1812
  runChildren_check (recombination)  BergerRigoutsosNode.C:725-760
Parity with the reference's own output is unpinned (the reference cannot be
built here); the tests pin this restatement by hand-worked cases and by the
invariants every clustering must satisfy.
"""
import numpy as np


def _hist(tags, lo, hi):
    sub = tags[tuple(slice(l, h + 1) for l, h in zip(lo, hi))]
    dim = tags.ndim
    return [sub.sum(axis=tuple(a for a in range(dim) if a != d)).astype(int) for d in range(dim)]


def _zero_swath(h, lo, hi, minb):
    lo_lim, hi_lim = lo + minb, hi - minb
    cl = ch = (lo + hi) // 2
    while cl >= lo_lim and ch <= hi_lim:
        if h[cl - lo] == 0:
            ch = cl
            while cl > lo_lim and h[cl - lo - 1] == 0:
                cl -= 1
            return cl, ch
        if h[ch - lo] == 0:
            cl = ch
            while ch < hi_lim and h[ch - lo + 1] == 0:
                ch += 1
            return cl, ch
        cl -= 1
        ch += 1
    return None


def _inflection(h, lo, minb):
    n = len(h)
    if n < 4:
        cut = 1
        for i in range(2, n):
            if abs(h[cut] - h[cut - 1]) < abs(h[i] - h[i - 1]):
                cut = i
        return cut + lo, 0
    mid = n // 2
    maxd = int(1.0 * n / 2)
    lo_lim = max(minb, mid - maxd)
    hi_lim = min(n - 1 - minb + 1, mid + maxd)
    lap = lambda i: h[i - 1] - 2 * h[i] + h[i + 1]
    cut, inf = mid, abs(lap(mid) - lap(mid - 1))
    cl, ch = mid - 1, mid + 1
    while cl > lo_lim or ch < hi_lim:
        for c, ok in ((cl, cl > lo_lim), (ch, ch < hi_lim)):
            if ok:
                la, lb = lap(c), lap(c - 1)
                if la * lb <= 0 and abs(la - lb) > inf:
                    cut, inf = c, abs(la - lb)
        cl -= 1
        ch += 1
    return cut + lo, inf


def cluster(tags, lo0=None, efficiency_tolerance=0.8, combine_efficiency=0.8, min_box=None,
            max_box=None):
    """Boxes ((lo...), (hi...)) covering the True cells of tags (indices
    offset by lo0)."""
    tags = np.asarray(tags, dtype=bool)
    dim = tags.ndim
    lo0 = tuple(lo0) if lo0 is not None else (0,) * dim
    minb = [max(1, m) for m in (min_box or (1,) * dim)]
    maxb = list(max_box or (1 << 30,) * dim)
    out = []

    def node(lo, hi):
        lo, hi = list(lo), list(hi)
        h = _hist(tags, lo, hi)
        for d in range(dim):
            size, beg, end = hi[d] - lo[d] + 1, 0, hi[d] - lo[d]
            while h[d][beg] == 0 and size > minb[d]:
                lo[d] += 1
                beg += 1
                size -= 1
            while h[d][end] == 0 and size > minb[d]:
                hi[d] -= 1
                end -= 1
                size -= 1
            h[d] = h[d][beg:end + 1]
        ntag = int(h[0].sum())
        dims = [hi[d] - lo[d] + 1 for d in range(dim)]
        vol = int(np.prod(dims))
        oversize = any(dims[d] > maxb[d] for d in range(dim))
        if not oversize:
            if (ntag / vol if vol else 1.0) >= efficiency_tolerance:
                out.append((tuple(lo), tuple(hi)))
                return True
            if ntag == 0:
                return False
        margin = [dims[d] - 2 * minb[d] for d in range(dim)]
        if all(m < 0 for m in margin):
            out.append((tuple(lo), tuple(hi)))
            return True
        order = list(range(dim))
        for a in range(dim - 1):
            for b in range(a + 1, dim):
                if margin[order[a]] < margin[order[b]]:
                    order[a], order[b] = order[b], order[a]
        left_hi, right_lo = list(hi), list(lo)
        for d in order:
            if margin[d] < 0:
                continue
            sw = _zero_swath(h[d], lo[d], hi[d], minb[d])
            if sw is not None:
                left_hi[d], right_lo[d] = sw[0] - 1, sw[1] + 1
                break
        else:
            leave = dims[order[0]] - 1
            best, cdir, cpt = -1, 0, 0
            for d in range(dim):
                if margin[d] < 0 or dims[d] <= leave:
                    continue
                pt, inf = _inflection(h[d], lo[d], minb[d])
                if best < inf or (best == inf and margin[d] > margin[cdir]):
                    best, cdir, cpt = inf, d, pt
            left_hi[cdir], right_lo[cdir] = cpt - 1, cpt
        mark = len(out)
        la = node(lo, left_hi)
        lbox = out[-1] if la else None
        ra = node(right_lo, hi)
        rbox = out[-1] if ra else None
        if la and ra and not oversize:
            cells = lambda b: int(np.prod([b[1][d] - b[0][d] + 1 for d in range(dim)]))
            if (cells(lbox) + cells(rbox)) / vol >= combine_efficiency:
                del out[mark:]
                out.append((tuple(lo), tuple(hi)))
                return True
        return False

    node([0] * dim, [s - 1 for s in tags.shape])
    return [(tuple(l + o for l, o in zip(b[0], lo0)), tuple(h + o for h, o in zip(b[1], lo0)))
            for b in out]
