"""Berger-Rigoutsos clustering (host; sfac_berger_rigoutsos).

The library's clustering follows mesh::BergerRigoutsosNode for one process
(BergerRigoutsosNode.C:725-760, 1251-1812).  The reference cannot run here,
so the checks are: a hand-worked cut (the L-shape below, worked through the
histogram Laplacians by hand in the comment), a zero-swath cut between two
separated blocks, the invariants every clustering must satisfy (every
tagged cell covered, boxes disjoint, inside the tag box, within max_box
unless no cut can respect min_box, no cut below min_box, no box without
tags), and equality with the Python
restatement oracle/clustering.py on seeded random tag fields.
"""
import numpy as np
import pytest

from oracle.clustering import cluster


def _fac():
    from samrai_amd import fac
    return fac


def test_l_shape_is_cut_at_the_inflection():
    # tags: x in [0,1] (all y) and y in [0,1] (all x) on 8 x 8.  Histograms
    # along x and y: [8, 8, 2, 2, 2, 2, 2, 2]; Laplacians h[i-1]-2h[i]+h[i+1]:
    # i=1: -6, i=2: 6, i=3..6: 0.  The strongest sign change is between i=1
    # and i=2 (|6 - (-6)| = 12), so the cut falls at x = 2 (tie with y: x has
    # the same margin and comes first).  Left [0,1] x [0,7] is full; right
    # [2,7] x [0,7] shrinks to [2,7] x [0,1], full; recombination would need
    # (16 + 12) / 64 >= 0.8, so both stay.
    t = np.zeros((8, 8), bool)
    t[0:2, :] = True
    t[:, 0:2] = True
    want = [((0, 0), (1, 7)), ((2, 0), (7, 1))]
    assert cluster(t) == want
    assert _fac().berger_rigoutsos(t) == want


def test_zero_swath_separates_blocks():
    t = np.zeros((20, 10), bool)
    t[2:5, 1:4] = True
    t[12:16, 3:9] = True
    want = [((2, 1), (4, 3)), ((12, 3), (15, 8))]
    assert cluster(t) == want
    assert _fac().berger_rigoutsos(t) == want
    # with an index offset
    got = _fac().berger_rigoutsos(t, lo=(-7, 5))
    assert got == [((-5, 6), (-3, 8)), ((5, 8), (8, 13))]


def _random_tags(rng, dim):
    shp = tuple(int(x) for x in rng.integers(4, 20 if dim == 3 else 40, dim))
    t = np.zeros(shp, bool)
    for _ in range(int(rng.integers(1, 5))):
        c = [int(rng.integers(0, s)) for s in shp]
        r = [int(rng.integers(1, 6)) for _ in shp]
        sl = tuple(slice(max(0, c[d] - r[d]), min(shp[d], c[d] + r[d] + 1)) for d in range(dim))
        t[sl] = rng.random(t[sl].shape) < 0.8
    return t


@pytest.mark.parametrize("dim", [2, 3])
def test_matches_restatement_and_invariants(dim):
    rng = np.random.default_rng(11 + dim)
    for trial in range(60):
        t = _random_tags(rng, dim)
        mn = tuple(int(rng.integers(1, 4)) for _ in range(dim))
        mx = tuple(int(rng.integers(4, 16)) for _ in range(dim))
        lo = tuple(int(rng.integers(-5, 5)) for _ in range(dim))
        eff, comb = [(0.8, 0.8), (0.5, 0.9), (0.95, 0.6)][trial % 3]
        got = _fac().berger_rigoutsos(t, lo, eff, comb, mn, mx)
        assert got == cluster(t, lo, eff, comb, mn, mx), (trial, got)
        cover = np.zeros(t.shape, int)
        for blo, bhi in got:
            sl = tuple(slice(blo[d] - lo[d], bhi[d] - lo[d] + 1) for d in range(dim))
            side = [bhi[d] - blo[d] + 1 for d in range(dim)]
            for d in range(dim):
                assert 0 <= blo[d] - lo[d] <= bhi[d] - lo[d] < t.shape[d]
                assert side[d] >= min(mn[d], t.shape[d])
            # a box over max_box is kept only when no direction can be cut
            # into two boxes of the minimum size (acceptOrSplitBox)
            if any(side[d] > mx[d] for d in range(dim)):
                assert all(side[d] < 2 * mn[d] for d in range(dim))
            assert t[sl].any()
            cover[sl] += 1
        assert cover.max() <= 1          # disjoint
        assert np.all(cover[t] == 1)     # every tag covered


def test_rejects_bad_parameters():
    fac = _fac()
    t = np.ones((4, 4), bool)
    with pytest.raises(fac.SolverError):
        fac.berger_rigoutsos(t, efficiency_tolerance=1.5)
    with pytest.raises(fac.SolverError):
        fac.berger_rigoutsos(t, min_box=(4, 4), max_box=(2, 2))
