"""CPU checks that pin the triplet oracle (3iso, 3ani; periodic or not).

* cpp32 3iso vs the reference's compiled CPU program (cpp/3PCFisotropic/normal)
  on its hard-coded configuration (N=10,000, 30 bins, dmax 50): the program
  prints 6 significant digits, so the oracle's values are compared through the
  same formatting.
* the cell-grid enumeration vs the literal 64-combination projection formula
  of the reference (brute force over all ordered triplets).
* structural identities: XXX is symmetric under permutations of its three side
  bins; XXY is symmetric in its last two; the 3ani histogram summed over the two
  angle bins is the 3iso histogram with the d23/d31 axes exchanged.
* the reference's own relation between its CPU and GPU layouts:
  CPU XXX = GPU XXX, CPU XXY = Sym6(GPU XXY)/6, with FP32 symmetrisation.
"""
import itertools
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import GOLDEN, catalog
from desi_amd import synth


def _tokens(path):
    return open(path).read().split()


def _fmt6(v):
    # std::ostream default: 6 significant digits, %g style
    return ["%g" % x for x in v]


@pytest.mark.parametrize("which", ["DDD", "RRR", "DDR", "DRR"])
def test_cpp32_3iso_matches_compiled_reference(which):
    d = catalog("data_10K.dat")
    r = catalog("rand0_10K.dat")
    args = {"DDD": (d, None), "RRR": (r, None), "DDR": (d, r), "DRR": (r, d)}[which]
    h = O.cpp32_3iso(*args, bins=30, dmax=50.0)
    ref = _tokens(os.path.join(GOLDEN, "ref_out", "3iso_normal", f"{which}iso_mesh_3D_ref.dat"))
    assert len(ref) == 27000
    assert _fmt6(h) == ref


def _sym6(h, n):
    h = h.reshape(n, n, n)
    return sum(np.transpose(h, p) for p in itertools.permutations(range(3)))


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("bpc", [False, True])
@pytest.mark.parametrize("auto", [True, False])
def test_triplet_grid_equals_brute(ani, bpc, auto):
    n = 36 if bpc else 70
    a = synth.uniform(n, 40.0, seed=41, weights="uniform")
    b = synth.uniform(n - 10, 40.0, seed=42, weights="uniform")
    bins = 5 if ani else 7
    for dmax in (9.0, 17.0):
        c1, w1 = O.triplets(a, None if auto else b, ani=ani, bpc=bpc, box=40.0, bins=bins,
                            dmax=dmax)
        c2, w2 = O.triplets(a, None if auto else b, ani=ani, bpc=bpc, box=40.0, bins=bins,
                            dmax=dmax, brute=True)
        assert c1.sum() > 0
        assert np.array_equal(c1, c2)
        assert np.allclose(w1, w2, rtol=1e-13, atol=0)


@pytest.mark.parametrize("bpc", [False, True])
def test_triplet_symmetries_and_ani_marginal(bpc):
    a = synth.neyman_scott(12, 16, 3.0, 60.0, seed=5)
    b = synth.uniform(150, 60.0, seed=6)
    n = 8
    xxx, _ = O.triplets(a, bpc=bpc, box=60.0, bins=n, dmax=20.0)
    h = xxx.reshape(n, n, n)
    for p in itertools.permutations(range(3)):
        assert np.array_equal(h, np.transpose(h, p))
    xxy, _ = O.triplets(a, b, bpc=bpc, box=60.0, bins=n, dmax=20.0)
    g = xxy.reshape(n, n, n)
    assert np.array_equal(g, np.transpose(g, (0, 2, 1)))
    # 3ani summed over the angle bins: layout [d12][d31][d23] vs iso [d12][d23][d31].
    # The angle bins drop a triplet only when cos = 1 exactly (t or p = bins).
    ani, _ = O.triplets(a, bpc=bpc, ani=True, box=60.0, bins=n, dmax=20.0)
    marg = ani.reshape(n, n, n, n, n).sum(axis=(3, 4))
    iso_swapped = np.transpose(h, (0, 2, 1))
    assert np.all(marg <= iso_swapped)
    assert marg.sum() >= iso_swapped.sum() * 0.999


def test_cpu_layout_relations():
    # CPU XXX == GPU XXX and CPU XXY == Sym6(GPU XXY)/6 up to the FP32
    # symmetrisation, when both use the same arithmetic (lattice: exact).
    a = synth.lattice(300, 32.0, 1.0, seed=8)
    b = synth.lattice(250, 32.0, 1.0, seed=9)
    n = 6
    gpu_xxx, _ = O.triplets(a, bins=n, dmax=6.0)
    cpu_xxx = O.cpp32_3iso(a, bins=n, dmax=6.0)
    # The CPU program does not exclude coincident points (d = 0); the lattice
    # has duplicates, so compare after removing them.
    _, idx = np.unique(a[:, :3], axis=0, return_index=True)
    au = a[np.sort(idx)]
    gpu_xxx, _ = O.triplets(au, bins=n, dmax=6.0)
    cpu_xxx = O.cpp32_3iso(au, bins=n, dmax=6.0)
    assert np.array_equal(gpu_xxx.astype(np.float64), cpu_xxx)
    _, idx = np.unique(b[:, :3], axis=0, return_index=True)
    bu = b[np.sort(idx)]
    gpu_xxy, _ = O.triplets(au, bu, bins=n, dmax=6.0)
    cpu_xxy = O.cpp32_3iso(au, bu, bins=n, dmax=6.0)
    sym = _sym6(gpu_xxy.astype(np.float64), n).ravel() / 6.0
    third = float(np.float32(1.0) / np.float32(3.0))
    expect = (np.round(sym * 3) * third).astype(np.float32).astype(np.float64)
    assert np.array_equal(cpu_xxy, expect)


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("auto", [True, False])
def test_triplet_grid_equals_brute_wide_periodic(ani, auto):
    """Periodic with 2*dmax >= box (dmax < box): a point can count under its
    regular and its projected image on one axis, and the grid enumeration
    (one cell spanning the box, neighbour images -1, 0, +1) must reproduce
    the literal 64-combination formula."""
    a = synth.uniform(26, 20.0, seed=43, weights="uniform")
    b = synth.uniform(21, 20.0, seed=44, weights="uniform")
    bins = 4 if ani else 6
    for dmax in (10.5, 13.0, 19.0):
        c1, w1 = O.triplets(a, None if auto else b, ani=ani, bpc=True, box=20.0, bins=bins,
                            dmax=dmax)
        c2, w2 = O.triplets(a, None if auto else b, ani=ani, bpc=True, box=20.0, bins=bins,
                            dmax=dmax, brute=True)
        assert c1.sum() > 0
        assert np.array_equal(c1, c2)
        assert np.allclose(w1, w2, rtol=1e-13, atol=0)


def test_wide_periodic_triplet_by_hand():
    """Three points on a line in a box of side 10 with dmax 8 (> box / 2):
    A = 5, B = 2, C = 8.  Per axis the reference takes |dx| - L*f for the 1-2
    and 1-3 sides (f = 1: the projected image) and |x3 - x2| - L*|f2 - f3| for
    the 2-3 side (3isoBPC_k.cuh:56-75), so from A the pair (B, C) counts four
    ways: (f2, f3) = (0,0) sides 3, 6, 3; (1,0) 7, 4, 3; (0,1) 3, 4, 7; and
    (1,1) 7, 6, 7, where the 2-3 side is |8 - 2| although the two projected
    images lie 14 apart.  Worked out for all six orders by hand, the 24 ordered
    (triplet, combination) entries are each permutation of {3, 6, 3}, of
    {7, 4, 3} and of {7, 6, 7}, twice each (bins = floor(d), ds = 1)."""
    pts = np.array([[5, 1, 1, 1], [2, 1, 1, 1], [8, 1, 1, 1]], np.float32)
    c, w = O.triplets(pts, bpc=True, box=10.0, bins=8, dmax=8.0)
    h = c.reshape(8, 8, 8)
    expect = np.zeros((8, 8, 8), np.uint64)
    for tri in ((3, 6, 3), (7, 4, 3), (7, 6, 7)):
        for p in set(itertools.permutations(tri)):
            expect[p] += 2
    assert h.sum() == 24
    assert np.array_equal(h, expect), np.argwhere(h != expect)
    assert np.array_equal(w.reshape(8, 8, 8), expect.astype(np.float64))
    # At dmax 4.5 (< box / 2) the three points form no triplet, although B
    # and C are 4 apart through the wall: from A both are at their regular
    # images (f = 0), so the 2-3 side is |8 - 2| = 6; from B (or C) the third
    # point is projected (f3 = 1) and the 2-3 side becomes |8 - 5| - 10.
    c2, _ = O.triplets(pts, bpc=True, box=10.0, bins=9, dmax=4.5)
    assert c2.sum() == 0
