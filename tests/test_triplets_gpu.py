"""GPU parity of the triplet histograms (3iso, 3ani; periodic or not) against the
FP64 oracle through the C ABI: counts bit-exact, weighted sums within 1e-12
relative.  Plus the reference's own GPU outputs (results/*.example)."""
import itertools
import os

import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O
from conftest import GOLDEN, catalog

pytestmark = pytest.mark.gpu

WREL = 1e-12


def _run(engine, a, b, *, ani, bpc, box, bins, dmax, cpr=0, prec=da.PCF_PREC_MIXED):
    ca = engine.upload(a)
    cb = engine.upload(b) if b is not None else -1
    try:
        job = da.make_job(da.PCF_3ANI if ani else da.PCF_3ISO,
                          da.PCF_XXX if b is None else da.PCF_XXY, bins=bins, dmax=dmax,
                          cat_x=ca, cat_y=cb, bpc=bpc, box=box, cells_per_rmax=cpr,
                          precision=prec)
        c, w = engine.histogram(job)
        st = engine.stats()
    finally:
        engine.release(ca)
        if cb >= 0:
            engine.release(cb)
    return c, w, st


def _check(engine, a, b, *, ani, bpc, box, bins, dmax, cpr=0, prec=da.PCF_PREC_MIXED):
    print(f"case 3{'ani' if ani else 'iso'} bpc={bpc} n={len(a)} cross={b is not None} "
          f"bins={bins} dmax={dmax} cpr={cpr} prec={prec}", flush=True)
    c, w, st = _run(engine, a, b, ani=ani, bpc=bpc, box=box, bins=bins, dmax=dmax, cpr=cpr,
                    prec=prec)
    co, wo = O.triplets(a, b, ani=ani, bpc=bpc, box=box, bins=bins, dmax=dmax)
    bad = np.nonzero(c != co)[0]
    assert len(bad) == 0, (f"{len(bad)} bins differ, e.g. {bad[:8]}: gpu {c[bad[:8]]} "
                           f"oracle {co[bad[:8]]}; sums {c.sum()} vs {co.sum()}")
    tol = WREL * np.maximum(np.abs(wo), 1e-300)
    assert np.all(np.abs(w - wo) <= tol), "weighted sums differ"
    assert co.sum() > 0
    return st


def test_empty_and_tiny(engine):
    """Catalogs of 0 to 4 points (one duplicated point: its zero separation is
    excluded) as XXX and XXY, both ends empty included: every histogram equals
    the oracle's, all-zero ones too."""
    one = np.array([[3, 2, 2, 1.5]], np.float32)
    three = np.array([[1, 1, 1, 1], [2, 1.5, 1, 0.5], [1.5, 3, 2, 2]], np.float32)
    dup = np.ascontiguousarray(np.concatenate([three, three[:1]]))
    empty = np.zeros((0, 4), np.float32)
    cases = ((one, None), (three[:2], None), (three, None), (dup, None), (three, one),
             (one, three), (three[:2], one), (three, empty), (empty, three))
    nonzero = 0
    for ani in (False, True):
        for bpc in (False, True):
            for a, b in cases:
                print(f"case 3{'ani' if ani else 'iso'} bpc={bpc} n={len(a)} "
                      f"m={None if b is None else len(b)}", flush=True)
                c, w, _ = _run(engine, a, b, ani=ani, bpc=bpc, box=50.0, bins=5, dmax=10.0)
                co, wo = O.triplets(a, b, ani=ani, bpc=bpc, box=50.0, bins=5, dmax=10.0)
                assert np.array_equal(c, co)
                assert np.all(np.abs(w - wo) <= WREL * np.abs(wo))
                nonzero += int(co.sum() > 0)
    assert nonzero == 16


@pytest.mark.parametrize("cross", [False, True])
@pytest.mark.parametrize("prec", [da.PCF_PREC_MIXED, da.PCF_PREC_FP64])
def test_3iso_many_bins_global_histogram(engine, cross, prec):
    """Weighted 3iso with 120 bins: one b12 row of the histogram (120 x 120
    bins x 12 bytes) exceeds the LDS budget, so the kernel deposits straight
    into the global histogram; counts and weights equal the oracle's."""
    a = synth.uniform(400, 50.0, seed=5, weights="uniform")
    b = synth.uniform(300, 50.0, seed=6, weights="uniform")
    _check(engine, a, b if cross else None, ani=False, bpc=False, box=50.0, bins=120,
           dmax=12.0, prec=prec)


def test_3ani_more_than_40_bins(engine):
    """3ani with 41 bins per dimension (41^5 = 115,856,201 bins per
    histogram, about 1.9 GB of counts and weights): the global-atomic kernel,
    against the oracle."""
    a = synth.neyman_scott(4, 10, 3.0, 60.0, seed=7, weights="uniform")
    _check(engine, a, None, ani=True, bpc=True, box=60.0, bins=41, dmax=20.0)


@pytest.mark.parametrize("bpc", [False, True])
@pytest.mark.parametrize("cross", [False, True])
def test_3iso_sample(engine, bpc, cross):
    d = catalog("data_1K.dat")
    r = catalog("rand0_1K.dat")
    _check(engine, d, r if cross else None, ani=False, bpc=bpc, box=250.0, bins=30, dmax=80.0)


@pytest.mark.parametrize("bpc", [False, True])
@pytest.mark.parametrize("cross", [False, True])
def test_3ani_sample(engine, bpc, cross):
    d = catalog("data_500.dat")
    r = catalog("rand0_500.dat")
    _check(engine, d, r if cross else None, ani=True, bpc=bpc, box=250.0, bins=15, dmax=60.0)


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("bpc", [False, True])
def test_plain_fp64_path(engine, ani, bpc):
    d = synth.neyman_scott(32, 16, 5.0, 250.0, seed=23, weights="uniform")
    r = synth.uniform(700, 250.0, seed=24, weights="uniform")
    bins = 15 if ani else 30
    _check(engine, d, r, ani=ani, bpc=bpc, box=250.0, bins=bins, dmax=80.0,
           prec=da.PCF_PREC_FP64)


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("bpc", [False, True])
def test_weighted_clustered(engine, ani, bpc):
    d = synth.neyman_scott(64, 32, 5.0, 250.0, seed=21, weights="uniform")
    r = synth.uniform(2048, 250.0, seed=22, weights="uniform")
    bins = 15 if ani else 30
    dmax = 60.0 if ani else 80.0
    _check(engine, d, None, ani=ani, bpc=bpc, box=250.0, bins=bins, dmax=dmax)
    _check(engine, d, r, ani=ani, bpc=bpc, box=250.0, bins=bins, dmax=dmax)
    _check(engine, r, d, ani=ani, bpc=bpc, box=250.0, bins=bins, dmax=dmax)


@pytest.mark.parametrize("ani", [False, True])
def test_pair_list_reuse(engine, ani):
    """Resident catalogs through DDD, DDR, DRR, RRR (the second of each two
    keeps the first's pair list) and then jobs that must not keep it: another
    range, a rebuilt grid, and a new catalog under a released one's id.
    Every histogram equals the oracle's."""
    d = synth.neyman_scott(48, 24, 4.0, 200.0, seed=31, weights="uniform")
    r = synth.uniform(1200, 200.0, seed=32, weights="uniform")
    r2 = synth.uniform(1100, 200.0, seed=33, weights="uniform")
    bins, dmax = (10, 40.0) if ani else (20, 50.0)
    typ = da.PCF_3ANI if ani else da.PCF_3ISO
    cd, cr = engine.upload(d), engine.upload(r)
    cats = {cd: d, cr: r}

    def run(x, y, *, dm=dmax, prec=da.PCF_PREC_MIXED, rebuild=False):
        print(f"case 3{'ani' if ani else 'iso'} x={x} y={y} dmax={dm} prec={prec} "
              f"rebuild={rebuild}", flush=True)
        job = da.make_job(typ, da.PCF_XXX if y < 0 else da.PCF_XXY, bins=bins, dmax=dm,
                          cat_x=x, cat_y=y, bpc=True, box=200.0, precision=prec,
                          rebuild_grid=rebuild)
        c, w = engine.histogram(job)
        co, wo = O.triplets(cats[x], None if y < 0 else cats[y], ani=ani, bpc=True, box=200.0,
                            bins=bins, dmax=dm)
        assert np.array_equal(c, co), f"counts differ: {c.sum()} vs {co.sum()}"
        assert np.all(np.abs(w - wo) <= WREL * np.abs(wo)), "weighted sums differ"
        assert co.sum() > 0

    try:
        run(cd, -1)
        run(cd, cr)
        run(cr, cd)
        run(cr, -1)
        run(cr, -1, prec=da.PCF_PREC_FP64)
        run(cr, -1, dm=dmax * 0.75)
        run(cr, -1, rebuild=True)
        run(cr, cd)
        engine.release(cr)
        cr2 = engine.upload(r2)
        cats[cr2] = r2
        run(cr2, -1)
        run(cr2, cd)
        engine.release(cr2)
    finally:
        engine.release(cd)


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("bpc", [False, True])
def test_lattice_exact_edges(engine, ani, bpc):
    a = synth.lattice(1500, 32.0, 0.5, seed=31)
    b = synth.lattice(1200, 32.0, 0.5, seed=32)
    st = _check(engine, a, None, ani=ani, bpc=bpc, box=32.0, bins=6, dmax=6.0)
    assert st["rechecks"] > 0
    _check(engine, a, b, ani=ani, bpc=bpc, box=32.0, bins=6, dmax=6.0)


@pytest.mark.parametrize("cpr", [1, 2, 3])
def test_cell_size_invariance(engine, cpr):
    d = synth.uniform(3000, 250.0, seed=41)
    _check(engine, d, None, ani=False, bpc=True, box=250.0, bins=20, dmax=50.0, cpr=cpr)


def test_reference_example_outputs(engine):
    """results/*.example of the reference: 3iso -f data.dat -r rand0.dat -n 32768
    -b 30 -d 60, no BPC, written by the reference GPU program (FP32 distances
    with FMA contraction).  The FP64 semantics differ only for triplets with a
    side within an FP32 rounding of a bin edge or the range limit, so the
    histograms agree to ~1e-5 of the total and exactly on most bins."""
    d = catalog("data.dat")
    r = catalog("rand0.dat")
    cd, cr = engine.upload(d), engine.upload(r)
    try:
        for name, (x, y) in (("DDD", (cd, -1)), ("RRR", (cr, -1)), ("DDR", (cd, cr)),
                             ("DRR", (cr, cd))):
            job = da.make_job(da.PCF_3ISO, da.PCF_XXX if y < 0 else da.PCF_XXY, bins=30,
                              dmax=60.0, cat_x=x, cat_y=y)
            c, w = engine.histogram(job)
            ref = np.loadtxt(os.path.join(GOLDEN, "examples", f"{name}iso.example")).ravel()
            assert ref.shape == c.shape
            diff = np.abs(c.astype(np.float64) - ref)
            print(name, "total", c.sum(), "ref", ref.sum(), "differing bins",
                  int((diff > 0).sum()), "sum |diff|", diff.sum(), flush=True)
            # Edge flips move triplets between adjacent bins (the ~1e10 triplets
            # make most populated bins see a few); the totals match to the
            # triplets within a rounding of dmax, and empty bins stay empty.
            assert diff.sum() <= 2e-5 * ref.sum()
            assert abs(float(c.sum()) - ref.sum()) <= 1e-6 * ref.sum()
            assert np.array_equal(c == 0, ref == 0)
            # corner bins identified from the file (exact)
            h = c.reshape(30, 30, 30)
            if name == "DDD":
                assert (h[0, 0, 0], h[0, 0, 1], h[0, 1, 1], h[1, 1, 1]) == (130896, 80130, 268682,
                                                                          450300)
    finally:
        engine.release(cd)
        engine.release(cr)


def test_cfg2_shape_symmetry(engine):
    """Config 2 shape at full size (N=32,768 clustered, box 250, 30 bins, rmax 80,
    periodic, weighted): structural checks at a size the oracle cannot run."""
    d = synth.neyman_scott(1024, 32, 5.0, 250.0, seed=1, weights="uniform")
    cd = engine.upload(d)
    try:
        c, w, = engine.histogram(da.make_job(da.PCF_3ISO, da.PCF_XXX, bins=30, dmax=80.0,
                                             cat_x=cd, bpc=True, box=250.0))
        st = engine.stats()
        print("cfg2 DDD", c.sum(), st, flush=True)
        h = c.reshape(30, 30, 30)
        hw = w.reshape(30, 30, 30)
        for p in itertools.permutations(range(3)):
            assert np.array_equal(h, np.transpose(h, p))
            assert np.allclose(hw, np.transpose(hw, p), rtol=1e-12, atol=0)
        # weights are products of values in [0.5, 1.5]
        nz = c > 0
        ratio = w[nz] / c[nz]
        assert np.all(ratio >= 0.125 - 1e-12) and np.all(ratio <= 3.375 + 1e-12)
    finally:
        engine.release(cd)


@pytest.mark.parametrize("cross", [False, True])
def test_3ani_many_bins_global_histogram(engine, cross):
    """bins = 24 weighted: one [b31][b23][p] slice of the LDS kernel does not fit,
    so the 5-D histogram goes through the global-atomic kernel."""
    d = synth.neyman_scott(10, 30, 5.0, 250.0, seed=61, weights="uniform")
    r = synth.uniform(300, 250.0, seed=62, weights="uniform")
    _check(engine, d, r if cross else None, ani=True, bpc=True, box=250.0, bins=24, dmax=60.0)


@pytest.mark.parametrize("bins,weighted", [(4, True), (15, True), (21, True), (22, True),
                                           (15, False), (30, False)])
@pytest.mark.parametrize("cross", [False, True])
def test_3ani_kernel_domains(engine, bins, weighted, cross):
    """3ani mixed precision across the bin counts where the LDS kernel's slice
    (one t value of a b12 row, bins^3 rows of 4 + 8 bytes) goes from many t
    per slice down to one (22 bins weighted) and, unweighted, up to 30 bins;
    against the oracle, periodic, auto and cross."""
    wk = "uniform" if weighted else None
    d = synth.neyman_scott(12, 24, 4.0, 120.0, seed=70 + bins, weights=wk)
    r = synth.uniform(260, 120.0, seed=90 + bins, weights=wk)
    _check(engine, d, r if cross else None, ani=True, bpc=True, box=120.0, bins=bins,
           dmax=30.0)


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("cross", [False, True])
def test_wide_periodic_triplets(engine, ani, cross):
    """Periodic boxes with 2*dmax >= box (dmax < box): a point can count
    under its regular and its projected image on an axis, and the reference's
    2-3 side |x3 - x2| - L*|f2 - f3| is no longer a distance in point 1's
    frame.  Requested in mixed precision (the library runs such jobs on its
    plain FP64 kernels) and in FP64, against the oracle."""
    d = synth.uniform(90, 20.0, seed=101, weights="uniform")
    r = synth.uniform(70, 20.0, seed=102, weights="uniform")
    bins = 5 if ani else 8
    for dmax in (10.5, 17.0):
        for prec in (da.PCF_PREC_MIXED, da.PCF_PREC_FP64):
            _check(engine, d, r if cross else None, ani=ani, bpc=True, box=20.0, bins=bins,
                   dmax=dmax, prec=prec)


def test_wide_periodic_triplet_by_hand_gpu(engine):
    """The three points of the oracle's hand-worked case (A = 5, B = 2,
    C = 8 on a line, box 10, dmax 8): 24 ordered entries, among them the
    (7, 6, 7) sides whose 2-3 side is |8 - 2| although both projected images
    lie 14 apart."""
    pts = np.array([[5, 1, 1, 1], [2, 1, 1, 1], [8, 1, 1, 1]], np.float32)
    c, _, _ = _run(engine, pts, None, ani=False, bpc=True, box=10.0, bins=8, dmax=8.0)
    co, _ = O.triplets(pts, bpc=True, box=10.0, bins=8, dmax=8.0)
    assert np.array_equal(c, co) and c.sum() == 24
    assert c.reshape(8, 8, 8)[7, 6, 7] == 2


def test_periodic_dmax_at_box_is_unsupported(engine):
    """dmax >= box would reach far images, which no reference term is."""
    d = synth.uniform(50, 20.0, seed=103)
    with pytest.raises(da.PcfError) as ei:
        _run(engine, d, None, ani=False, bpc=True, box=20.0, bins=5, dmax=20.0)
    assert ei.value.rc == da.PCF_ERR_UNSUPPORTED
