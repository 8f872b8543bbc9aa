"""GPU parity of the Gram-form pair path (squared distances on the matrix cores
as |xi|^2 + |xj|^2 - 2 xi.xj in tile-local coordinates; 2ani as perpendicular
and line-of-sight products) against the FP64 oracle and against the direct FP32
path, single and batched, through the C ABI.  Counts must be
bit-exact.  Catalogs are chosen so that some tiles exceed the Gram path's error
bound and run on the direct path inside the same launch (direct_items > 0),
and so that many pairs sit exactly on bin edges (dyadic lattices)."""
import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _hist(engine, a, b, *, bpc, box, bins, dmax, kernel, cpr=0, prec=da.PCF_PREC_MIXED):
    ca = engine.upload(a)
    cb = engine.upload(b) if b is not None else -1
    try:
        job = da.make_job(da.PCF_2ISO, da.PCF_XX if b is None else da.PCF_XY, bins=bins,
                          dmax=dmax, cat_x=ca, cat_y=cb, bpc=bpc, box=box, precision=prec,
                          cells_per_rmax=cpr, pair_kernel=kernel)
        c, _ = engine.histogram(job)
        st = engine.stats()
    finally:
        engine.release(ca)
        if cb >= 0:
            engine.release(cb)
    return c, st


def _both(engine, a, b, **kw):
    print(f"case n={len(a)} cross={b is not None} {kw}", flush=True)
    cg, sg = _hist(engine, a, b, kernel=da.PCF_PAIR_GRAM, **kw)
    cd, _ = _hist(engine, a, b, kernel=da.PCF_PAIR_DIRECT, **kw)
    assert np.array_equal(cg, cd), f"gram {cg} direct {cd}"
    return cg, sg


@pytest.mark.parametrize("bpc", [False, True])
@pytest.mark.parametrize("cross", [False, True])
def test_gram_uniform_vs_oracle(engine, bpc, cross):
    d = synth.uniform(30000, 300.0, seed=21)
    r = synth.uniform(25000, 300.0, seed=22)
    b = r if cross else None
    c, st = _both(engine, d, b, bpc=bpc, box=300.0, bins=30, dmax=80.0)
    co, _ = O.pairs(d, b, bpc=bpc, box=300.0, bins=30, dmax=80.0)
    assert np.array_equal(c, co)
    assert st["evaluations"] > 0


@pytest.mark.parametrize("cpr", [1, 2, 3, 4])
def test_gram_clustered_with_sparse_tiles(engine, cpr):
    # Dense clumps plus a thin uniform background: background tiles span many
    # cells and exceed the Gram bound, so both paths run in one launch.
    clumps = synth.neyman_scott(256, 64, 3.0, 400.0, seed=31)
    bg = synth.uniform(3000, 400.0, seed=32)
    d = np.ascontiguousarray(np.concatenate([clumps, bg]))
    c, st = _both(engine, d, None, bpc=False, box=400.0, bins=25, dmax=60.0, cpr=cpr)
    co, _ = O.pairs(d, None, bins=25, dmax=60.0)
    assert np.array_equal(c, co)
    print("direct items", st["direct_items"], "of", st["work_items"], flush=True)


def test_gram_fallback_happens(engine):
    # Very sparse catalog with the smallest cells: 64-point tiles are long.
    clumps = synth.neyman_scott(64, 64, 2.0, 500.0, seed=41)
    bg = synth.uniform(2000, 500.0, seed=42)
    d = np.ascontiguousarray(np.concatenate([clumps, bg]))
    c, st = _both(engine, d, None, bpc=False, box=500.0, bins=20, dmax=40.0, cpr=4)
    co, _ = O.pairs(d, None, bins=20, dmax=40.0)
    assert np.array_equal(c, co)
    assert st["direct_items"] > 0
    assert st["direct_items"] < st["work_items"]


@pytest.mark.parametrize("bpc", [False, True])
def test_gram_lattice_edges(engine, bpc):
    # Dyadic lattice: FP32, FP64 and exact distances agree; pairs sit exactly
    # on bin edges, on the range limit and at d = 0.
    a = synth.lattice(6000, 64.0, 0.5, seed=13)
    b = synth.lattice(4000, 64.0, 0.5, seed=14)
    for bb in (None, b):
        c, st = _both(engine, a, bb, bpc=bpc, box=64.0, bins=10, dmax=10.0)
        co, _ = O.pairs(a, bb, bpc=bpc, box=64.0, bins=10, dmax=10.0)
        assert np.array_equal(c, co)
        assert st["rechecks"] > 0


def test_gram_large_vs_fp64_plain(engine):
    """A north-star-like shape at reduced size (N = 262,144 in box 630, rmax 80,
    30 bins; the same density as N = 1,048,576 in box 1000): Gram, direct and
    plain FP64 agree on every count."""
    d = synth.uniform(1 << 18, 630.0, seed=51)
    r = synth.uniform(1 << 18, 630.0, seed=52)
    for b in (None, r):
        c, st = _both(engine, d, b, bpc=False, box=630.0, bins=30, dmax=80.0)
        c64, _ = _hist(engine, d, b, bpc=False, box=630.0, bins=30, dmax=80.0,
                       kernel=da.PCF_PAIR_AUTO, prec=da.PCF_PREC_FP64)
        assert np.array_equal(c, c64)
        assert st["direct_items"] < st["work_items"] // 10


def _ani_both(engine, a, b, **kw):
    print(f"case 2ani n={len(a)} cross={b is not None} {kw}", flush=True)
    out = []
    for kern in (da.PCF_PAIR_GRAM, da.PCF_PAIR_DIRECT):
        ca = engine.upload(a)
        cb = engine.upload(b) if b is not None else -1
        try:
            job = da.make_job(da.PCF_2ANI, da.PCF_XX if b is None else da.PCF_XY, cat_x=ca,
                              cat_y=cb, pair_kernel=kern, **kw)
            c, _ = engine.histogram(job)
            out.append((c, engine.stats()))
        finally:
            engine.release(ca)
            if cb >= 0:
                engine.release(cb)
    assert np.array_equal(out[0][0], out[1][0])
    return out[0]


@pytest.mark.parametrize("bpc", [False, True])
@pytest.mark.parametrize("cross", [False, True])
def test_gram_2ani_vs_oracle(engine, bpc, cross):
    d = synth.neyman_scott(300, 40, 5.0, 300.0, seed=71)
    r = synth.uniform(12000, 300.0, seed=72)
    b = r if cross else None
    c, st = _ani_both(engine, d, b, bins=20, dmax=60.0, bpc=bpc, box=300.0)
    co, _ = O.pairs(d, b, ani=True, bpc=bpc, box=300.0, bins=20, dmax=60.0)
    assert np.array_equal(c, co)


@pytest.mark.parametrize("bpc", [False, True])
def test_gram_2ani_lattice_edges(engine, bpc):
    a = synth.lattice(5000, 64.0, 0.5, seed=15)
    b = synth.lattice(3000, 64.0, 0.5, seed=16)
    for bb in (None, b):
        c, st = _ani_both(engine, a, bb, bins=10, dmax=10.0, bpc=bpc, box=64.0)
        co, _ = O.pairs(a, bb, ani=True, bpc=bpc, box=64.0, bins=10, dmax=10.0)
        assert np.array_equal(c, co)
        assert st["rechecks"] > 0


def test_gram_2ani_batch_vs_direct_batch(engine):
    """The batched persistent launch in both forms: every histogram equal, and
    equal to the oracle."""
    d = synth.neyman_scott(40, 32, 8.0, 1000.0, seed=81)
    rs = [synth.uniform(n, 1000.0, seed=90 + n) for n in (900, 2500, 6000)]
    cd = engine.upload(d)
    crs = [engine.upload(r) for r in rs]
    try:
        res = {}
        for kern in (da.PCF_PAIR_GRAM, da.PCF_PAIR_DIRECT):
            jobs = [da.make_job(da.PCF_2ANI, da.PCF_XX, bins=30, dmax=120.0, cat_x=cd,
                                pair_kernel=kern)]
            for cr in crs:
                jobs.append(da.make_job(da.PCF_2ANI, da.PCF_XX, bins=30, dmax=120.0, cat_x=cr,
                                        pair_kernel=kern))
                jobs.append(da.make_job(da.PCF_2ANI, da.PCF_XY, bins=30, dmax=120.0, cat_x=cd,
                                        cat_y=cr, pair_kernel=kern))
            res[kern] = engine.histogram_batch(jobs)
        for (c1, _), (c2, _) in zip(res[da.PCF_PAIR_GRAM], res[da.PCF_PAIR_DIRECT]):
            assert np.array_equal(c1, c2)
        co, _ = O.pairs(d, rs[1], ani=True, bins=30, dmax=120.0)
        assert np.array_equal(res[da.PCF_PAIR_GRAM][4][0], co)
    finally:
        engine.release(cd)
        for c in crs:
            engine.release(c)


@pytest.mark.parametrize("cross", [False, True])
def test_long_runs_and_block_windows(engine, cross):
    """Work items carry step windows per 16-point block (blocks outside them
    are not evaluated) unless a run has more than 255 steps of 16.  A dense
    catalog with one-rmax cells gives runs of thousands of points (no
    windows) beside the windowed items of a sparse second catalog; both forms
    equal the oracle, and with windows fewer pairs are evaluated than the
    items hold."""
    dense = synth.uniform(30000, 100.0, seed=41)  # ~7,500-point runs at 1 cell per rmax
    c, st = _both(engine, dense, None, bpc=False, box=100.0, bins=20, dmax=45.0, cpr=1)
    co, _ = O.pairs(dense, None, bins=20, dmax=45.0)
    assert np.array_equal(c, co)
    sparse = synth.uniform(20000, 400.0, seed=42)
    b = synth.uniform(15000, 400.0, seed=43) if cross else None
    c, st = _both(engine, sparse, b, bpc=True, box=400.0, bins=25, dmax=60.0, cpr=2)
    co, _ = O.pairs(sparse, b, bpc=True, box=400.0, bins=25, dmax=60.0)
    assert np.array_equal(c, co)
    assert st["evaluations"] > 0


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("bpc", [False, True])
def test_slab_windows_at_the_reach(engine, ani, bpc):
    """Resident grids order each cell's points by z sub-slab, and a block's
    step window keeps only the j slabs within the column's z reach of its own.
    Pairs separated almost entirely along z, just inside and just outside
    the range, at several lateral offsets, in a catalog dense enough for
    windows that cut through cells: both forms equal the oracle, so no window
    drops an in-range pair."""
    rng = np.random.default_rng(7)
    box, dmax, bins = 200.0, 30.0, 15
    base = rng.uniform(5.0, box - 5.0, (4000, 3))
    parts = [base]
    for eps in (1e-3, 3e-5, -3e-5):
        for gx in (0.0, 7.0, 19.0):
            dz = np.sqrt(dmax * dmax - gx * gx) - eps
            q = base[:250].copy()
            q[:, 0] += gx
            q[:, 2] += dz
            parts.append(q)
    xyz = np.concatenate(parts)
    if bpc:
        xyz = np.mod(xyz, box)
        xyz[xyz >= box] = 0.0
    a = np.zeros((len(xyz), 4), np.float32)
    a[:, :3] = xyz
    a[:, 3] = 1.0
    b = a[::-1].copy()[:3000]
    for y in (None, b):
        print(f"case ani={ani} bpc={bpc} cross={y is not None}", flush=True)
        for kernel in (da.PCF_PAIR_GRAM, da.PCF_PAIR_DIRECT):
            ca = engine.upload(a)
            cb = engine.upload(y) if y is not None else -1
            try:
                job = da.make_job(da.PCF_2ANI if ani else da.PCF_2ISO,
                                  da.PCF_XX if y is None else da.PCF_XY, bins=bins, dmax=dmax,
                                  cat_x=ca, cat_y=cb, bpc=bpc, box=box, cells_per_rmax=2,
                                  pair_kernel=kernel)
                c, _ = engine.histogram(job)
            finally:
                engine.release(ca)
                if cb >= 0:
                    engine.release(cb)
            co, _ = O.pairs(a, y, ani=ani, bpc=bpc, box=box, bins=bins, dmax=dmax)
            assert np.array_equal(c, co), f"kernel {kernel}: {c.sum()} vs {co.sum()}"
