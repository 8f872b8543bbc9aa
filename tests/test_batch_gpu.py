"""Batched pair histograms in one persistent launch (pcf_histogram_batch):
every histogram of the batch equals the single-job result (counts bit-exact,
weights to 1e-12 relative) and the oracle; a repeated batch replays the
captured graph; a changed catalog invalidates the plan."""
import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _jobs(type_, d, rs, *, bins, dmax, bpc=False, box=0.0):
    jobs = [da.make_job(type_, da.PCF_XX, bins=bins, dmax=dmax, cat_x=d, bpc=bpc, box=box)]
    for r in rs:
        jobs.append(da.make_job(type_, da.PCF_XX, bins=bins, dmax=dmax, cat_x=r, bpc=bpc, box=box))
        jobs.append(da.make_job(type_, da.PCF_XY, bins=bins, dmax=dmax, cat_x=d, cat_y=r, bpc=bpc,
                                box=box))
    return jobs


def _singles(engine, jobs):
    out = []
    for j in jobs:
        out.append(engine.histogram(j))
    return out


def _same(batch, single, rel=1e-12):
    for k, ((cb, wb), (cs, ws)) in enumerate(zip(batch, single)):
        assert np.array_equal(cb, cs), f"job {k}: counts differ"
        assert np.all(np.abs(wb - ws) <= rel * np.maximum(np.abs(ws), 1e-300)), f"job {k} weights"


@pytest.mark.parametrize("type_", [da.PCF_2ISO, da.PCF_2ANI])
def test_ragged_batch_with_empty_and_single_point_catalogs(engine, type_):
    """Randoms of 0, 1, 2 and a few thousand points in one batch (jobs with no
    work items beside large ones): every histogram equals its single run and
    the oracle."""
    d = synth.neyman_scott(24, 32, 6.0, 400.0, seed=71)
    rcats = [np.zeros((0, 4), np.float32), synth.uniform(1, 400.0, seed=72),
             np.array([[10, 10, 10, 1], [12, 11, 10, 1]], np.float32),
             synth.uniform(3000, 400.0, seed=73)]
    print("case ragged batch type", type_, flush=True)
    cd = engine.upload(d)
    crs = [engine.upload(r) for r in rcats]
    try:
        jobs = _jobs(type_, cd, crs, bins=12, dmax=60.0)
        res = engine.histogram_batch(jobs)
        _same(res, _singles(engine, jobs))
        ani = type_ == da.PCF_2ANI
        for k, r in enumerate(rcats):
            c, _ = O.pairs(r, None, ani=ani, bins=12, dmax=60.0) if len(r) else (None, None)
            if c is not None:
                assert np.array_equal(res[1 + 2 * k][0], c)
            else:
                assert res[1 + 2 * k][0].sum() == 0
            c, _ = O.pairs(d, r, ani=ani, bins=12, dmax=60.0) if len(r) else (None, None)
            if c is not None:
                assert np.array_equal(res[2 + 2 * k][0], c)
            else:
                assert res[2 + 2 * k][0].sum() == 0
    finally:
        engine.release(cd)
        for c in crs:
            engine.release(c)


@pytest.mark.parametrize("weighted", [False, True])
def test_2ani_many_randoms(engine, weighted):
    """The 64-random 2ani shape at small size: clustered D, uneven uniform
    randoms, rmax 120, 30 x 30 bins, no BPC."""
    w = "uniform" if weighted else None
    d = synth.neyman_scott(48, 32, 8.0, 1000.0, seed=51, weights=w)
    rng = np.random.default_rng(52)
    sizes = [int(round(np.exp(rng.uniform(np.log(300), np.log(3000))))) for _ in range(6)]
    rcats = [synth.uniform(n, 1000.0, seed=60 + i, weights=w) for i, n in enumerate(sizes)]
    print("case 2ani batch weighted", weighted, "sizes", sizes, flush=True)
    cd = engine.upload(d)
    crs = [engine.upload(r) for r in rcats]
    try:
        jobs = _jobs(da.PCF_2ANI, cd, crs, bins=30, dmax=120.0)
        res = engine.histogram_batch(jobs)
        st = engine.stats()
        assert st["graph_replay"] == 0 and st["chunks"] > 0
        # replay of the captured graph (before single jobs re-grid the catalogs
        # on their own geometries and evict the batch's grids)
        res2 = engine.histogram_batch(jobs)
        assert engine.stats()["graph_replay"] == 1
        for (c1, w1), (c2, w2) in zip(res, res2):
            assert np.array_equal(c1, c2)
            assert np.allclose(w1, w2, rtol=1e-12, atol=0)
        _same(res, _singles(engine, jobs))
        # the oracle on DD and one (RR, DR)
        c, _ = O.pairs(d, None, ani=True, bins=30, dmax=120.0)
        assert np.array_equal(res[0][0], c)
        c, _ = O.pairs(d, rcats[2], ani=True, bins=30, dmax=120.0)
        assert np.array_equal(res[6][0], c)
    finally:
        engine.release(cd)
        for c in crs:
            engine.release(c)


def test_2iso_periodic_zero_rules(engine):
    """XX excludes coincident points, XY keeps them: duplicated lattice points
    in one batch with both kinds of job."""
    a = synth.lattice(800, 32.0, 0.5, seed=71)
    a = np.concatenate([a, a[:100]])  # exact duplicates
    b = np.concatenate([synth.lattice(600, 32.0, 0.5, seed=72), a[:50]])
    ca, cb = engine.upload(a), engine.upload(b)
    try:
        jobs = [da.make_job(da.PCF_2ISO, da.PCF_XX, bins=8, dmax=6.0, cat_x=ca, bpc=True, box=32.0),
                da.make_job(da.PCF_2ISO, da.PCF_XY, bins=8, dmax=6.0, cat_x=ca, cat_y=cb, bpc=True,
                            box=32.0),
                da.make_job(da.PCF_2ISO, da.PCF_XX, bins=8, dmax=6.0, cat_x=cb, bpc=True, box=32.0)]
        res = engine.histogram_batch(jobs)
        for (c, _), (x, y) in zip(res, ((a, None), (a, b), (b, None))):
            co, _ = O.pairs(x, y, bpc=True, box=32.0, bins=8, dmax=6.0)
            assert np.array_equal(c, co)
        _same(res, _singles(engine, jobs))
    finally:
        engine.release(ca)
        engine.release(cb)


def test_plan_invalidated_by_new_catalog(engine):
    d = synth.uniform(1500, 500.0, seed=81)
    r1 = synth.uniform(1200, 500.0, seed=82)
    r2 = synth.uniform(900, 500.0, seed=83)
    cd, cr = engine.upload(d), engine.upload(r1)
    try:
        jobs = _jobs(da.PCF_2ISO, cd, [cr], bins=20, dmax=60.0)
        engine.histogram_batch(jobs)
        engine.release(cr)
        cr2 = engine.upload(r2)  # may reuse the id
        cr = cr2
        jobs = _jobs(da.PCF_2ISO, cd, [cr], bins=20, dmax=60.0)
        res = engine.histogram_batch(jobs)
        assert engine.stats()["graph_replay"] == 0
        c, _ = O.pairs(d, r2, bins=20, dmax=60.0)
        assert np.array_equal(res[2][0], c)
    finally:
        engine.release(cd)
        engine.release(cr)


def test_batch_argument_errors(engine):
    d = synth.uniform(200, 100.0, seed=91)
    cd = engine.upload(d)
    try:
        # jobs of one batch share the histogram type and periodicity
        jobs = [da.make_job(da.PCF_2ISO, da.PCF_XX, bins=10, dmax=20.0, cat_x=cd),
                da.make_job(da.PCF_2ANI, da.PCF_XX, bins=10, dmax=20.0, cat_x=cd)]
        with pytest.raises(da.PcfError):
            engine.histogram_batch(jobs)
        jobs = [da.make_job(da.PCF_2ISO, da.PCF_XX, bins=10, dmax=20.0, cat_x=cd),
                da.make_job(da.PCF_2ISO, da.PCF_XX, bins=10, dmax=20.0, cat_x=cd, bpc=True,
                            box=100.0)]
        with pytest.raises(da.PcfError):
            engine.histogram_batch(jobs)
        jobs = [da.make_job(da.PCF_3ISO, da.PCF_XXX, bins=10, dmax=20.0, cat_x=cd)]
        with pytest.raises(da.PcfError):
            engine.histogram_batch(jobs)
    finally:
        engine.release(cd)


@pytest.mark.parametrize("type_", [da.PCF_2ISO, da.PCF_2ANI])
@pytest.mark.parametrize("bpc", [False, True])
def test_heterogeneous_batch(engine, type_, bpc):
    """Jobs of one launch with different bins and dmax (each with its own bin
    table, range and scale on the grid of the largest dmax), weighted and
    not: every histogram equals its single run, and the oracle."""
    d = synth.neyman_scott(60, 32, 6.0, 300.0, seed=81, weights="uniform")
    r1 = synth.uniform(2500, 300.0, seed=82)
    r2 = synth.uniform(1800, 300.0, seed=83, weights="uniform")
    cd, c1, c2 = engine.upload(d), engine.upload(r1), engine.upload(r2)
    print(f"case heterogeneous batch type={type_} bpc={bpc}", flush=True)
    try:
        J = da.make_job
        kw = dict(bpc=bpc, box=300.0)
        jobs = [J(type_, da.PCF_XX, bins=20, dmax=40.0, cat_x=cd, **kw),
                J(type_, da.PCF_XY, bins=30, dmax=80.0, cat_x=cd, cat_y=c1, **kw),
                J(type_, da.PCF_XX, bins=40, dmax=110.0, cat_x=c2, **kw),
                J(type_, da.PCF_XY, bins=20, dmax=40.0, cat_x=c1, cat_y=c2, **kw),
                J(type_, da.PCF_XY, bins=30, dmax=65.5, cat_x=cd, cat_y=c2, **kw)]
        res = engine.histogram_batch(jobs)
        assert [len(c) for c, _ in res] == [da.hist_size(j.type, j.bins) for j in jobs]
        _same(res, _singles(engine, jobs))
        cats = {cd: d, c1: r1, c2: r2}
        ani = type_ == da.PCF_2ANI
        for j, (c, w) in zip(jobs, res):
            b = None if j.hist == da.PCF_XX else cats[j.cat_y]
            co, wo = O.pairs(cats[j.cat_x], b, ani=ani, bpc=bpc, box=300.0, bins=j.bins,
                             dmax=j.dmax)
            assert np.array_equal(c, co)
            assert np.all(np.abs(w - wo) <= 1e-12 * np.maximum(np.abs(wo), 1e-300))
        # the plan replays
        res2 = engine.histogram_batch(jobs)
        assert engine.stats()["graph_replay"] == 1
        _same(res2, res)
    finally:
        for c in (cd, c1, c2):
            engine.release(c)


@pytest.mark.parametrize("type_", [da.PCF_2ISO, da.PCF_2ANI])
@pytest.mark.parametrize("ncat", [5, 10])
def test_regrid_batch_joint_sort(engine, type_, ncat):
    """Batches that re-grid their catalogs: up to eight catalogs are sorted in
    one joint counting sort, more one by one.  Catalogs of 0, 1 and thousands
    of points, periodic box: every histogram equals its single run (grids
    built one catalog at a time) and the oracle, on the first run and on the
    graph replays."""
    box, dmax = 120.0, 17.0
    sizes = [2500, 0, 1, 1700, 4000, 900, 3, 2200, 1200, 60][:ncat]
    cats = [synth.neyman_scott(max(1, n // 20), 20, 2.0, box, seed=300 + k, weights="uniform")[:n]
            if k % 2 == 0 else synth.uniform(n, box, seed=300 + k, weights="uniform")
            for k, n in enumerate(sizes)]
    ids = [engine.upload(x) for x in cats]
    try:
        jobs = [da.make_job(type_, da.PCF_XX, bins=12, dmax=dmax, cat_x=ids[0], bpc=True, box=box,
                            rebuild_grid=True)]
        for k in range(1, ncat):
            jobs.append(da.make_job(type_, da.PCF_XX, bins=12, dmax=dmax, cat_x=ids[k], bpc=True,
                                    box=box, rebuild_grid=True))
            jobs.append(da.make_job(type_, da.PCF_XY, bins=12, dmax=dmax, cat_x=ids[0],
                                    cat_y=ids[k], bpc=True, box=box, rebuild_grid=True))
        single = _singles(engine, jobs)
        for rep in range(3):
            _same(engine.histogram_batch(jobs), single)
        for k in (0, 3, 4):
            c, w = single[0 if k == 0 else 2 * k - 1]
            co, wo = O.pairs(cats[k], None, ani=type_ == da.PCF_2ANI, bpc=True, box=box, bins=12,
                             dmax=dmax)
            assert np.array_equal(c, co), f"catalog {k}"
            assert np.all(np.abs(w - wo) <= 1e-12 * np.maximum(np.abs(wo), 1e-300))
    finally:
        for i in ids:
            engine.release(i)
