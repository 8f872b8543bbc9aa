"""Seeded random sweep of pair histograms against the FP64 oracle through the
C ABI: 2iso and 2ani, auto and cross, periodic or open boxes, uniform or
clustered catalogs, with and without weights, every kernel form and cell size,
resident or rebuilt grids, and batches of jobs in one launch.  Counts must be
bit-exact and weighted sums within 1e-12 relative.

The draws cover what the fixed cases leave between them: sparse catalogs whose
16-point blocks span several z cells, dense ones whose windows cut through a
cell's sub-slabs, reaches from a few percent of the box to just under it
(periodic reaches of half the box and more take the exact kernels), and single
or many bins.
"""
import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu

WREL = 1e-12
KERNELS = (da.PCF_PAIR_AUTO, da.PCF_PAIR_DIRECT, da.PCF_PAIR_GRAM)


def _draw(seed):
    rng = np.random.default_rng(seed)
    box = float(rng.uniform(20.0, 300.0))
    bpc = bool(rng.random() < 0.5)
    case = {
        "ani": bool(rng.random() < 0.4),
        "bpc": bpc,
        "box": box,
        "dmax": float(box * rng.uniform(0.04, 0.7 if bpc else 0.6)),
        "bins": int(rng.choice([1, 5, 17, 30])),
        "cross": bool(rng.random() < 0.5),
        "cpr": int(rng.choice([0, 1, 2, 3])),
        "kernel": int(rng.choice(KERNELS)),
        "rebuild": bool(rng.random() < 0.3),
    }
    n = int(rng.integers(30, 2500))
    weights = "uniform" if rng.random() < 0.3 else None

    def cat(s):
        if rng.random() < 0.5:
            return synth.uniform(n, box, seed=s, weights=weights)
        kids = int(rng.integers(4, 33))
        return synth.neyman_scott(max(1, n // kids), kids, float(box * rng.uniform(0.005, 0.05)),
                                  box, seed=s, weights=weights)

    a = cat(seed * 7 + 1)
    b = cat(seed * 7 + 2) if case["cross"] else None
    return case, a, b


def _check(c, w, a, b, case):
    co, wo = O.pairs(a, b, ani=case["ani"], bpc=case["bpc"], box=case["box"],
                     bins=case["bins"], dmax=case["dmax"])
    bad = np.nonzero(c != co)[0]
    assert len(bad) == 0, (f"{len(bad)} bins differ, e.g. {bad[:6]}: gpu {c[bad[:6]]} oracle "
                           f"{co[bad[:6]]}; totals {c.sum()} vs {co.sum()}")
    assert np.all(np.abs(w - wo) <= WREL * np.maximum(np.abs(wo), 1e-300)), "weighted sums differ"


@pytest.mark.parametrize("seed", range(150))
def test_random_pairs(engine, seed):
    case, a, b = _draw(seed)
    print(f"case seed={seed} n={len(a)} m={None if b is None else len(b)} {case}", flush=True)
    ca = engine.upload(a)
    cb = engine.upload(b) if b is not None else -1
    try:
        job = da.make_job(da.PCF_2ANI if case["ani"] else da.PCF_2ISO,
                          da.PCF_XX if b is None else da.PCF_XY, bins=case["bins"],
                          dmax=case["dmax"], cat_x=ca, cat_y=cb, bpc=case["bpc"], box=case["box"],
                          cells_per_rmax=case["cpr"], pair_kernel=case["kernel"],
                          rebuild_grid=case["rebuild"])
        c, w = engine.histogram(job)
        # The same job again on the resident (or once more rebuilt) grids.
        c2, w2 = engine.histogram(job)
    finally:
        engine.release(ca)
        if cb >= 0:
            engine.release(cb)
    _check(c, w, a, b, case)
    assert np.array_equal(c, c2)
    assert np.all(np.abs(w2 - w) <= WREL * np.maximum(np.abs(w), 1e-300))


@pytest.mark.parametrize("seed", range(1000, 1024))
def test_random_batches(engine, seed):
    """Three histograms (XX of both catalogs, XY) in one batched launch, each
    equal to the oracle's."""
    case, a, b = _draw(seed)
    if b is None:
        b = synth.uniform(len(a), case["box"], seed=seed)
    print(f"case seed={seed} n={len(a)} m={len(b)} {case}", flush=True)
    ca, cb = engine.upload(a), engine.upload(b)
    try:
        typ = da.PCF_2ANI if case["ani"] else da.PCF_2ISO
        jobs = [da.make_job(typ, h, bins=case["bins"], dmax=case["dmax"], cat_x=x, cat_y=y,
                            bpc=case["bpc"], box=case["box"], cells_per_rmax=case["cpr"],
                            pair_kernel=case["kernel"], rebuild_grid=case["rebuild"])
                for h, x, y in ((da.PCF_XX, ca, -1), (da.PCF_XX, cb, -1), (da.PCF_XY, ca, cb))]
        out = engine.histogram_batch(jobs)
    finally:
        engine.release(ca)
        engine.release(cb)
    for (c, w), (x, y) in zip(out, ((a, None), (b, None), (a, b))):
        _check(c, w, x, y, case)
