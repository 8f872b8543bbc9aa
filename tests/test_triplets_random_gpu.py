"""Seeded random sweep of triplet histograms against the FP64 oracle through
the C ABI: 3iso and 3ani, XXX and XXY, periodic or open boxes (periodic
reaches of a third of the box and more take the exact kernels), uniform or
clustered catalogs, with and without weights, several bin counts and cell
sizes, and the pair list kept from one job to the next (XXX then XXY over the
same first catalog).  Counts must be bit-exact and weighted sums within 1e-12
relative."""
import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu

WREL = 1e-12


@pytest.mark.parametrize("seed", range(30))
def test_random_triplets(engine, seed):
    rng = np.random.default_rng(5000 + seed)
    box = float(rng.uniform(30.0, 200.0))
    bpc = bool(rng.random() < 0.5)
    ani = bool(rng.random() < 0.4)
    dmax = float(box * rng.uniform(0.05, 0.45 if bpc else 0.35))
    bins = int(rng.choice([2, 6, 15]))
    cpr = int(rng.choice([0, 1, 2]))
    n = int(rng.integers(20, 500))
    weights = "uniform" if rng.random() < 0.5 else None

    def cat(s):
        if rng.random() < 0.5:
            return synth.uniform(n, box, seed=s, weights=weights)
        kids = int(rng.integers(4, 17))
        return synth.neyman_scott(max(1, n // kids), kids, float(box * rng.uniform(0.01, 0.05)),
                                  box, seed=s, weights=weights)

    a, b = cat(seed * 11 + 1), cat(seed * 11 + 2)
    typ = da.PCF_3ANI if ani else da.PCF_3ISO
    print(f"case seed={seed} ani={ani} bpc={bpc} box={box:.3f} dmax={dmax:.3f} bins={bins} "
          f"cpr={cpr} n={n} weights={weights}", flush=True)
    ca, cb = engine.upload(a), engine.upload(b)
    try:
        for hist, y, cy in ((da.PCF_XXX, None, -1), (da.PCF_XXY, b, cb)):
            job = da.make_job(typ, hist, bins=bins, dmax=dmax, cat_x=ca, cat_y=cy, bpc=bpc,
                              box=box, cells_per_rmax=cpr)
            c, w = engine.histogram(job)
            co, wo = O.triplets(a, y, ani=ani, bpc=bpc, box=box, bins=bins, dmax=dmax)
            bad = np.nonzero(c != co)[0]
            assert len(bad) == 0, (f"{'XXY' if y is not None else 'XXX'}: {len(bad)} bins "
                                   f"differ; totals {c.sum()} vs {co.sum()}")
            assert np.all(np.abs(w - wo) <= WREL * np.maximum(np.abs(wo), 1e-300))
    finally:
        engine.release(ca)
        engine.release(cb)
