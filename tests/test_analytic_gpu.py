"""GPU parity of the analytic periodic 3iso path (pcf_analytic_3iso, PCF 3iso
-bpc -a) against the oracle, through the C ABI.

Unit-weight catalogs give integer fine pair histograms on both sides, so
ff_av and ff_av_ref (index-order sums of identical FP64 expressions) are
bit-exact; the sub-cube sums of the partial bin cubes are reduced in another
order on the GPU and agree to 1e-12 relative.  At the reference's own
refinement (ptt 100, bn_ref 200) the result is checked against the compiled
reference program's ff_av / ff_av_ref and through refinement monotonicity."""
import itertools
import os

import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O
from conftest import GOLDEN, catalog

pytestmark = pytest.mark.gpu

L = 250.0


def _gpu(engine, d, **kw):
    c = engine.upload(d)
    try:
        return engine.analytic_3iso(c, **kw)
    finally:
        engine.release(c)


def _close(a, b, rel):
    scale = np.maximum(np.abs(b), 1e-300)
    bad = np.nonzero(np.abs(a - b) > rel * scale)[0]
    assert len(bad) == 0, f"{len(bad)} values differ, e.g. {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


def test_parity_unit_weights(engine):
    d = catalog("data.dat")[:5000]
    print("case data.dat[:5000] bins 30 dmax 60 ptt 100 bn_ref 8", flush=True)
    g = _gpu(engine, d, bins=30, dmax=60.0, box=L, ptt=100, bn_ref=8)
    o = O.analytic_3iso(d, bins=30, dmax=60.0, box=L, ptt=100, bn_ref=8)
    assert np.array_equal(g[2], o[2]), "ff_av"
    assert np.array_equal(g[3], o[3]), "ff_av_ref"
    _close(g[1], o[1], 1e-12)  # RRR
    _close(g[0], o[0], 1e-12)  # DDR
    assert np.array_equal(g[1] == 0, o[1] == 0)


@pytest.mark.parametrize("seed", [3, 4])
def test_parity_weighted_clustered(engine, seed):
    d = synth.neyman_scott(60, 40, 4.0, L, seed=seed, weights="uniform")
    print(f"case neyman_scott seed {seed} n={len(d)} bins 12 dmax 80 ptt 20 bn_ref 6",
          flush=True)
    g = _gpu(engine, d, bins=12, dmax=80.0, box=L, ptt=20, bn_ref=6)
    o = O.analytic_3iso(d, bins=12, dmax=80.0, box=L, ptt=20, bn_ref=6)
    # weighted fine sums differ in summation order (1e-15 relative per bin);
    # ff = sum r (DD/RR - 1) can cancel, so compare against the term scale
    for k in (2, 3):
        assert np.all(np.abs(g[k] - o[k]) <= 1e-11 * (np.abs(o[k]) + 1.0))
    _close(g[1], o[1], 1e-12)
    _close(g[0], o[0], 1e-10)


def test_reference_refinement(engine):
    """ptt 100 / bn_ref 200, the reference's constants, on the compiled CPU
    program's configuration (5000 points of data.dat, 30 bins, dmax 60)."""
    d = catalog("data.dat")[:5000]
    print("case data.dat[:5000] defaults", flush=True)
    ddr, rrr, ff, ffr = _gpu(engine, d, bins=30, dmax=60.0, box=L)
    st = engine.stats()
    print("analytic 3iso kernel ms", st["kernel_ms"], "total ms", st["total_ms"], flush=True)
    ref_dir = os.path.join(GOLDEN, "ref_out", "3iso_bpca")
    rff = np.loadtxt(os.path.join(ref_dir, "ff_av.dat"))
    rffr = np.loadtxt(os.path.join(ref_dir, "ff_av_ref.dat"))
    rel = np.abs(ff - rff) / np.abs(rff)
    assert np.sum(rel > 1e-6) <= 3 and rel.max() < 1e-4
    rel2 = np.abs(ffr - rffr) / np.maximum(np.abs(rffr), 1e-300)
    assert np.median(rel2) < 1e-6
    assert np.sum(np.abs(ffr - rffr)) <= 1e-3 * np.sum(np.abs(rffr))
    # ff_av / ff_av_ref at the default refinement equal the oracle's exactly
    _, _, off, offr = O.analytic_3iso(d, bins=30, dmax=60.0, box=L, cubes=False)
    assert np.array_equal(ff, off) and np.array_equal(ffr, offr)
    # refinement: full cubes do not depend on bn_ref, partial cubes only grow
    d8, r8, _, _ = _gpu(engine, d, bins=30, dmax=60.0, box=L, ptt=100, bn_ref=8)
    full = r8 == rrr
    assert full.sum() > 10000
    assert np.all(rrr >= r8 * (1 - 1e-12))
    # symmetric under permutations of the three sides
    h = rrr.reshape(30, 30, 30)
    f = ddr.reshape(30, 30, 30)
    for p in itertools.permutations(range(3)):
        assert np.allclose(h, np.transpose(h, p), rtol=1e-12, atol=0)
        assert np.allclose(f, np.transpose(f, p), rtol=1e-12, atol=0)


def test_argument_errors(engine):
    d = synth.uniform(100, L, seed=1)
    c = engine.upload(d)
    try:
        with pytest.raises(da.PcfError):
            engine.analytic_3iso(c, bins=10, dmax=260.0, box=L)  # dmax >= box
        with pytest.raises(da.PcfError):
            engine.analytic_3iso(c, bins=10, dmax=60.0, box=100.0)  # points outside box
    finally:
        engine.release(c)


def test_parity_wide_periodic(engine):
    """2*dmax >= box: the fine pair histograms count a pair under both of its
    images where both are in range, as the pair kernels do."""
    d = synth.uniform(900, 100.0, seed=9)
    print("case uniform 900 box 100 bins 6 dmax 60 ptt 10 bn_ref 4", flush=True)
    g = _gpu(engine, d, bins=6, dmax=60.0, box=100.0, ptt=10, bn_ref=4)
    o = O.analytic_3iso(d, bins=6, dmax=60.0, box=100.0, ptt=10, bn_ref=4)
    assert np.array_equal(g[2], o[2]), "ff_av"
    assert np.array_equal(g[3], o[3]), "ff_av_ref"
    _close(g[1], o[1], 1e-12)
    _close(g[0], o[0], 1e-12)
