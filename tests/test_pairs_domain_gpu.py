"""GPU parity of the pair histograms beyond the mixed kernel's envelope, against
the FP64 oracle, through the C ABI and the command line.

* Periodic boxes with 2*dmax >= L (dmax < L): a pair counts under its regular
  separation and under each close image in range, as the reference's per-axis
  node flags allow (cuda/kernels/2isoBPC_k.cuh:43-47).  The reference's own CPU
  program runs 2ani-BPC at dmax 150 in a box of 250
  (cpp/2PCFanisotropic/BPC/main.cpp:24-26).
* Bin counts past the mixed kernel's u8 table (2iso > 253, 2ani > 120) or its
  LDS histogram: the plain FP64 kernel, with an LDS or a global-memory
  histogram (the reference deposits with global atomics for any -b,
  cuda/kernels/2iso_k.cuh:51-55).
"""
import os

import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O
from conftest import GOLDEN, catalog
from test_pairs_gpu import _check, _close

pytestmark = pytest.mark.gpu

FIX = os.path.join(GOLDEN, "oracle_out")


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("cross", [False, True])
@pytest.mark.parametrize("prec", [da.PCF_PREC_MIXED, da.PCF_PREC_FP64])
def test_periodic_2dmax_ge_box(engine, ani, cross, prec):
    d = catalog("data.dat")[:2500]
    r = catalog("rand0.dat")[:2000]
    for dmax in (150.0, 240.0):
        _check(engine, d, r if cross else None, ani=ani, bpc=True, box=250.0, bins=20,
               dmax=dmax, prec=prec)


def test_periodic_2dmax_ge_box_lattice(engine):
    # Exact edges: dyadic lattice in a box of 32 with dmax 20.
    a = synth.lattice(1500, 32.0, 0.5, seed=31)
    b = synth.lattice(1200, 32.0, 0.5, seed=32)
    for ani in (False, True):
        _check(engine, a, None, ani=ani, bpc=True, box=32.0, bins=10, dmax=20.0)
        _check(engine, a, b, ani=ani, bpc=True, box=32.0, bins=10, dmax=20.0)


def test_periodic_weighted_2dmax_ge_box(engine):
    d = synth.neyman_scott(40, 32, 6.0, 100.0, seed=33, weights="uniform")
    r = synth.uniform(1500, 100.0, seed=34, weights="uniform")
    for ani in (False, True):
        _check(engine, d, None, ani=ani, bpc=True, box=100.0, bins=12, dmax=70.0)
        _check(engine, d, r, ani=ani, bpc=True, box=100.0, bins=12, dmax=70.0)


def test_periodic_dmax_ge_box_refused(engine):
    a = synth.uniform(100, 50.0, seed=1)
    ca = engine.upload(a)
    try:
        with pytest.raises(da.PcfError) as e:
            engine.histogram(da.make_job(da.PCF_2ISO, da.PCF_XX, bins=10, dmax=50.0, cat_x=ca,
                                         bpc=True, box=50.0))
        assert e.value.rc == da.PCF_ERR_UNSUPPORTED
    finally:
        engine.release(ca)


def _fixture(name):
    return np.loadtxt(os.path.join(FIX, f"2ani_bpc_d150_s250_b50_{name}.txt"), dtype=np.uint64)


def test_2ani_bpc_reference_program_config(engine):
    """The reference CPU program's own 2ani-BPC configuration on the full
    sample catalogs (32,768 points each), against the oracle's histograms
    (tests/golden/make_oracle_fixtures.py)."""
    d, r = catalog("data.dat"), catalog("rand0.dat")
    cd, cr = engine.upload(d), engine.upload(r)
    try:
        for name, x, y in (("DD", cd, -1), ("RR", cr, -1), ("DR", cd, cr)):
            print("case 2ani bpc d150 s250 b50", name, flush=True)
            job = da.make_job(da.PCF_2ANI, da.PCF_XX if y < 0 else da.PCF_XY, bins=50,
                              dmax=150.0, cat_x=x, cat_y=y, bpc=True, box=250.0)
            c, w = engine.histogram(job)
            ref = _fixture(name)
            assert np.array_equal(c, ref), name
            assert np.array_equal(w, c.astype(np.float64))
    finally:
        engine.release(cd)
        engine.release(cr)


@pytest.mark.parametrize("bins", [300, 1100, 20000])
def test_many_bins_2iso(engine, bins):
    d = synth.neyman_scott(30, 32, 5.0, 200.0, seed=41, weights="uniform")
    r = synth.uniform(1500, 200.0, seed=42)
    for bpc in (False, True):
        _check(engine, d, None, ani=False, bpc=bpc, box=200.0, bins=bins, dmax=60.0)
        _check(engine, d, r, ani=False, bpc=bpc, box=200.0, bins=bins, dmax=60.0)
        _check(engine, r, None, ani=False, bpc=bpc, box=200.0, bins=bins, dmax=60.0)


@pytest.mark.parametrize("bins", [121, 200, 400])
def test_many_bins_2ani(engine, bins):
    d = synth.neyman_scott(30, 32, 5.0, 200.0, seed=43, weights="uniform")
    r = synth.uniform(1200, 200.0, seed=44)
    for bpc in (False, True):
        _check(engine, d, None, ani=True, bpc=bpc, box=200.0, bins=bins, dmax=60.0)
        _check(engine, r, d, ani=True, bpc=bpc, box=200.0, bins=bins, dmax=60.0)
        _check(engine, r, None, ani=True, bpc=bpc, box=200.0, bins=bins, dmax=60.0)


def test_weighted_2ani_120_bins(engine):
    # Weighted at 120 bins: the mixed kernel's LDS histogram does not fit.
    d = synth.neyman_scott(30, 32, 5.0, 200.0, seed=45, weights="uniform")
    _check(engine, d, None, ani=True, bpc=False, box=200.0, bins=120, dmax=60.0)


def test_batch_many_bins_equals_single(engine):
    d = synth.uniform(2000, 200.0, seed=46)
    rs = [synth.uniform(n, 200.0, seed=47 + n) for n in (900, 1300)]
    cd = engine.upload(d)
    crs = [engine.upload(r) for r in rs]
    try:
        for type_, bins in ((da.PCF_2ISO, 400), (da.PCF_2ANI, 150)):
            jobs = [da.make_job(type_, da.PCF_XX, bins=bins, dmax=50.0, cat_x=cd)]
            jobs += [da.make_job(type_, da.PCF_XY, bins=bins, dmax=50.0, cat_x=cd, cat_y=cr)
                     for cr in crs]
            got = engine.histogram_batch(jobs)
            for j, (c, w) in zip(jobs, got):
                c1, w1 = engine.histogram(j)
                assert np.array_equal(c, c1)
                assert _close(w, w1)
    finally:
        engine.release(cd)
        for cr in crs:
            engine.release(cr)
