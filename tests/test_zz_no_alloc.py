"""pcf_histogram_noalloc (no device allocation; buffers in a static device
array) against pcf_histogram, through the C ABI: counts bit-exact, weights
within 1e-12 relative, on two calls in a row with the second smaller (the
array keeps the first call's contents), for 2iso and 2ani, auto and cross,
periodic or not, mixed precision and plain FP64, plus the size limit."""
import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth

pytestmark = pytest.mark.gpu

WREL = 1e-12


def _same(a, b):
    (c1, w1), (c2, w2) = a, b
    assert np.array_equal(c1, c2)
    assert np.all(np.abs(w1 - w2) <= WREL * np.maximum(np.abs(w2), 1e-300))


@pytest.mark.parametrize("type_", [da.PCF_2ISO, da.PCF_2ANI])
@pytest.mark.parametrize("bpc", [False, True])
@pytest.mark.parametrize("prec", [da.PCF_PREC_MIXED, da.PCF_PREC_FP64])
def test_noalloc_equals_histogram(engine, type_, bpc, prec):
    big_d = synth.neyman_scott(400, 32, 5.0, 250.0, seed=61, weights="uniform")
    big_r = synth.uniform(16000, 250.0, seed=62)
    small_d = synth.uniform(3000, 250.0, seed=63)
    small_r = synth.uniform(2000, 250.0, seed=64, weights="uniform")
    for d, r in ((big_d, big_r), (small_d, small_r)):   # second call smaller
        print(f"case type={type_} bpc={bpc} prec={prec} nd={len(d)} nr={len(r)}", flush=True)
        cd, cr = engine.upload(d), engine.upload(r)
        try:
            for hist, y in ((da.PCF_XX, -1), (da.PCF_XY, cr)):
                job = da.make_job(type_, hist, bins=30, dmax=80.0, cat_x=cd, cat_y=y, bpc=bpc,
                                  box=250.0, precision=prec)
                _same(engine.histogram_noalloc(job), engine.histogram(job))
        finally:
            engine.release(cd)
            engine.release(cr)


def test_noalloc_ns_sized_cross(engine):
    # About the largest cross job the array holds: 1M x 1M uniform points,
    # rmax 80 in a box of 1000 (two such calls in a row, then a smaller one).
    d = synth.uniform(1 << 20, 1000.0, seed=65)
    r = synth.uniform(1 << 20, 1000.0, seed=66)
    s = synth.uniform(50000, 1000.0, seed=67)
    cd, cr, cs = engine.upload(d), engine.upload(r), engine.upload(s)
    try:
        for x, y in ((cd, cr), (cd, cr), (cs, cr)):
            print(f"case NS-sized cross x={x} y={y}", flush=True)
            job = da.make_job(da.PCF_2ISO, da.PCF_XY, bins=30, dmax=80.0, cat_x=x, cat_y=y)
            _same(engine.histogram_noalloc(job), engine.histogram(job))
    finally:
        for c in (cd, cr, cs):
            engine.release(c)


def test_noalloc_rejects_jobs_beyond_the_array(engine):
    a = synth.uniform(3 << 20, 1000.0, seed=68)
    ca = engine.upload(a)
    try:
        job = da.make_job(da.PCF_2ISO, da.PCF_XX, bins=30, dmax=80.0, cat_x=ca)
        with pytest.raises(da.PcfError) as e:
            engine.histogram_noalloc(job)
        assert e.value.rc == da.PCF_ERR_NOMEM
        assert "bytes of scratch" in str(e.value)
        # The ordinary entry point still serves it, and a small job after the
        # refusal is still right.
        engine.histogram(job)
        b = synth.uniform(5000, 1000.0, seed=69)
        cb = engine.upload(b)
        try:
            j2 = da.make_job(da.PCF_2ANI, da.PCF_XX, bins=20, dmax=60.0, cat_x=cb)
            _same(engine.histogram_noalloc(j2), engine.histogram(j2))
        finally:
            engine.release(cb)
    finally:
        engine.release(ca)


def test_noalloc_refuses_triplets(engine):
    a = synth.uniform(500, 100.0, seed=70)
    ca = engine.upload(a)
    try:
        with pytest.raises(da.PcfError) as e:
            engine.histogram_noalloc(da.make_job(da.PCF_3ISO, da.PCF_XXX, bins=10, dmax=20.0,
                                                 cat_x=ca))
        assert e.value.rc == da.PCF_ERR_UNSUPPORTED
    finally:
        engine.release(ca)
