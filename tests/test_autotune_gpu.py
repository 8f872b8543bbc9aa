"""Autotuner of the pair kernels (pcf_autotune) and the cross-configuration
equality it relies on: every cell size, LDS histogram layout and occupancy
gives the same counts (bit-exact) and weighted sums (1e-12 relative)."""
import itertools

import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ani,bpc", [(False, False), (False, True), (True, False), (True, True)])
def test_cross_config_equality(engine, ani, bpc):
    d = synth.neyman_scott(40, 40, 6.0, 300.0, seed=101, weights="uniform")
    r = synth.uniform(2500, 300.0, seed=102, weights="uniform")
    type_ = da.PCF_2ANI if ani else da.PCF_2ISO
    bins = 12 if ani else 25
    cd, cr = engine.upload(d), engine.upload(r)
    print(f"case cross-config ani={ani} bpc={bpc}", flush=True)
    try:
        for hist, y in ((da.PCF_XX, -1), (da.PCF_XY, cr)):
            ref_c, ref_w = None, None
            for m, lds, pc in itertools.product((1, 2, 3, 5), (0, 1, 3, 7), (0, 1, 3)):
                job = da.make_job(type_, hist, bins=bins, dmax=45.0, cat_x=cd, cat_y=y, bpc=bpc,
                                  box=300.0, cells_per_rmax=m, lds_copies_log2=lds,
                                  blocks_per_cu=pc)
                c, w = engine.histogram(job)
                if ref_c is None:
                    ref_c, ref_w = c, w
                    co, _ = O.pairs(d, None if y < 0 else r, ani=ani, bpc=bpc, box=300.0,
                                    bins=bins, dmax=45.0)
                    assert np.array_equal(c, co)
                assert np.array_equal(c, ref_c), (m, lds, pc)
                assert np.all(np.abs(w - ref_w) <= 1e-12 * np.abs(ref_w)), (m, lds, pc)
    finally:
        engine.release(cd)
        engine.release(cr)


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("bpc", [False, True])
def test_cross_config_equality_pair_kernels(engine, ani, bpc):
    # Unweighted 2iso: the Gram path (matrix cores, thin z cells) and the
    # direct path under every cell size and LDS layout give the same counts.
    d = synth.neyman_scott(60, 40, 6.0, 300.0, seed=103)
    r = synth.uniform(3000, 300.0, seed=104)
    cd, cr = engine.upload(d), engine.upload(r)
    print(f"case pair kernels ani={ani} bpc={bpc}", flush=True)
    type_ = da.PCF_2ANI if ani else da.PCF_2ISO
    bins = 12 if ani else 25
    try:
        for hist, y in ((da.PCF_XX, -1), (da.PCF_XY, cr)):
            co, _ = O.pairs(d, None if y < 0 else r, ani=ani, bpc=bpc, box=300.0, bins=bins,
                            dmax=45.0)
            for kern, m, lds in itertools.product((da.PCF_PAIR_GRAM, da.PCF_PAIR_DIRECT),
                                                  (1, 2, 3, 5), (0, 2, 7)):
                job = da.make_job(type_, hist, bins=bins, dmax=45.0, cat_x=cd, cat_y=y,
                                  bpc=bpc, box=300.0, cells_per_rmax=m, lds_copies_log2=lds,
                                  pair_kernel=kern)
                c, _ = engine.histogram(job)
                assert np.array_equal(c, co), (kern, m, lds)
    finally:
        engine.release(cd)
        engine.release(cr)


def test_autotune_single_job_is_cached_and_used(engine):
    d = synth.uniform(60000, 500.0, seed=111)
    cd = engine.upload(d)
    try:
        job = da.make_job(da.PCF_2ISO, da.PCF_XX, bins=30, dmax=40.0, cat_x=cd)
        ref, _ = engine.histogram(job)
        t = engine.autotune(job)
        print("tuning", t, flush=True)
        assert 1 <= t["cells_per_rmax"] <= 6 and t["trials"] >= 6
        assert t["pair_kernel"] in (da.PCF_PAIR_DIRECT, da.PCF_PAIR_GRAM)
        assert t["best_ms"] <= t["default_ms"] * 1.0001
        c, _ = engine.histogram(job)  # knobs 0: the cached choice applies
        assert engine.stats()["cells_per_rmax"] == t["cells_per_rmax"]
        assert np.array_equal(c, ref)
    finally:
        engine.release(cd)


def test_autotune_batch(engine):
    d = synth.neyman_scott(40, 32, 8.0, 1000.0, seed=121)
    rs = [synth.uniform(n, 1000.0, seed=130 + n) for n in (700, 2000, 5000)]
    cd = engine.upload(d)
    crs = [engine.upload(r) for r in rs]
    try:
        jobs = [da.make_job(da.PCF_2ANI, da.PCF_XX, bins=20, dmax=120.0, cat_x=cd)]
        for cr in crs:
            jobs.append(da.make_job(da.PCF_2ANI, da.PCF_XX, bins=20, dmax=120.0, cat_x=cr))
            jobs.append(da.make_job(da.PCF_2ANI, da.PCF_XY, bins=20, dmax=120.0, cat_x=cd,
                                    cat_y=cr))
        ref = engine.histogram_batch(jobs)
        t = engine.autotune(jobs)
        print("batch tuning", t, flush=True)
        res = engine.histogram_batch(jobs)
        assert engine.stats()["cells_per_rmax"] == t["cells_per_rmax"]
        for (c1, _), (c2, _) in zip(ref, res):
            assert np.array_equal(c1, c2)
    finally:
        engine.release(cd)
        for c in crs:
            engine.release(c)


def test_autotune_rejects_triplets(engine):
    d = synth.uniform(500, 100.0, seed=141)
    cd = engine.upload(d)
    try:
        with pytest.raises(da.PcfError):
            engine.autotune(da.make_job(da.PCF_3ISO, da.PCF_XXX, bins=10, dmax=20.0, cat_x=cd))
    finally:
        engine.release(cd)


def test_tuning_cache_file(tmp_path):
    """pcf_tune_save / pcf_tune_load / pcf_tune_lookup: a tuning written by
    one context is found by another after loading the file, and the histogram
    under the loaded tuning has the tuned run's counts; malformed files and
    files from another device model are refused.  (Fresh contexts: a
    context's cache holds every workload it tuned.)"""
    d = synth.uniform(3000, 250.0, seed=151)
    path = str(tmp_path / "tune.cache")
    with da.Engine(0) as e1:
        cd = e1.upload(d)
        job = da.make_job(da.PCF_2ISO, da.PCF_XX, bins=20, dmax=40.0, cat_x=cd)
        assert e1.tune_lookup(job) is None
        t = e1.autotune(job)
        assert e1.tune_lookup(job) == t
        c_ref, _ = e1.histogram(job)
        e1.tune_save(path)
    lines = open(path).read().splitlines()
    assert lines[0] == "pcf-tuning-cache 1" and lines[1].startswith("device ")
    assert len(lines) == 3 and "\t" in lines[2]
    with da.Engine(0) as e2:
        cd2 = e2.upload(d)
        job2 = da.make_job(da.PCF_2ISO, da.PCF_XX, bins=20, dmax=40.0, cat_x=cd2)
        assert e2.tune_lookup(job2) is None
        e2.tune_load(path)
        got = e2.tune_lookup(job2)
        assert got is not None
        for k in ("cells_per_rmax", "lds_copies_log2", "blocks_per_cu", "pair_kernel",
                  "trials"):
            assert got[k] == t[k], k
        c2, _ = e2.histogram(job2)
        assert np.array_equal(c2, c_ref)
        assert e2.stats()["cells_per_rmax"] == t["cells_per_rmax"]
        # Another job's workload is not matched by the entry.
        other = da.make_job(da.PCF_2ISO, da.PCF_XX, bins=21, dmax=40.0, cat_x=cd2)
        assert e2.tune_lookup(other) is None
        bad = tmp_path / "bad.cache"
        bad.write_text(lines[0] + "\n" + lines[1] + "\nkey-without-tab 1 2 3\n")
        with pytest.raises(da.PcfError):
            e2.tune_load(str(bad))
        alien = tmp_path / "alien.cache"
        alien.write_text(lines[0] + "\ndevice gfx000 1\n" + lines[2] + "\n")
        with pytest.raises(da.PcfError):
            e2.tune_load(str(alien))
        with pytest.raises(da.PcfError):
            e2.tune_load(str(tmp_path / "missing.cache"))
