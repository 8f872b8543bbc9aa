"""GPU parity of the reference-defect mode (pcf_set_ref_bugs; PCF --ref-bugs)
against the oracle's reproductions (pinned by hand-worked cases in
tests/test_oracle_refbugs.py): the cumulative triplet weight on the
reference's node grid, the XY2ani weight read, the 32-bit analytic
overflows, and the command line's file names and 3ani slices."""
import os
import subprocess

import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu

WREL = 1e-12


@pytest.fixture
def ref_engine():
    e = da.Engine(0)
    e.set_ref_bugs(True)
    yield e
    e.close()


def _close(w, wo):
    assert np.all(np.abs(w - wo) <= WREL * np.maximum(np.abs(wo), 1e-300))


@pytest.mark.parametrize("ani", [False, True])
@pytest.mark.parametrize("bpc", [False, True])
@pytest.mark.parametrize("cross", [False, True])
@pytest.mark.parametrize("partitions", [3, 35])
def test_cumulative_triplet_weight(ref_engine, ani, bpc, cross, partitions):
    """Weighted triplets under the reference's chain: counts as always, sums
    equal to the oracle's literal restatement (3 partitions: long chains)."""
    d = synth.neyman_scott(6, 10, 3.0, 60.0, seed=81, weights="uniform")
    r = synth.uniform(50, 60.0, seed=82, weights="uniform")
    bins = 4 if ani else 6
    print(f"case refbugs 3{'ani' if ani else 'iso'} bpc={bpc} cross={cross} p={partitions}",
          flush=True)
    ca = ref_engine.upload(d)
    cb = ref_engine.upload(r) if cross else -1
    try:
        job = da.make_job(da.PCF_3ANI if ani else da.PCF_3ISO,
                          da.PCF_XXY if cross else da.PCF_XXX, bins=bins, dmax=15.0, cat_x=ca,
                          cat_y=cb, bpc=bpc, box=60.0, partitions=partitions)
        c, w = ref_engine.histogram(job)
    finally:
        ref_engine.release(ca)
        if cb >= 0:
            ref_engine.release(cb)
    co, wo = O.triplets_refbugs(d, r if cross else None, ani=ani, bpc=bpc, box=60.0,
                                partitions=partitions, bins=bins, dmax=15.0)
    c0, w0 = O.triplets(d, r if cross else None, ani=ani, bpc=bpc, box=60.0, bins=bins,
                        dmax=15.0)
    assert co.sum() > 0
    assert np.array_equal(c, co) and np.array_equal(c, c0)
    _close(w, wo)
    if partitions == 3:
        assert np.abs(wo - w0).max() > 1e-3 * np.abs(w0).max()


def test_xy2ani_weight(ref_engine):
    """2ani DR without periodic boundaries: each random point carries the data
    weight of its ordered slot; other histograms are unchanged."""
    d = synth.neyman_scott(20, 20, 4.0, 100.0, seed=83, weights="uniform")
    r = synth.uniform(400, 100.0, seed=84, weights="uniform")
    alias = O.xy2ani_alias_weights(d, r, box=100.0, partitions=7)
    ra = r.copy()
    ra[:, 3] = alias
    ca, cb = ref_engine.upload(d), ref_engine.upload(r)
    try:
        job = da.make_job(da.PCF_2ANI, da.PCF_XY, bins=10, dmax=30.0, cat_x=ca, cat_y=cb,
                          box=100.0, partitions=7)
        c, w = ref_engine.histogram(job)
        co, wo = O.pairs(d, ra, ani=True, bins=10, dmax=30.0)
        assert np.array_equal(c, co)
        _close(w, wo)
        # XX and periodic DR keep their true weights.
        for j, (a, b, bpc) in ((da.make_job(da.PCF_2ANI, da.PCF_XX, bins=10, dmax=30.0,
                                            cat_x=ca, box=100.0, partitions=7), (d, None, False)),
                               (da.make_job(da.PCF_2ANI, da.PCF_XY, bins=10, dmax=30.0,
                                            cat_x=ca, cat_y=cb, bpc=True, box=100.0,
                                            partitions=7), (d, r, True))):
            c, w = ref_engine.histogram(j)
            co, wo = O.pairs(a, b, ani=True, bpc=bpc, box=100.0, bins=10, dmax=30.0)
            assert np.array_equal(c, co)
            _close(w, wo)
        # The alias is formed per call: batches refuse such jobs.
        with pytest.raises(da.PcfError) as ei:
            ref_engine.histogram_batch([job, job])
        assert ei.value.rc == da.PCF_ERR_UNSUPPORTED
    finally:
        ref_engine.release(ca)
        ref_engine.release(cb)
    # A random catalog longer than the data catalog would read past it.
    d2 = synth.uniform(100, 100.0, seed=85, weights="uniform")
    ca, cb = ref_engine.upload(d2), ref_engine.upload(r)
    try:
        with pytest.raises(da.PcfError) as ei:
            ref_engine.histogram(da.make_job(da.PCF_2ANI, da.PCF_XY, bins=10, dmax=30.0,
                                             cat_x=ca, cat_y=cb, box=100.0, partitions=7))
        assert ei.value.rc == da.PCF_ERR_UNSUPPORTED
    finally:
        ref_engine.release(ca)
        ref_engine.release(cb)


@pytest.mark.parametrize("np_,bins", [(46341, 30), (65536, 30), (1000, 26800)])
def test_analytic_rr_overflows(ref_engine, np_, bins):
    g = ref_engine.analytic_rr(bins, np_, 80.0, 250.0)
    o = O.analytic_rr(bins, np_, 80.0, 250.0, ref_bugs=True)
    assert np.array_equal(g, o)


def test_analytic_3iso_overflows(ref_engine):
    """np = 46341 (np^2 wraps negative) and 600 x 50 fine bins (the shell count
    wraps from bin 26755 on)."""
    d = synth.uniform(2000, 250.0, seed=86)
    c = ref_engine.upload(d)
    try:
        g = ref_engine.analytic_3iso(c, bins=6, dmax=60.0, box=250.0, np_=46341, ptt=100,
                                     bn_ref=50)
    finally:
        ref_engine.release(c)
    o = O.analytic_3iso(d, bins=6, dmax=60.0, box=250.0, np_=46341, ptt=100, bn_ref=50,
                        ref_bugs=True)
    assert np.array_equal(g[2], o[2]) and np.array_equal(g[3], o[3])
    for a, b in ((g[0], o[0]), (g[1], o[1])):
        assert np.all(np.abs(a - b) <= 1e-12 * np.maximum(np.abs(b), 1e-300))


def _write(path, pts):
    with open(path, "w") as f:
        for p in pts:
            f.write("%.9g %.9g %.9g %.9g\n" % tuple(p))


def _load(path):
    return np.array(open(path).read().split(), dtype=np.float64)


def test_command_line_ref_bugs(tmp_path):
    """PCF 3iso and 3ani -rd --ref-bugs on weighted catalogs (files written
    here): DDD equals the oracle's chain-weighted DDD, DDR of the second
    random file has the accumulated name, and 3ani random file 1's file holds
    values bins^3 .. 2 bins^3 - 1 of the first random file's DDR."""
    data = tmp_path / "data"
    (data / "rand").mkdir(parents=True)
    d = synth.neyman_scott(5, 10, 3.0, 60.0, seed=87, weights="uniform")
    r0 = synth.uniform(50, 60.0, seed=88, weights="uniform")
    r1 = synth.uniform(50, 60.0, seed=89, weights="uniform")
    _write(data / "d.dat", d)
    _write(data / "rand" / "r0.dat", r0)
    _write(data / "rand" / "r1.dat", r1)
    d, r0, r1 = (np.loadtxt(p, dtype=np.float32) for p in
                 (data / "d.dat", data / "rand" / "r0.dat", data / "rand" / "r1.dat"))
    names = [n for n in os.listdir(data / "rand")]  # readdir order, as --ref-bugs keeps it
    for kind, bins in (("3iso", 5), ("3ani", 3)):
        res = tmp_path / f"res_{kind}"
        res.mkdir()
        env = dict(os.environ, PCF_DATA_DIR=str(data), PCF_RESULTS_DIR=str(res))
        args = [da.CLI_PATH, kind, "-n", "50", "-f", "d.dat", "-rd", "rand/", "-b", str(bins),
                "-d", "15", "-s", "60", "-p", "4", "--ref-bugs"]
        print("cli", " ".join(args[1:]), flush=True)
        p = subprocess.run(args, capture_output=True, text=True, env=env, timeout=300)
        assert p.returncode == 0, p.stdout + p.stderr
        out = sorted(os.listdir(res))
        tag = "iso_" if kind == "3iso" else "ani_"
        assert f"DDR{tag}{names[0]}{names[1]}" in out, out
        ani = kind == "3ani"
        c, w = O.triplets_refbugs(d, None, ani=ani, box=60.0, partitions=4, bins=bins, dmax=15.0)
        b3 = bins ** 3
        got = _load(res / f"DDD{tag}d.dat")
        # Weighted sums are printed with 12 significant digits (a relative
        # rounding of up to 5e-12).
        assert np.allclose(got, w[:b3], rtol=1e-11, atol=0)
        rfirst = {"r0.dat": r0, "r1.dat": r1}[names[0]]
        c, w = O.triplets_refbugs(d, rfirst, ani=ani, box=60.0, partitions=4, bins=bins,
                                  dmax=15.0)
        got = _load(res / f"DDR{tag}{names[0]}{names[1]}")
        want = w[b3:2 * b3] if ani else O.triplets_refbugs(
            d, {"r0.dat": r0, "r1.dat": r1}[names[1]], box=60.0, partitions=4, bins=bins,
            dmax=15.0)[1]
        assert np.allclose(got, want, rtol=1e-11, atol=0)
