"""End to end through the drop-in command line (desi_amd/bin/PCF): the files it
writes hold the oracle's histograms (counts printed with 12 significant digits,
so integers below 1e12 are exact), under the reference's file names."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import desi_amd as da
import oracle_lib as O
from conftest import GOLDEN, catalog

pytestmark = pytest.mark.gpu

CATS = os.path.join(GOLDEN, "catalogs")


def _run(tmp_path, *args, data_dir=CATS):
    res = tmp_path / "results"
    res.mkdir(exist_ok=True)
    env = dict(os.environ, PCF_DATA_DIR=str(data_dir), PCF_RESULTS_DIR=str(res))
    print("cli", " ".join(args), flush=True)
    p = subprocess.run([da.CLI_PATH, *args], capture_output=True, text=True, env=env,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return res


def _load(path):
    return np.array(open(path).read().split(), dtype=np.float64)


def test_2iso_files(tmp_path):
    res = _run(tmp_path, "2iso", "-n", "1000", "-f", "data_1K.dat", "-r", "rand0_1K.dat",
               "-b", "30", "-d", "80")
    d, r = catalog("data_1K.dat"), catalog("rand0_1K.dat")
    box = float(np.ceil(max(d[:, :3].max(), r[:, :3].max()) + 1))
    for name, (a, b) in (("DDiso_data_1K.dat", (d, None)), ("RRiso_rand0_1K.dat", (r, None)),
                         ("DRiso_rand0_1K.dat", (d, r))):
        c, _ = O.pairs(a, b, bins=30, dmax=80.0, box=box)
        assert np.array_equal(_load(res / name), c.astype(np.float64)), name


def test_2ani_bpc_files(tmp_path):
    res = _run(tmp_path, "2ani", "-n", "500", "-f", "data_500.dat", "-r", "rand0_500.dat",
               "-b", "15", "-d", "60", "-bpc", "-s", "250")
    d, r = catalog("data_500.dat"), catalog("rand0_500.dat")
    c, _ = O.pairs(d, r, ani=True, bpc=True, box=250.0, bins=15, dmax=60.0)
    assert np.array_equal(_load(res / "DRani_BPC_rand0_500.dat"), c.astype(np.float64))


def test_3iso_bpc_files(tmp_path):
    res = _run(tmp_path, "3iso", "-n", "1000", "-f", "data_1K.dat", "-r", "rand0_1K.dat",
               "-b", "20", "-d", "60", "-bpc", "-s", "250")
    d, r = catalog("data_1K.dat"), catalog("rand0_1K.dat")
    for name, (a, b) in (("DDDiso_BPC_data_1K.dat", (d, None)),
                         ("DDRiso_BPC_rand0_1K.dat", (d, r)),
                         ("DRRiso_BPC_rand0_1K.dat", (r, d))):
        c, _ = O.triplets(a, b, bpc=True, box=250.0, bins=20, dmax=60.0)
        assert np.array_equal(_load(res / name), c.astype(np.float64)), name


def test_random_directory(tmp_path):
    """-rd: every file of the directory is a random catalog, in sorted order,
    each giving its own RR and DR files."""
    rd = tmp_path / "cats"
    (rd / "rands").mkdir(parents=True)
    shutil.copy(os.path.join(CATS, "data_500.dat"), rd / "data_500.dat")
    shutil.copy(os.path.join(CATS, "rand0_500.dat"), rd / "rands" / "b.dat")
    r2 = catalog("rand0_1K.dat")[500:1000]
    np.savetxt(rd / "rands" / "a.dat", r2, fmt="%.6f")
    res = _run(tmp_path, "2iso", "-n", "500", "-f", "data_500.dat", "-rd", "rands/", "-b", "10",
               "-d", "50", "-bpc", "-s", "250", data_dir=rd)
    d = catalog("data_500.dat")
    for fname, r in (("a.dat", np.loadtxt(rd / "rands" / "a.dat", dtype=np.float32)),
                     ("b.dat", catalog("rand0_500.dat"))):
        c, _ = O.pairs(r, None, bpc=True, box=250.0, bins=10, dmax=50.0)
        assert np.array_equal(_load(res / f"RRiso_BPC_{fname}"), c.astype(np.float64))
        c, _ = O.pairs(d, r, bpc=True, box=250.0, bins=10, dmax=50.0)
        assert np.array_equal(_load(res / f"DRiso_BPC_{fname}"), c.astype(np.float64))


def test_3iso_analytic_files(tmp_path, engine):
    res = _run(tmp_path, "3iso", "-n", "5000", "-f", "data.dat", "-b", "30", "-d", "60", "-bpc",
               "-a", "-s", "250")
    d = catalog("data.dat")[:5000]
    ddr = _load(res / "DDRiso_BPCanalytic_data.dat")
    assert np.array_equal(ddr, _load(res / "DRRiso_BPCanalytic_data.dat"))
    rrr = _load(res / "RRRiso_BPCanalytic_data.dat")
    c = engine.upload(d)
    try:
        g_ddr, g_rrr, _, _ = engine.analytic_3iso(c, bins=30, dmax=60.0, box=250.0)
        job = da.make_job(da.PCF_3ISO, da.PCF_XXX, bins=30, dmax=60.0, cat_x=c, bpc=True,
                          box=250.0)
        ddd, _ = engine.histogram(job)
    finally:
        engine.release(c)
    np.testing.assert_allclose(ddr, g_ddr, rtol=1e-11, atol=0)
    np.testing.assert_allclose(rrr, g_rrr, rtol=1e-11, atol=0)
    assert np.array_equal(_load(res / "DDDiso_BPCanalytic_data.dat"), ddd.astype(np.float64))


def test_estimator_files(tmp_path):
    """-e writes the Landy-Szalay estimator beside the histograms, equal (to
    the 12 printed digits) to the scripts' formula on the oracle's counts."""
    res = _run(tmp_path, "2iso", "-n", "1000", "-f", "data_1K.dat", "-r", "rand0_1K.dat",
               "-b", "30", "-d", "80", "-bpc", "-s", "250", "-e")
    d, r = catalog("data_1K.dat"), catalog("rand0_1K.dat")
    hs = [O.pairs(a, b, bpc=True, box=250.0, bins=30, dmax=80.0)[0].astype(np.float64)
          for a, b in ((d, None), (d, r), (r, None))]
    xi = O.estimator(hs)
    np.testing.assert_allclose(_load(res / "LSiso_BPC_rand0_1K.dat"), xi, rtol=1e-11, atol=1e-12)
    res = _run(tmp_path, "3iso", "-n", "500", "-f", "data_500.dat", "-r", "rand0_500.dat",
               "-b", "8", "-d", "40", "-e")
    d, r = catalog("data_500.dat"), catalog("rand0_500.dat")
    box = float(np.ceil(max(d[:, :3].max(), r[:, :3].max()) + 1))
    hs = [O.triplets(a, b, box=box, bins=8, dmax=40.0)[0].astype(np.float64)
          for a, b in ((d, None), (d, r), (r, d), (r, None))]
    xi3 = O.estimator(hs)
    np.testing.assert_allclose(_load(res / "LS3iso_rand0_500.dat"), xi3, rtol=1e-11, atol=1e-12)


def test_3ani_full_5d_file(tmp_path):
    """PCF_SAVE_5D=1 writes the whole bins^5 histogram in save_histogram5D's
    layout (cuda/create_grid.cuh:230-266); by default the bins^3 slab."""
    args = ("3ani", "-n", "300", "-f", "data_500.dat", "-r", "rand0_500.dat", "-b", "5", "-d",
            "40", "-bpc", "-s", "250")
    res = _run(tmp_path, *args)
    slab = _load(res / "DDDani_BPC_data_500.dat")
    res5 = tmp_path / "r5"
    res5.mkdir()
    env = dict(os.environ, PCF_DATA_DIR=CATS, PCF_RESULTS_DIR=str(res5), PCF_SAVE_5D="1")
    p = subprocess.run([da.CLI_PATH, *args], capture_output=True, text=True, env=env,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    full = _load(res5 / "DDDani_BPC_data_500.dat")
    d = catalog("data_500.dat")[:300]
    c, _ = O.triplets(d, None, ani=True, bpc=True, box=250.0, bins=5, dmax=40.0)
    assert np.array_equal(full, c.astype(np.float64))
    assert np.array_equal(slab, full[:5 ** 3])


def test_2ani_bpc_2dmax_ge_box_files(tmp_path):
    """`PCF 2ani -bpc -d 150 -s 250 -b 50` (the reference CPU program's
    configuration, 2*dmax > box) on the full sample catalogs: the written
    DD/RR/DR equal the oracle's (tests/golden/oracle_out)."""
    res = _run(tmp_path, "2ani", "-n", "32768", "-f", "data.dat", "-r", "rand0.dat",
               "-b", "50", "-d", "150", "-bpc", "-s", "250")
    fix = os.path.join(GOLDEN, "oracle_out")
    for name, ref in (("DDani_BPC_data.dat", "DD"), ("RRani_BPC_rand0.dat", "RR"),
                      ("DRani_BPC_rand0.dat", "DR")):
        want = np.loadtxt(os.path.join(fix, f"2ani_bpc_d150_s250_b50_{ref}.txt"))
        assert np.array_equal(_load(res / name), want), name


def test_tuning_cache_env(tmp_path):
    """PCF_TUNE_CACHE + PCF_AUTOTUNE=1: the first run tunes its pair
    histograms and writes the cache, the second reuses it; both runs write the
    same histograms as a run without tuning."""
    cache = tmp_path / "tune.cache"
    args = ("2iso", "-n", "5000", "-f", "data_5K.dat", "-r", "rand0_5K.dat", "-b", "20", "-d",
            "60")
    (tmp_path / "plain").mkdir()
    base = _run(tmp_path / "plain", *args)
    outs = []
    for k in range(2):
        res = tmp_path / f"t{k}" / "results"
        res.mkdir(parents=True)
        env = dict(os.environ, PCF_DATA_DIR=CATS, PCF_RESULTS_DIR=str(res),
                   PCF_TUNE_CACHE=str(cache), PCF_AUTOTUNE="1")
        print("cli tuned run", k, flush=True)
        p = subprocess.run([da.CLI_PATH, *args], capture_output=True, text=True, env=env,
                           timeout=300)
        assert p.returncode == 0, p.stdout + p.stderr
        assert "ignoring" not in p.stderr, p.stderr
        outs.append(res)
        lines = cache.read_text().splitlines()
        assert lines[0] == "pcf-tuning-cache 1"
        # DD, RR and DR: three workloads (RR and DD share sizes and differ in
        # nothing the signature holds, so they may share one entry).
        assert 2 <= len(lines) - 2 <= 3, lines
    for name in os.listdir(base):
        ref = _load(base / name)
        for res in outs:
            assert np.array_equal(_load(res / name), ref), name


@pytest.mark.parametrize("kind", ["3iso", "3ani"])
def test_triplets_bpc_2dmax_ge_box_files(tmp_path, kind):
    """`PCF 3iso|3ani -bpc -d 150 -s 250` (2*dmax > box): the written
    histograms equal the oracle's (a triplet counts under every combination of
    regular and projected images in range)."""
    bins = 6 if kind == "3iso" else 4
    res = _run(tmp_path, kind, "-n", "100", "-f", "data_500.dat", "-r", "rand0_500.dat",
               "-b", str(bins), "-d", "150", "-bpc", "-s", "250")
    d, r = catalog("data_500.dat")[:100], catalog("rand0_500.dat")[:100]
    tag = "iso" if kind == "3iso" else "ani"
    for name, (a, b) in ((f"DDD{tag}_BPC_data_500.dat", (d, None)),
                         (f"DDR{tag}_BPC_rand0_500.dat", (d, r)),
                         (f"DRR{tag}_BPC_rand0_500.dat", (r, d)),
                         (f"RRR{tag}_BPC_rand0_500.dat", (r, None))):
        c, _ = O.triplets(a, b, ani=kind == "3ani", bpc=True, box=250.0, bins=bins,
                          dmax=150.0)
        got = _load(res / name)
        # 3ani files hold the first bins^3 values of the bins^5 histogram.
        assert np.array_equal(got, c[:got.size].astype(np.float64)), name
