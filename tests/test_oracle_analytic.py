"""CPU checks that pin the analytic periodic 3iso oracle (PCF 3iso -bpc -a).

* ff_av / ff_av_ref vs the reference's compiled CPU program
  (cpp/3PCFisotropic/BPC, which computes exactly these two arrays on its
  hard-coded configuration: the first 5000 points of data.dat, 30 bins, dmax 60,
  box 250, ptt 100, bn_ref 200; its RRR/DDR loop is commented out).  The
  program bins FP32 distances into 600,000 fine bins, so pairs within an FP32
  rounding of a fine-bin edge land in the neighbouring bin: the comparison
  allows that noise and nothing systematic.
* physics: the analytic RRR is the expected ordered-triplet count of uniform
  randoms, 8 pi^2 N n^2 r1 r2 r3 dr^3 (n = N/L^3), so it must match the
  counted XXX of a uniform periodic catalog; the analytic DDR is the expected
  XXY(D, R) of a uniform R with |R| = |D| averaged over which side joins the
  two data points, so its total must match the counted XXY.
* the closed form of the bin cubes whose 8 corners all form triangles
  (the integral of r1 r2 r3 over the cube), and for the partial cubes the
  monotone convergence of the sub-cube sum to the integral over the part of the
  cube that satisfies the triangle inequality (an independent midpoint rule).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import GOLDEN, catalog
from desi_amd import synth

L = 250.0


def test_ff_av_matches_cpu_program():
    d = catalog("data.dat")[:5000]
    _, _, ff, ffr = O.analytic_3iso(d, bins=30, dmax=60.0, box=L, ptt=100, bn_ref=200,
                                    cubes=False)
    ref_dir = os.path.join(GOLDEN, "ref_out", "3iso_bpca")
    rff = np.loadtxt(os.path.join(ref_dir, "ff_av.dat"))
    rffr = np.loadtxt(os.path.join(ref_dir, "ff_av_ref.dat"))
    assert ff.shape == rff.shape and ffr.shape == rffr.shape
    rel = np.abs(ff - rff) / np.abs(rff)
    # 30 coarse bins: all but the bins touched by fine-edge flips agree to the
    # FP32-vs-FP64 constants (dr, beta) of the two programs.
    assert np.sum(rel > 1e-6) <= 3, rel
    assert rel.max() < 1e-4
    rel2 = np.abs(ffr - rffr) / np.maximum(np.abs(rffr), 1e-300)
    assert np.median(rel2) < 1e-6
    assert np.sum(np.abs(ffr - rffr)) <= 1e-3 * np.sum(np.abs(rffr))


def test_rrr_is_the_uniform_triplet_count():
    bins, dmax = 10, 60.0
    r = synth.uniform(2000, L, seed=5)
    _, rrr, _, _ = O.analytic_3iso(r, bins=bins, dmax=dmax, box=L, ptt=20, bn_ref=16)
    c, _ = O.triplets(r, None, bpc=True, box=L, bins=bins, dmax=dmax)
    c = c.astype(np.float64)
    assert abs(c.sum() / rrr.sum() - 1) < 0.02
    m = c > 1e4
    assert m.sum() > 100
    assert np.median(np.abs(c[m] / rrr[m] - 1)) < 0.05
    # zero exactly where no triangle fits
    assert np.all(c[rrr == 0] == 0)


def test_ddr_is_the_expected_mixed_count():
    bins, dmax = 10, 60.0
    d = synth.neyman_scott(80, 25, 5.0, L, seed=6)
    r = synth.uniform(len(d), L, seed=7)
    ddr, rrr, _, _ = O.analytic_3iso(d, bins=bins, dmax=dmax, box=L, ptt=20, bn_ref=16)
    c, _ = O.triplets(d, r, bpc=True, box=L, bins=bins, dmax=dmax)
    assert abs(c.sum() / ddr.sum() - 1) < 0.03
    # clustering raises DDR above RRR
    assert ddr.sum() > 1.05 * rrr.sum()


def _corners_in(r1, r2, r3, h):
    v = 0
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                s1, s2, s3 = r1 + a * h, r2 + b * h, r3 + c * h
                v += (s1 + s2 >= s3) and (s1 + s3 >= s2) and (s2 + s3 >= s1)
    return v


def test_full_cubes_closed_form():
    bins, dmax, n = 8, 60.0, 1500
    r = synth.uniform(n, L, seed=8)
    _, rrr, _, _ = O.analytic_3iso(r, bins=bins, dmax=dmax, box=L, ptt=10, bn_ref=4)
    rrr = rrr.reshape(bins, bins, bins)
    gama = 8 * np.pi ** 2 * n * (n / L ** 3) ** 2
    dr = dmax / bins
    nfull = 0
    for i in range(bins):
        for j in range(bins):
            for k in range(bins):
                if _corners_in(i * dr, j * dr, k * dr, dr) != 8:
                    continue
                nfull += 1
                integral = np.prod([((x + 1) ** 2 - x ** 2) * dr * dr / 2 for x in (i, j, k)])
                assert rrr[i, j, k] == pytest.approx(gama * integral, rel=1e-12)
    assert nfull > 100


def test_partial_cubes_converge_to_the_triangle_integral():
    bins, dmax, n = 6, 60.0, 1000
    r = synth.uniform(n, L, seed=9)
    gama = 8 * np.pi ** 2 * n * (n / L ** 3) ** 2
    dr = dmax / bins
    res = {}
    for bn_ref in (4, 8, 16):
        _, rrr, _, _ = O.analytic_3iso(r, bins=bins, dmax=dmax, box=L, ptt=10, bn_ref=bn_ref)
        res[bn_ref] = rrr.reshape(bins, bins, bins)
    m = 80  # midpoint rule inside each cube
    t = (np.arange(m) + 0.5) / m * dr
    checked = 0
    for i in range(bins):
        for j in range(bins):
            for k in range(bins):
                v = _corners_in(i * dr, j * dr, k * dr, dr)
                if v == 0 or v == 8:
                    continue
                a = i * dr + t[:, None, None]
                b = j * dr + t[None, :, None]
                c = k * dr + t[None, None, :]
                inside = (a + b >= c) & (a + c >= b) & (b + c >= a)
                integral = gama * np.sum(np.where(inside, a * b * c, 0.0)) * (dr / m) ** 3
                r4, r8, r16 = res[4][i, j, k], res[8][i, j, k], res[16][i, j, k]
                assert r4 <= r8 * (1 + 1e-12) and r8 <= r16 * (1 + 1e-12)
                assert r16 <= integral * 1.02 + 1e-9
                gap8, gap16 = integral - r8, integral - r16
                if gap8 > 1e-3 * integral:
                    # first-order convergence: halving the sub-cube halves the gap
                    assert gap16 < 0.75 * gap8
                checked += 1
    assert checked > 20
