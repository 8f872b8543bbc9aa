"""The FP64 oracle against the outputs of the reference's compiled CPU programs
(tests/golden/ref_out, written by tests/golden/make_ref_golden.sh) at their
hard-coded configurations (N = 32,768 of data.dat, dmax 150, box 250).

Those programs form distances in FP32, so pairs within an FP32 rounding of a
bin edge or of dmax can land one bin over: the totals agree to the pairs at
the range limit and the per-bin differences are that edge noise (a few per
million).  What these pin that the brute-force and lattice checks do not is the
reference's own choice of images, layout and zero rules on real catalogs,
including dmax > L/2, where the regular and the periodic image of a pair are
both in range and both count."""
import os

import numpy as np

import oracle_lib as O
from conftest import GOLDEN, catalog


def _ref(prog, name):
    return np.loadtxt(os.path.join(GOLDEN, "ref_out", prog, name)).ravel()


def _edge_noise_only(c, ref):
    c = c.astype(np.float64)
    assert c.shape == ref.shape
    assert abs(c.sum() - ref.sum()) <= 1e-6 * ref.sum()
    assert np.abs(c - ref).sum() <= 1e-5 * ref.sum()
    assert np.array_equal(c == 0, ref == 0)


def test_2ani_non_periodic():
    # cpp/2PCFanisotropic/normal/main.cpp: bn 20, dmax 150
    d = catalog("data.dat")
    c, _ = O.pairs(d, None, ani=True, bins=20, dmax=150.0)
    _edge_noise_only(c, _ref("2ani_normal", "DDani_mesh_3D_ref.dat"))


def test_2ani_periodic_dmax_above_half_box():
    # cpp/2PCFanisotropic/BPC/main.cpp: bn 50, dmax 150, box 250
    d = catalog("data.dat")
    c, _ = O.pairs(d, None, ani=True, bpc=True, box=250.0, bins=50, dmax=150.0)
    _edge_noise_only(c, _ref("2ani_bpc", "DDani_mesh_3D_ref.dat"))


def test_2iso_analytic_rr():
    # cpp/2PCFisotropic/BPC_analitic: RR analytic (double constants there,
    # FP32 dr and beta in the GPU reference the oracle follows)
    rr = O.analytic_rr(30, 32768, 150.0, 250.0)
    ref = _ref("2iso_bpca", "RRiso_mesh_3D_ref.dat")
    assert np.allclose(rr, ref, rtol=2e-7, atol=0)
