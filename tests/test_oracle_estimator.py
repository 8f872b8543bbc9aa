"""The oracle's estimator restatement pinned by identities the formulas fix:
no excess over randoms gives 0, a doubled DD gives 1, the reference scripts'
zero and NaN guards, and the pair / triplet normalisations."""
import numpy as np

import oracle_lib as O


def test_ls_2pcf_identities():
    rr = np.array([4.0, 10.0, 7.0, 0.0])
    # DD = DR = RR: no clustering, xi = 0 (the zero bin becomes 1e-11 / 1e-11).
    assert np.array_equal(O.estimator((rr, rr, rr)), np.zeros(4))
    # DD = 2 RR, DR = RR: xi = (2RR - 2RR + RR) / RR = 1.
    xi = O.estimator((2 * rr, rr, rr))
    assert np.array_equal(xi[:3], np.ones(3))
    # Zero guard (graph_2DPF.py:12-14): DD = 0, DR = RR = 5 -> (1e-11 - 10 + 5) / 5.
    assert O.estimator(([0.0], [5.0], [5.0]))[0] == (1e-11 - 2 * 5.0 + 5.0) / 5.0
    # Zero RR only: the denominator is 1e-11.
    assert O.estimator(([3.0], [2.0], [0.0]))[0] == (3.0 - 4.0 + 1e-11) / 1e-11


def test_ls_3pcf_identities():
    r = np.array([8.0, 3.0, 0.0, 1.0])
    # All equal: (1 - 3 + 3 - 1) R / R = 0; the empty bin's NaN -> 0.
    assert np.array_equal(O.estimator((r, r, r, r)), np.zeros(4))
    # DDD = 2RRR, DDR = DRR = RRR: xi = 1.
    xi = O.estimator((2 * r, r, r, r))
    assert np.array_equal(xi, np.array([1.0, 1.0, 0.0, 1.0]))
    # RRR = 0 with a nonzero numerator: inf -> 0 (graph_3PCF.py:22).
    assert O.estimator(([5.0], [1.0], [1.0], [0.0]))[0] == 0.0


def test_normalisation():
    # Uniform expectation: DD ~ N_D(N_D-1) p, DR ~ N_D N_R p, RR ~ N_R(N_R-1) p.
    nd, nr, p = 1000.0, 4000.0, np.array([1e-4, 3e-4])
    xi = O.estimator((nd * (nd - 1) * p, nd * nr * p, nr * (nr - 1) * p), nd, nr, True)
    assert np.allclose(xi, 0.0, atol=1e-12)
    xi3 = O.estimator((nd * (nd - 1) * (nd - 2) * p, nd * (nd - 1) * nr * p,
                       nd * nr * (nr - 1) * p, nr * (nr - 1) * (nr - 2) * p), nd, nr, True)
    assert np.allclose(xi3, 0.0, atol=1e-12)
