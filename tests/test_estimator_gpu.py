"""GPU parity of pcf_estimator (Landy-Szalay estimators on the device) with
the oracle's restatement of the reference's analysis scripts: elementwise FP64
with the same operation order, so results agree to the last bit."""
import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("normalize", [False, True])
def test_random_histograms(engine, normalize):
    rng = np.random.default_rng(7)
    for k in (3, 4):
        hs = [rng.integers(0, 50, 5000).astype(np.float64) for _ in range(k)]
        print(f"case k={k} normalize={normalize}", flush=True)
        g = engine.estimator(hs, 1000, 3000, normalize)
        o = O.estimator(hs, 1000, 3000, normalize)
        assert np.array_equal(g, o)


def test_end_to_end_clustered(engine):
    """DD, DR, RR of a clustered catalog against uniform randoms, then the
    estimator, on the GPU; the oracle's counts give the same xi, and the
    clustering shows as xi > 0 at small separations."""
    d = synth.neyman_scott(200, 20, 3.0, 200.0, seed=61)
    r = synth.uniform(4000, 200.0, seed=62)
    cd, cr = engine.upload(d), engine.upload(r)
    try:
        hs = []
        for hist, x, y in ((da.PCF_XX, cd, -1), (da.PCF_XY, cd, cr), (da.PCF_XX, cr, -1)):
            c, _ = engine.histogram(da.make_job(da.PCF_2ISO, hist, bins=20, dmax=40.0, cat_x=x,
                                                cat_y=y))
            hs.append(c.astype(np.float64))
        xi = engine.estimator(hs, len(d), len(r), True)
    finally:
        engine.release(cd)
        engine.release(cr)
    ho = [O.pairs(a, b, bins=20, dmax=40.0)[0].astype(np.float64)
          for a, b in ((d, None), (d, r), (r, None))]
    assert np.array_equal(xi, O.estimator(ho, len(d), len(r), True))
    assert xi[0] > 1.0 and xi[1] > 0.5


def test_bad_arguments(engine):
    with pytest.raises(ValueError):
        engine.estimator([np.ones(3), np.ones(3)])
    with pytest.raises(da.PcfError):
        engine.estimator([np.ones(3)] * 3, 1, 5, True)  # fewer points than the order
