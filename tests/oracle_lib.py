"""ctypes binding of the CPU oracle (oracle/liboracle_pcf.so).

Test infrastructure: imported only by tests/, the smoke check and the CPU
baseline leg of bench.py.  The oracle is the checker, never the product.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_pcf.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_i32 = ctypes.c_int


def _configure(L):
    L.pcf_oracle_pairs.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _i32, _i32, ctypes.c_float,
                                   _i32, ctypes.c_float, _i32, _u64p, _f64p]
    L.pcf_oracle_pairs_brute.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _i32, _i32,
                                         ctypes.c_float, _i32, ctypes.c_float, _u64p, _f64p]
    L.pcf_oracle_cpp32_2iso.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _i32, ctypes.c_float,
                                        _i32, _f64p]
    L.pcf_oracle_analytic_rr.argtypes = [_i32, _i64, ctypes.c_float, ctypes.c_float, _f64p]
    L.pcf_oracle_analytic_rr.restype = None
    if hasattr(L, "pcf_oracle_analytic_rr_refbugs"):
        L.pcf_oracle_analytic_rr_refbugs.argtypes = [_i32, _i64, ctypes.c_float, ctypes.c_float,
                                                     _f64p]
        L.pcf_oracle_analytic_rr_refbugs.restype = None
        L.pcf_oracle_analytic_3iso_refbugs.argtypes = [_f32p, _i64, _i32, ctypes.c_float,
                                                       ctypes.c_float, _i64, _i32, _i32, _i32,
                                                       _f64p, _f64p, _f64p, _f64p]
        L.pcf_oracle_triplets_refbugs.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _i32, _i32,
                                                  ctypes.c_float, _i32, _i32, ctypes.c_float,
                                                  _u64p, _f64p]
        L.pcf_oracle_xy2ani_alias_weights.argtypes = [_f32p, _i64, _f32p, _i64, ctypes.c_float,
                                                      _i32, _f32p]
        L.pcf_oracle_xy2ani_alias_weights.restype = None
    L.pcf_oracle_cpp32_3iso.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _i32, ctypes.c_float,
                                        _i32, _f64p]
    L.pcf_oracle_analytic_3iso.argtypes = [_f32p, _i64, _i32, ctypes.c_float, ctypes.c_float,
                                           _i64, _i32, _i32, _i32, _f64p, _f64p, _f64p,
                                           _f64p]
    for name in ("pcf_oracle_triplets", "pcf_oracle_triplets_brute"):
        if hasattr(L, name):
            getattr(L, name).argtypes = [_f32p, _i64, _f32p, _i64, _i32, _i32, _i32,
                                         ctypes.c_float, _i32, ctypes.c_float, _i32,
                                         _u64p, _f64p]
    return L


def load(path):
    """A configured handle on an oracle build at `path` (the mutation tests
    load deliberately broken builds of the oracle this way)."""
    return _configure(ctypes.CDLL(path))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = load(_LIB_PATH)
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _cat(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def pairs(a, b=None, *, ani=False, bpc=False, box=0.0, bins=30, dmax=80.0, threads=0,
          brute=False):
    """Ordered-pair histogram (counts uint64, weights float64) in the FP64
    canonical mode.  b=None means the auto histogram (XX)."""
    a = _cat(a)
    auto = b is None
    b = a if auto else _cat(b)
    n = bins * bins if ani else bins
    c = np.zeros(n, np.uint64)
    w = np.zeros(n, np.float64)
    if brute:
        rc = lib().pcf_oracle_pairs_brute(_p(a, _f32p), len(a), _p(b, _f32p), len(b), int(auto),
                                          int(ani), int(bpc), box, bins, dmax,
                                          _p(c, _u64p), _p(w, _f64p))
    else:
        rc = lib().pcf_oracle_pairs(_p(a, _f32p), len(a), _p(b, _f32p), len(b), int(auto),
                                    int(ani), int(bpc), box, bins, dmax, threads,
                                    _p(c, _u64p), _p(w, _f64p))
    if rc != 0:
        raise ValueError(f"oracle pairs rc={rc}")
    return c, w


def cpp32_2iso(a, b=None, *, bins=20, dmax=150.0, threads=0):
    a = _cat(a)
    auto = b is None
    b = a if auto else _cat(b)
    h = np.zeros(bins, np.float64)
    rc = lib().pcf_oracle_cpp32_2iso(_p(a, _f32p), len(a), _p(b, _f32p), len(b), int(auto), bins,
                                     dmax, threads, _p(h, _f64p))
    if rc != 0:
        raise ValueError(f"oracle cpp32 rc={rc}")
    return h


def cpp32_3iso(a, b=None, *, bins=30, dmax=50.0, threads=0):
    """The CPU program's 3iso histogram (after its FP32 symmetrisation)."""
    a = _cat(a)
    auto = b is None
    b = a if auto else _cat(b)
    h = np.zeros(bins ** 3, np.float64)
    rc = lib().pcf_oracle_cpp32_3iso(_p(a, _f32p), len(a), _p(b, _f32p), len(b), int(auto),
                                     bins, dmax, threads, _p(h, _f64p))
    if rc != 0:
        raise ValueError(f"oracle cpp32 3iso rc={rc}")
    return h


def analytic_rr(bins, np_, dmax, box, ref_bugs=False):
    """Analytic periodic RR; ref_bugs: the reference's 32-bit int overflows."""
    h = np.zeros(bins, np.float64)
    f = lib().pcf_oracle_analytic_rr_refbugs if ref_bugs else lib().pcf_oracle_analytic_rr
    f(bins, np_, dmax, box, _p(h, _f64p))
    return h


def analytic_3iso(a, *, bins, dmax, box, np_=0, ptt=100, bn_ref=200, threads=0, cubes=True,
                  ref_bugs=False):
    """Analytic periodic 3iso: (DDR, RRR, ff_av, ff_av_ref); cubes=False stops
    after ff_av / ff_av_ref (DDR and RRR are then None); ref_bugs: the
    reference's 32-bit int overflows in the fine RR."""
    a = _cat(a)
    ddr = np.zeros(bins ** 3, np.float64) if cubes else None
    rrr = np.zeros(bins ** 3, np.float64) if cubes else None
    ff = np.zeros(bins, np.float64)
    ffr = np.zeros(bins * bn_ref, np.float64)
    f = lib().pcf_oracle_analytic_3iso_refbugs if ref_bugs else lib().pcf_oracle_analytic_3iso
    rc = f(_p(a, _f32p), len(a), bins, dmax, box, int(np_ if np_ > 0 else len(a)), ptt, bn_ref,
           threads, _p(ddr, _f64p) if cubes else None, _p(rrr, _f64p) if cubes else None,
           _p(ff, _f64p), _p(ffr, _f64p))
    if rc != 0:
        raise ValueError(f"oracle analytic 3iso rc={rc}")
    return ddr, rrr, ff, ffr


def triplets(a, b=None, *, ani=False, bpc=False, box=0.0, bins=30, dmax=80.0, threads=0,
             brute=False, ddr=True, L=None):
    """Ordered-triplet histogram.  b=None: XXX over a.  Otherwise XXY with the
    pair (1,2) from a and point 3 from b.  L: an oracle handle from load()
    (default: the oracle build in oracle/)."""
    a = _cat(a)
    auto = b is None
    b = a if auto else _cat(b)
    n = bins ** 5 if ani else bins ** 3
    c = np.zeros(n, np.uint64)
    w = np.zeros(n, np.float64)
    L = lib() if L is None else L
    if brute:
        rc = L.pcf_oracle_triplets_brute(_p(a, _f32p), len(a), _p(b, _f32p), len(b), int(auto),
                                         int(ani), int(bpc), box, bins, dmax, 1,
                                         _p(c, _u64p), _p(w, _f64p))
    else:
        rc = L.pcf_oracle_triplets(_p(a, _f32p), len(a), _p(b, _f32p), len(b), int(auto),
                                   int(ani), int(bpc), box, bins, dmax, threads,
                                   _p(c, _u64p), _p(w, _f64p))
    if rc != 0:
        raise ValueError(f"oracle triplets rc={rc}")
    return c, w


def triplets_refbugs(a, b=None, *, ani=False, bpc=False, box, partitions=35, bins, dmax):
    """Triplet histogram with the reference GPU's cumulative triplet weight
    (literal form, small inputs)."""
    a = _cat(a)
    auto = b is None
    b = a if auto else _cat(b)
    n = bins ** 5 if ani else bins ** 3
    c = np.zeros(n, np.uint64)
    w = np.zeros(n, np.float64)
    rc = lib().pcf_oracle_triplets_refbugs(_p(a, _f32p), len(a), _p(b, _f32p), len(b), int(auto),
                                           int(ani), int(bpc), box, partitions, bins, dmax,
                                           _p(c, _u64p), _p(w, _f64p))
    if rc != 0:
        raise ValueError(f"oracle triplets_refbugs rc={rc}")
    return c, w


def xy2ani_alias_weights(d, r, *, box, partitions=35):
    """The data weights the reference's XY2ani reads for each random point."""
    d, r = _cat(d), _cat(r)
    out = np.zeros(len(r), np.float32)
    lib().pcf_oracle_xy2ani_alias_weights(_p(d, _f32p), len(d), _p(r, _f32p), len(r), box,
                                          partitions, _p(out, _f32p))
    return out


def estimator(hists, n_data=0.0, n_rand=0.0, normalize=False):
    """Landy-Szalay estimators as the reference's analysis scripts form them,
    elementwise in FP64 (numpy).  3 histograms (DD, DR, RR):
    (DD - 2 DR + RR) / RR with zeros of all three replaced by 1e-11
    (cpp/2PCFisotropic/normal/graph_2DPF.py:8-14).  4 histograms (DDD, DDR,
    DRR, RRR): (DDD - 3 DDR + 3 DRR - RRR) / RRR with NaN and inf set to 0
    (cpp/3PCFisotropic/normal/graph_3PCF.py:5-6,21-22).  normalize divides each
    histogram by its number of ordered pairs (triplets) of distinct points."""
    h = [np.array(x, dtype=np.float64, copy=True).ravel() for x in hists]
    nd, nr = float(n_data), float(n_rand)
    if len(h) == 3:
        if normalize:
            for x, s in zip(h, (nd * (nd - 1), nd * nr, nr * (nr - 1))):
                x /= s
        dd, dr, rr = h
        for x in (dd, rr, dr):
            x[x == 0] = 0.00000000001
        return (dd - 2 * dr + rr) / rr
    if normalize:
        for x, s in zip(h, (nd * (nd - 1) * (nd - 2), nd * (nd - 1) * nr, nd * nr * (nr - 1),
                            nr * (nr - 1) * (nr - 2))):
            x /= s
    ddd, ddr, drr, rrr = h
    with np.errstate(divide="ignore", invalid="ignore"):
        xi = (ddd - 3 * ddr + 3 * drr - rrr) / rrr
    xi[np.isnan(xi)] = 0.0
    xi[np.isinf(xi)] = 0.0
    return xi
