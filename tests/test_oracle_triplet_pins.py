"""Checks that pin the triplet oracle's angle bins and periodic terms
independently of the oracle's own code, and a mutation test showing that each
check catches a plausible mistake.

The reference rules being pinned:

* 3ani layout and angle bins (cuda/kernels/3ani_k.cuh:75-92): deposit at
  [b(d12)][b(d31)][b(d23)][t][p] with t = (int)((1 + dz12/d12) * bn/2) and
  p = (int)((1 + dz31/d31) * bn/2), skipped when t or p > bn - 1 (cos = 1).
* periodic 3ani (cuda/kernels/3aniBPC_k.cuh:57,73): dz = |z2 - z1| - L*f, so
  an unprojected pair has dz >= 0 whatever its direction, and a projected one
  dz < 0.
* periodic triplet sides (cuda/kernels/3isoBPC_k.cuh:56-75):
  |d12| - L*f2, |d31| - L*f3, |d23| - L*|f2 - f3| per axis.  With
  3*dmax < L this is the minimum-image distance of every side (every side in
  range is then shorter than L - 2*dmax on each axis, so the consistent
  embedding through point 1 and the minimum image agree).

The checks (each written without the oracle's code):

* known-answer triplets built by hand, whose bins follow from their geometry
  (values worked out in the comments, all well away from bin edges);
* the exchange of points 2 and 3 in XXX: H[a][b][c][t][p] == H[b][a][c][p][t];
* z-reflection of a non-periodic catalog: t -> bn-1-t and p -> bn-1-p;
* a numpy minimum-image brute force on integer lattices (exact distances,
  bin = isqrt(d2) with ds = 1) for periodic 3iso and 3ani, XXX and XXY;
* translation invariance mod L of periodic 3iso.

The mutation test compiles copies of oracle/pcf_oracle.cpp with one planted
mistake each into a temporary directory and shows the checks fail on it.
"""
import json
import math
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from desi_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


# ---------------------------------------------------------------------------
# Known-answer triplets
# ---------------------------------------------------------------------------
def _cat(points):
    out = np.zeros((len(points), 4), np.float32)
    out[:, :3] = np.array(points, np.float32)
    out[:, 3] = 1.0
    return out


def _deposits(c, bins):
    """{(a, b, c, t, p): count} of the non-zero 3ani bins."""
    idx = np.nonzero(c)[0]
    return {tuple(int(v) for v in np.unravel_index(i, (bins,) * 5)): int(c[i]) for i in idx}


def check_known_angles(L=None):
    # 15 bins to dmax 60: ds = 0.25 (a side of length r is in bin floor(r/4)),
    # angle scale bn/2 = 7.5.
    bins, dmax = 15, 60.0
    p1 = (100.0, 100.0, 100.0)
    p2 = (112.0, 100.0, 108.0)   # p1 + (12, 0, 8): |d| = sqrt(208) = 14.42
    p3 = (100.0, 120.0, 95.0)    # p1 + (0, 20, -5): |d| = sqrt(425) = 20.62
    # |p3 - p2| = |(-12, 20, -13)| = sqrt(713) = 26.70
    # (1, 2, 3): sides 14.42 -> 3, d31 20.62 -> 5, d23 26.70 -> 6;
    #   cos12 = 8/14.42 = 0.555 -> (1.555)(7.5) = 11.66 -> t = 11;
    #   cos31 = -5/20.62 = -0.243 -> (0.757)(7.5) = 5.68 -> p = 5.
    # (2, 1, 3): d12 14.42 -> 3, d31 = |p3 - p2| 26.70 -> 6, d23 20.62 -> 5;
    #   cos12 = -8/14.42 -> (0.445)(7.5) = 3.34 -> t = 3;
    #   cos31 = -13/26.70 = -0.487 -> (0.513)(7.5) = 3.85 -> p = 3.
    c, _ = O.triplets(_cat([p1, p2]), _cat([p3]), ani=True, bins=bins, dmax=dmax, L=L)
    assert _deposits(c, bins) == {(3, 5, 6, 11, 5): 1, (3, 6, 5, 3, 3): 1}, "tilted triplet"

    # Pair in the xy plane (cos 0 -> t = floor(7.5) = 7) and a third point at
    # about 60 degrees to z: p1 + (7, 0, 4), cos = 4/8.06 = 0.496 ->
    # (1.496)(7.5) = 11.22 -> p = 11.  Sides: |(9, 12, 0)| = 15 -> 3,
    # |(7, 0, 4)| = 8.06 -> 2, |(-2, -12, 4)| = 12.81 -> 3.
    q2 = (109.0, 112.0, 100.0)
    q3 = (107.0, 100.0, 104.0)
    c, _ = O.triplets(_cat([p1, q2]), _cat([q3]), ani=True, bins=bins, dmax=dmax, L=L)
    dep = _deposits(c, bins)
    assert dep.get((3, 2, 3, 7, 11)) == 1, "xy-plane / 60-degree triplet"
    # (2, 1, 3): cos12 = 0 -> t = 7; d31 = |q3 - q2| = 12.81 -> 3, cos31 =
    # 4/12.81 = 0.312 -> (1.312)(7.5) = 9.84 -> p = 9; d23 = 8.06 -> 2.
    assert dep == {(3, 2, 3, 7, 11): 1, (3, 3, 2, 7, 9): 1}, "xy-plane / 60-degree triplet"

    # Along the line of sight: p2 = p1 + (0, 0, 10).  From p1, cos = 1 ->
    # (2)(7.5) = 15 > 14: skipped.  From p2, cos = -1 -> t = 0: kept.
    # Third point p1 + (6, 8, 0): d31 = 10 -> 2, cos 0 -> p = 7;
    # from p2: d31 = |(6, 8, -10)| = 14.14 -> 3, cos -0.707 -> 2.2 -> p = 2;
    # d23 from p1's view = 14.14 -> 3, from p2's view = 10 -> 2.
    r2 = (100.0, 100.0, 110.0)
    r3 = (106.0, 108.0, 100.0)
    c, _ = O.triplets(_cat([p1, r2]), _cat([r3]), ani=True, bins=bins, dmax=dmax, L=L)
    assert _deposits(c, bins) == {(2, 3, 2, 0, 2): 1}, "cos = 1 exclusion"

    # Periodic, box 250 (3 * 60 < 250).  Point 1 at z = 20.
    #   s2 = (112, 100, 12): dz = -8, not projected; the reference uses
    #   |dz| = +8 -> cos +0.555 -> t = 11.
    #   s3 = (100, 120, 245): dz = 225 -> projected,
    #   |225| - 250 = -25; separation (0, 20, -25): |d| = 32.02 -> 8,
    #   cos = -25/32.02 = -0.781 -> (0.219)(7.5) = 1.64 -> p = 1.
    #   d23: s3 - s2 = (-12, 20, 233) -> projected on z: (-12, 20, -17),
    #   |d| = sqrt(144 + 400 + 289) = 28.86 -> 7.
    s1 = (100.0, 100.0, 20.0)
    s2 = (112.0, 100.0, 12.0)
    s3 = (100.0, 120.0, 245.0)
    c, _ = O.triplets(_cat([s1, s2]), _cat([s3]), ani=True, bpc=True, box=250.0, bins=bins,
                      dmax=dmax, L=L)
    dep = _deposits(c, bins)
    assert dep.get((3, 8, 7, 11, 1)) == 1, "periodic sign rule"
    # (2, 1, 3): from s2, s1 is at dz = +8 (unprojected) -> t = 11 again;
    # s3 - s2 = (-12, 20, 233): projected, dz = 233 - 250 = -17, |d| = 28.86
    # -> 7, cos = -17/28.86 = -0.589 -> (0.411)(7.5) = 3.08 -> p = 3;
    # d23 = |s3 - s1| = 32.02 -> 8.
    assert dep == {(3, 8, 7, 11, 1): 1, (3, 7, 8, 11, 3): 1}, "periodic sign rule"


# ---------------------------------------------------------------------------
# Symmetries
# ---------------------------------------------------------------------------
def check_exchange_symmetry(L=None):
    bins = 6
    for bpc in (False, True):
        a = synth.neyman_scott(10, 14, 3.0, 50.0, seed=11)
        c, w = O.triplets(a, ani=True, bpc=bpc, box=50.0, bins=bins, dmax=14.0, L=L)
        h = c.reshape((bins,) * 5)
        assert h.sum() > 1000
        # exchange of points 2 and 3: d12 <-> d31, d23 fixed, t <-> p
        assert np.array_equal(h, np.transpose(h, (1, 0, 2, 4, 3))), f"exchange bpc={bpc}"


def _dyadic(n, box, seed):
    rng = np.random.default_rng(seed)
    out = np.ones((n, 4), np.float32)
    out[:, :3] = rng.integers(0, int(box * 256), (n, 3)) / 256.0
    out[:, 3] = (0.5 + rng.integers(0, 256, n) / 256.0)
    return out


def check_z_reflection(L=None):
    # z -> 64 - z is exact on multiples of 2^-8 in [0, 40): dz -> -dz, so
    # cos -> -cos and (1 - cos) * bn/2 = bn - (1 + cos) * bn/2: t -> bn-1-t,
    # p -> bn-1-p (no triplet of these catalogs sits on an angle-bin edge or
    # on the z axis).
    bins = 15
    a = _dyadic(70, 40.0, seed=3)
    b = _dyadic(60, 40.0, seed=4)
    ra, rb = a.copy(), b.copy()
    ra[:, 2] = 64.0 - a[:, 2]
    rb[:, 2] = 64.0 - b[:, 2]
    for x, y, rx, ry in ((a, None, ra, None), (a, b, ra, rb)):
        c, _ = O.triplets(x, y, ani=True, bins=bins, dmax=16.0, L=L)
        cr, _ = O.triplets(rx, ry, ani=True, bins=bins, dmax=16.0, L=L)
        h = c.reshape((bins,) * 5)
        hr = cr.reshape((bins,) * 5)
        assert h.sum() > 1000
        assert np.array_equal(hr, h[:, :, :, ::-1, ::-1]), "z reflection"


# ---------------------------------------------------------------------------
# Minimum-image brute force on integer lattices
# ---------------------------------------------------------------------------
def _min_image(d, box):
    return d - box * np.round(d / box).astype(np.int64)


def minimum_image_triplets(a, b, *, auto, ani, bins, dmax, box):
    """Ordered-triplet histogram of integer-lattice catalogs, periodic box
    `box` (integer), dmax = bins (so ds = 1 and a side's bin is isqrt(d2)).
    Every side is the minimum-image separation.  3ani: dz of a side from point
    1 is |dz_raw| when the minimum image is the unshifted point and
    |dz_raw| - box when it is shifted; t, p = (int)((1 + dz/d) * bins/2);
    XXY with point 3 on point 1 (d31 = 0) deposits p = 0."""
    assert 3 * dmax < box and float(bins) == float(dmax)
    A = a[:, :3].astype(np.int64)
    B = A if auto else b[:, :3].astype(np.int64)
    R = dmax * dmax
    isq = np.array([math.isqrt(v) for v in range(R)], np.int64)
    h = np.zeros(bins ** 5 if ani else bins ** 3, np.uint64)
    half = bins / 2.0
    for i in range(len(A)):
        raw12 = A - A[i]
        d12 = _min_image(raw12, box)
        s12 = (d12 * d12).sum(1)
        js = np.nonzero((s12 > 0) & (s12 < R))[0]
        raw13 = B - A[i]
        d13 = _min_image(raw13, box)
        s13 = (d13 * d13).sum(1)
        ok13 = s13 < R
        if auto:
            ok13 &= s13 > 0
        ks = np.nonzero(ok13)[0]
        if len(js) == 0 or len(ks) == 0:
            continue
        d23 = _min_image(B[ks][None, :, :] - A[js][:, None, :], box)
        s23 = (d23 * d23).sum(2)
        ok = s23 < R
        if auto:
            ok &= s23 > 0
        jj, kk = np.nonzero(ok)
        a12 = isq[s12[js][jj]]
        b23 = isq[s23[jj, kk]]
        c31 = isq[s13[ks][kk]]
        if not ani:
            flat = (a12 * bins + b23) * bins + c31
        else:
            def zterm(raw):
                m = np.abs(raw[:, 2])
                return np.where(m * 2 < box, m, m - box).astype(np.float64)

            dz12 = zterm(raw12[js][jj])
            dz31 = zterm(raw13[ks][kk])
            r12 = np.sqrt(s12[js][jj].astype(np.float64))
            r31 = np.sqrt(s13[ks][kk].astype(np.float64))
            t = np.floor((1.0 + dz12 / r12) * half).astype(np.int64)
            with np.errstate(invalid="ignore", divide="ignore"):
                pf = (1.0 + dz31 / r31) * half
            p = np.where(r31 == 0, 0, np.floor(np.nan_to_num(pf))).astype(np.int64)
            keep = (t <= bins - 1) & (p <= bins - 1)
            flat = ((((a12 * bins + c31) * bins + b23) * bins + t) * bins + p)[keep]
        h += np.bincount(flat, minlength=len(h)).astype(np.uint64)
    return h


def _lattice_pair(seed):
    a = synth.lattice(150, 32.0, 1.0, seed=seed)
    b = synth.lattice(120, 32.0, 1.0, seed=seed + 1)
    b[:30, :3] = a[:30, :3]   # coincident points: XXY with d31 = 0
    return a, b


def check_min_image_lattice(L=None):
    bins, box = 10, 32
    a, b = _lattice_pair(5)
    for ani in (False, True):
        for auto in (True, False):
            c, _ = O.triplets(a, None if auto else b, ani=ani, bpc=True, box=float(box),
                              bins=bins, dmax=float(bins), L=L)
            ref = minimum_image_triplets(a, b, auto=auto, ani=ani, bins=bins, dmax=bins,
                                         box=box)
            assert ref.sum() > 2000
            assert np.array_equal(c, ref), f"minimum image ani={ani} auto={auto}"


def check_translation_3iso(L=None):
    bins = 10
    a, b = _lattice_pair(7)
    shift = np.array([13, 29, 5], np.float32)
    sa, sb = a.copy(), b.copy()
    sa[:, :3] = np.mod(a[:, :3] + shift, 32.0)
    sb[:, :3] = np.mod(b[:, :3] + shift, 32.0)
    for y, sy in ((None, None), (b, sb)):
        c1, _ = O.triplets(a, y, bpc=True, box=32.0, bins=bins, dmax=10.0, L=L)
        c2, _ = O.triplets(sa, sy, bpc=True, box=32.0, bins=bins, dmax=10.0, L=L)
        assert c1.sum() > 2000
        assert np.array_equal(c1, c2), "translation"


CHECKS = {
    "known_angles": check_known_angles,
    "exchange": check_exchange_symmetry,
    "z_reflection": check_z_reflection,
    "min_image": check_min_image_lattice,
    "translation": check_translation_3iso,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_pin(name):
    CHECKS[name]()


# ---------------------------------------------------------------------------
# Mutation test: each planted mistake must fail at least one check.
# ---------------------------------------------------------------------------
ORACLE_SRC = os.path.join(ROOT, "oracle", "pcf_oracle.cpp")

# (name, text in oracle/pcf_oracle.cpp, replacement, checks expected to fail)
MUTANTS = [
    ("t_p_swapped",
     "return (((((int64_t)a * bn + b31) * bn + b23) * bn + t) * bn) + p;",
     "return (((((int64_t)a * bn + b31) * bn + b23) * bn + p) * bn) + t;",
     {"known_angles"}),
    ("one_plus_dropped",
     "const double cth12 = 1 + t12[2] / d12;",
     "const double cth12 = t12[2] / d12;",
     {"known_angles", "z_reflection", "min_image"}),
    ("dz_sign_flipped",
     "const double cth31 = 1 + t31[2] / d31;",
     "const double cth31 = 1 - t31[2] / d31;",
     {"known_angles", "min_image"}),
    ("angle_scale_wrong",
     "float ds_th = (float)bins / 2;",
     "float ds_th = (float)(bins - 1) / 2;",
     {"known_angles", "z_reflection", "min_image"}),
    ("side23_rule_or",
     "t23[a] = proj(d23, L, g2 != g3);",
     "t23[a] = proj(d23, L, g2 | g3);",
     {"min_image"}),
    ("side23_rule_f3",
     "t23[a] = proj(d23, L, g2 != g3);",
     "t23[a] = proj(d23, L, g3);",
     {"min_image"}),
    ("projection_unsigned",
     "return std::fabs(d) - L * (double)f; }",
     "return d - L * (double)f; }",
     {"known_angles", "min_image"}),
    ("ani_sides_swapped",
     "return (((((int64_t)a * bn + b31) * bn + b23) * bn + t) * bn) + p;",
     "return (((((int64_t)a * bn + b23) * bn + b31) * bn + t) * bn) + p;",
     {"known_angles", "exchange", "min_image"}),
    ("xxy_excludes_d31_zero",
     "if (!(s23 < r.R && s31 < r.R)) return -1;",
     "if (!(s23 < r.R && s31 < r.R && s31 > 0)) return -1;",
     {"min_image"}),
]

_RUNNER = r"""
import json, sys
sys.path.insert(0, {tests!r}); sys.path.insert(0, {root!r})
import oracle_lib as O
import test_oracle_triplet_pins as P
L = O.load({lib!r})
failed = []
for name, fn in P.CHECKS.items():
    try:
        fn(L)
    except AssertionError:
        failed.append(name)
print(json.dumps(failed))
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("mutant", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(mutant, tmp_path):
    name, old, new, expect = mutant
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"mutation site of {name} not found exactly once"
    mut = tmp_path / "pcf_oracle_mutant.cpp"
    mut.write_text(src.replace(old, new))
    so = tmp_path / "liboracle_mutant.so"
    subprocess.run(["g++", "-O1", "-fopenmp", "-ffp-contract=off", "-fPIC", "-std=c++17",
                    "-shared", "-w", "-o", str(so), str(mut)], check=True)
    code = _RUNNER.format(tests=HERE, root=ROOT, lib=str(so))
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=600, env=dict(os.environ, OMP_NUM_THREADS="4"))
    if res.returncode != 0:
        pytest.fail(f"mutant {name} crashed the checks: {res.stderr[-2000:]}")
    failed = set(json.loads(res.stdout.strip().splitlines()[-1]))
    assert expect <= failed, f"{name}: expected {sorted(expect)} to fail, got {sorted(failed)}"
