"""Seeded synthetic catalogs of the benchmark configurations.

uniform(n, box, seed): x, y, z i.i.d. uniform FP32 in [0, box), w = 1.
neyman_scott(nparents, nchild, sigma, box, seed): parent centres uniform in the
box, each with `nchild` children offset by an isotropic Gaussian of width sigma,
wrapped periodically into [0, box).
Weights: w = 1, or uniform in [0.5, 1.5] FP32 with weights="uniform".
"""
import numpy as np


def _wrap(v, box):
    v = np.mod(v, box).astype(np.float32)
    # float rounding of the modulo can land exactly on the box side
    top = np.nextafter(np.float32(box), np.float32(0))
    return np.minimum(v, top)


def _weights(n, weights, rng):
    if weights is None or weights == "one":
        return np.ones(n, np.float32)
    if weights == "uniform":
        return (0.5 + rng.random(n)).astype(np.float32)
    raise ValueError(weights)


def uniform(n, box, seed, weights=None):
    rng = np.random.default_rng(seed)
    xyz = _wrap(rng.random((n, 3)) * box, box)
    out = np.empty((n, 4), np.float32)
    out[:, :3] = xyz
    out[:, 3] = _weights(n, weights, rng)
    return out


def neyman_scott(nparents, nchild, sigma, box, seed, weights=None):
    rng = np.random.default_rng(seed)
    parents = rng.random((nparents, 3)) * box
    kids = parents[:, None, :] + rng.normal(0.0, sigma, (nparents, nchild, 3))
    xyz = _wrap(kids.reshape(-1, 3), box)
    n = len(xyz)
    out = np.empty((n, 4), np.float32)
    out[:, :3] = xyz
    out[:, 3] = _weights(n, weights, rng)
    return out


def lattice(n, box, step, seed, weights=None):
    """Points on a dyadic lattice (multiples of `step`, a power of two), so FP32
    and FP64 squared distances are exact integers times step^2: exercises
    d2 == edge and d == 0 deterministically."""
    rng = np.random.default_rng(seed)
    m = int(box / step)
    xyz = (rng.integers(0, m, (n, 3)) * step).astype(np.float32)
    out = np.empty((n, 4), np.float32)
    out[:, :3] = xyz
    out[:, 3] = _weights(n, weights, rng)
    return out
