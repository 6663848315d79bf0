"""Seeded synthetic light curves for the benchmark and the tests.

Observation model (shared): integer night indices uniform over 0..span days
plus a uniform within-night offset in [0, 0.3] d, sorted; a sinusoid of period
P log-uniform in [0.05, 10] d and uniform phase.  Draw order per object:
count, nights, offsets, P, phase, [mean magnitude, dmag], noise.

Workloads (angular frequency grid 1.005 .. 150.796 rad/day, max excluded):
  * single_object_3555      1 object, N = 3555, y normalised, 1e5 bins
  * lognormal_batch         K objects, N ~ lognormal(median, sigma) clipped
    - standard, 1000 x median 140 / sigma 0.6 / clip 20..1000, 1e6 bins
    - generalized ("with_error"): span 2000 d, y = m + A sin(...) + N(0, dmag),
      m ~ U(15, 20) not subtracted, dmag ~ logU(0.01, 0.2) per object, A = 0.3;
      with dmag_per_obs each observation draws its own dmag ~ logU(0.01, 0.2)
      (weights 1/dmag^2 spread over a factor of 400 within an object)
    - uneven, 100000 x median 60 / sigma 1.0 / clip 5..5000, 2e5 bins
    - refinement, 4096 x median 200 / sigma 0.5 / clip >= 5, 5e5 bins
"""
from __future__ import annotations

import numpy as np

OMEGA_MIN = 1.005
OMEGA_MAX = 150.796
AMPLITUDE_WITH_ERROR = 0.3


def _one(rng, n, span, with_error, dmag_per_obs=False):
    nights = rng.integers(0, span + 1, n)
    t = np.sort(nights + rng.uniform(0.0, 0.3, n))
    P = float(np.exp(rng.uniform(np.log(0.05), np.log(10.0))))
    phi = float(rng.uniform(0.0, 2.0 * np.pi))
    if with_error:
        m = float(rng.uniform(15.0, 20.0))
        if dmag_per_obs:
            dy = np.exp(rng.uniform(np.log(0.01), np.log(0.2), n))
        else:
            dy = np.full(n, float(np.exp(rng.uniform(np.log(0.01), np.log(0.2)))))
        y = m + AMPLITUDE_WITH_ERROR * np.sin(2.0 * np.pi * t / P + phi) + rng.normal(0.0, dy, n)
    else:
        y = np.sin(2.0 * np.pi * t / P + phi) + rng.normal(0.0, 0.3, n)
        y = (y - y.mean()) / y.std()
        dy = None
    return t, y, dy, P


def single_object_3555(seed=1000):
    """One object, 3555 observations over 1800 nights, normalised y."""
    rng = np.random.default_rng(seed)
    t, y, _, P = _one(rng, 3555, 1800, False)
    return {"obj_id": np.zeros(3555, np.uint32), "t": t, "y": y, "dy": None, "period": np.array([P])}


def lognormal_counts(rng, K, median, sigma, lo=None, hi=None):
    z = rng.standard_normal(K)
    n = np.rint(np.exp(np.log(median) + sigma * z))
    if lo is not None or hi is not None:
        n = np.clip(n, lo if lo is not None else -np.inf, hi if hi is not None else np.inf)
    return n.astype(np.int64)


def lognormal_batch(K=1000, median=140, sigma=0.6, lo=20, hi=1000, span=1800, with_error=False,
                    seed=1001, dmag_per_obs=False):
    """K objects with log-normal observation counts, rows object-contiguous."""
    rng = np.random.default_rng(seed)
    counts = lognormal_counts(rng, K, median, sigma, lo, hi)
    ts, ys, dys, Ps = [], [], [], []
    for k in range(K):
        t, y, dy, P = _one(rng, int(counts[k]), span, with_error, dmag_per_obs)
        ts.append(t)
        ys.append(y)
        dys.append(dy)
        Ps.append(P)
    obj = np.repeat(np.arange(K, dtype=np.uint32), counts)
    return {
        "obj_id": obj,
        "t": np.concatenate(ts),
        "y": np.concatenate(ys),
        "dy": np.concatenate(dys) if with_error else None,
        "counts": counts,
        "period": np.array(Ps),
    }


def standard_batch_1000(seed=1001):
    """1000 objects, median 140 obs (sigma 0.6, clip 20..1000), 1800 d."""
    return lognormal_batch(1000, 140, 0.6, 20, 1000, 1800, False, seed)


def generalized_batch_1000(seed=1002, dmag_per_obs=False):
    """The 1000-object layout over 2000 d, un-centred magnitudes with dmag
    (per object, or per observation with dmag_per_obs)."""
    return lognormal_batch(1000, 140, 0.6, 20, 1000, 2000, True, seed, dmag_per_obs)


def uneven_batch_100000(seed=1003, with_error=False):
    """100000 objects, median 60 obs (sigma 1.0, clip 5..5000)."""
    return lognormal_batch(100000, 60, 1.0, 5, 5000, 1800, with_error, seed)


def refinement_batch_4096(seed=1004, with_error=True):
    """4096 objects, median 200 obs (sigma 0.5, >= 5), dmag as the
    generalized batch."""
    return lognormal_batch(4096, 200, 0.5, 5, None, 2000, with_error, seed)
