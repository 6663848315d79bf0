"""High-precision Lomb-Scargle powers (test infrastructure).

The power of the input data, computed far beyond FP64 so that the FP64
results of the reference's CPU path (the oracle) and of the GPU can each be
measured against it.  Where the reference's own formula is ill-conditioned
(weak bins of the generalized variant, near-singular 2x2 systems of objects
with a handful of observations), two correct FP64 evaluations that differ
only in the last bit of a sin, cos or atan2 disagree by far more than 1e-12;
there the question is which of them is closer to the power itself.

Both variants are evaluated in the tau-free closed form (algebraically equal
to the reference's tau-frame sums, source/main.cu:1566-1599 standard and
1701-1777 generalized; the tau rotation only diagonalises the 2x2 system):

  standard      P = (SS XC^2 + CC XS^2 - 2 CS XC XS) / (2 (CC SS - CS^2)),
                times 2 / (N var) when normalised (main.cu:1295-1339)
  generalized   with w = dy^-2 / sum dy^-2, y_c = y - sum w y and the centred
                C = sum w cos, S = sum w sin, CC = sum w cos^2 - C^2,
                SS = sum w sin^2 - S^2, CS = sum w cos sin - C S,
                YC = sum w y_c cos, YS = sum w y_c sin, YY = sum w y_c^2:
                P = (SS YC^2 + CC YS^2 - 2 CS YC YS) / ((CC SS - CS^2) YY)

in x87 extended precision (64-bit significand).  The phase omega t is formed
exactly: both FP64 factors are split into 26/27-bit halves whose four
products are exact in extended precision, and cos(A + B) is expanded, so the
only phase error is that of cosl/sinl themselves.  tests/test_hiprec.py
checks this module against the C oracle on well-conditioned bins, against
the least-squares definition of the power, against an mpmath evaluation at
40 digits on ill-conditioned bins, and for invariance under a time shift.
"""
from __future__ import annotations

import numpy as np

LD = np.longdouble
_SPLIT = 134217729.0  # 2^27 + 1


def _halves(x):
    """x = hi + lo with hi of 26 and lo of at most 27 significant bits."""
    x = np.asarray(x, np.float64)
    c = _SPLIT * x
    hi = c - (c - x)
    return hi, x - hi


def trig(omegas, t, fp64_phase=False):
    """cos and sin of omega t, (len(omegas), len(t)) in extended precision;
    with fp64_phase, of the FP64-rounded product (the reference's phase)."""
    if fp64_phase:
        ph = (np.asarray(omegas, np.float64)[:, None] * np.asarray(t, np.float64)[None, :]).astype(LD)
        return np.cos(ph), np.sin(ph)
    oh, ol = _halves(omegas)
    th, tl = _halves(t)
    oh, ol = oh.astype(LD)[:, None], ol.astype(LD)[:, None]
    th, tl = th.astype(LD)[None, :], tl.astype(LD)[None, :]
    a = oh * th                          # exact (52 bits)
    b = oh * tl + ol * th + ol * tl      # each product exact; |b| ~ 2^-26 |a|
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    return ca * cb - sa * sb, sa * cb + ca * sb


def _chunks(n, step):
    for i in range(0, n, step):
        yield slice(i, min(n, i + step))


def std_powers(t, y, omegas, normalize=True, block=256, fp64_phase=False):
    t = np.asarray(t, np.float64)
    yl = np.asarray(y, np.float64).astype(LD)
    omegas = np.asarray(omegas, np.float64)
    out = np.empty(omegas.size, LD)
    for sl in _chunks(omegas.size, block):
        c, s = trig(omegas[sl], t, fp64_phase)
        xc, xs = (yl * c).sum(1), (yl * s).sum(1)
        cc, ss, cs = (c * c).sum(1), (s * s).sum(1), (c * s).sum(1)
        out[sl] = (ss * xc * xc + cc * xs * xs - 2 * cs * xc * xs) / (2 * (cc * ss - cs * cs))
    if normalize:
        n = LD(yl.size)
        yc = yl - yl.sum() / n
        out *= 2 / (n * ((yc * yc).sum() / n))
    return out


def gen_powers(t, y, dy, omegas, block=256, fp64_phase=False):
    t = np.asarray(t, np.float64)
    w = 1 / np.asarray(dy, np.float64).astype(LD) ** 2
    w /= w.sum()
    yl = np.asarray(y, np.float64).astype(LD)
    yc = yl - (w * yl).sum()
    yy = (w * yc * yc).sum()
    omegas = np.asarray(omegas, np.float64)
    out = np.empty(omegas.size, LD)
    for sl in _chunks(omegas.size, block):
        c, s = trig(omegas[sl], t, fp64_phase)
        C, S = (w * c).sum(1), (w * s).sum(1)
        cc = (w * c * c).sum(1) - C * C
        ss = (w * s * s).sum(1) - S * S
        cs = (w * c * s).sum(1) - C * S
        ycs, yss = (w * yc * c).sum(1), (w * yc * s).sum(1)
        out[sl] = (ss * ycs * ycs + cc * yss * yss - 2 * cs * ycs * yss) / ((cc * ss - cs * cs) * yy)
    return out


def powers(generalized, t, y, dy, omegas, normalize=True, fp64_phase=False):
    """Extended-precision powers of one object at the given angular
    frequencies (rounded to FP64 on return).  fp64_phase evaluates the
    power at the FP64-rounded phases fl(omega t) instead of the exact ones:
    each such phase is off by up to 2^-53 |omega t| (3e-11 rad at
    omega t = 3e5), which moves the FP64 powers of long time bases by
    ~1e-11 relative whatever the rest of the arithmetic."""
    if generalized:
        return gen_powers(t, y, dy, omegas, fp64_phase=fp64_phase).astype(np.float64)
    return std_powers(t, y, omegas, normalize, fp64_phase=fp64_phase).astype(np.float64)


def grid(fmin, fmax, nf, bins):
    """The reference's FP64 angular frequencies of the given bins
    (main.cu:1136-1152: fmin + i * ((fmax - fmin) / nf))."""
    return fmin + np.asarray(bins, np.float64) * ((fmax - fmin) / nf)


def pick_bins(got, ref, mask_frac, worst=512, rand=512, seed=0):
    """(object, bin) pairs to check in high precision: the `worst` bins of
    largest relative disagreement between two FP64 periodograms among those
    with ref >= mask_frac * the object's maximum, and `rand` random ones."""
    got = np.atleast_2d(np.asarray(got, np.float64))
    ref = np.atleast_2d(np.asarray(ref, np.float64))
    pm = ref.max(axis=1, keepdims=True)
    m = ref >= mask_frac * pm
    rel = np.where(m, np.abs(got - ref) / np.where(m, ref, 1.0), -1.0).ravel()
    order = np.argsort(rel)[::-1]
    pick = list(order[:worst][rel[order[:worst]] >= 0])
    cand = np.flatnonzero(rel >= 0)
    rng = np.random.default_rng(seed)
    if cand.size:
        pick += list(rng.choice(cand, min(rand, cand.size), replace=False))
    pick = np.unique(np.asarray(pick, np.int64))
    return pick // ref.shape[1], pick % ref.shape[1]


def compare(got, ref, objects, generalized, fmin, fmax, mask_frac=1e-3, worst=512, rand=512, seed=0,
            normalize=True):
    """Errors of two FP64 periodograms (GPU `got`, oracle `ref`, rows of
    objects) against the extended-precision power on picked bins.

    objects: list of (t, y, dy) of the rows.  Returns a dict with the max
    and median relative errors of each and the number of bins checked."""
    ko, kb = pick_bins(got, ref, mask_frac, worst, rand, seed)
    truth = np.empty(ko.size)
    for k in np.unique(ko):
        sel = ko == k
        t, y, dy = objects[k]
        truth[sel] = powers(generalized, t, y, dy, grid(fmin, fmax, np.atleast_2d(ref).shape[1], kb[sel]),
                            normalize)
    g = np.atleast_2d(got)[ko, kb].astype(np.float64)
    r = np.atleast_2d(ref)[ko, kb].astype(np.float64)
    eg, er = np.abs(g - truth) / truth, np.abs(r - truth) / truth
    return {"bins": int(ko.size), "gpu_max": float(eg.max()), "gpu_median": float(np.median(eg)),
            "ref_max": float(er.max()), "ref_median": float(np.median(er)),
            "gpu_vs_ref_max": float((np.abs(g - r) / r).max())}
