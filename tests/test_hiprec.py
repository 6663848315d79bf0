"""The extended-precision power of tests/hiprec.py, and the conditioning of
the reference's FP64 formula (CPU).

hiprec is what FP64 results are measured against where two correct FP64
evaluations cannot agree to 1e-12.  It is pinned here by routes that do not
share its code:

  * the C oracle (bit-equal to the reference's CPU build) on the bins where
    the FP64 formula is well conditioned, to 1e-12;
  * the least-squares definition of the power (the standard power is half
    the drop in sum y^2 from fitting a cos + b sin; the generalized power is
    the fraction of the weighted variance removed by fitting c + a cos +
    b sin), solved by numpy's lstsq, to 1e-9 (lstsq's own accuracy);
  * an mpmath evaluation at 40 significant digits on the worst-conditioned
    bins of a 6-observation object, to 1e-14;
  * invariance of the power under a shift of all times, to 1e-14.

The second half shows why parity of FP64 results is stated against hiprec
on the ill-conditioned bins: the reference's generalized tau-frame formula
(main.cu:1701-1777), evaluated in FP64 with tau moved by one ulp (as a
different libm's atan2 may return), changes weak bins of the 2000-day
generalized batch by far more than 1e-12, while hiprec shows the
unperturbed and perturbed results to be equally accurate.
"""
import mpmath
import numpy as np
import pytest

import hiprec
import oracle
from gpu_lomb_scargle_amd import synth

F64 = np.float64


def _objects(data, k0, k1):
    off = np.concatenate([[0], np.cumsum(data["counts"])])
    out = []
    for k in range(k0, k1):
        sl = slice(off[k], off[k + 1])
        out.append((data["t"][sl], data["y"][sl], None if data["dy"] is None else data["dy"][sl]))
    return out


def _oracle_rows(objs, gen, nf, normalize=True):
    rows = []
    for t, y, dy in objs:
        rows.append(oracle.single(t, y.copy(), dy, synth.OMEGA_MIN, synth.OMEGA_MAX, nf, gen, F64,
                                  normalize=normalize, nthreads=8)[0])
    return np.array(rows)


@pytest.mark.parametrize("gen", [False, True], ids=["standard", "generalized"])
def test_matches_oracle_where_well_conditioned(gen):
    data = synth.generalized_batch_1000() if gen else synth.standard_batch_1000()
    objs = _objects(data, 0, 4)
    if gen:
        # the tau frame's second pass rounds omega (t - tau), whose FP64
        # error (ulp(t) omega, 3e-11 rad at 2000 d) moves every bin by
        # ~1e-11: an 8-day base (t / 256, exact) leaves 1e-12 to the formula
        objs = [(t / 256.0, y, dy) for t, y, dy in objs]
    nf = 20000
    ref = _oracle_rows(objs, gen, nf)
    for k, (t, y, dy) in enumerate(objs):
        bins = np.flatnonzero(ref[k] >= 0.05 * ref[k].max())
        om = hiprec.grid(synth.OMEGA_MIN, synth.OMEGA_MAX, nf, bins)
        hp = hiprec.powers(gen, t, y, dy, om, fp64_phase=True)
        rel = np.abs(hp - ref[k, bins]) / ref[k, bins]
        exact = hiprec.powers(gen, t, y, dy, om)
        print(f"object {k}: oracle vs hiprec (FP64 phases) {rel.max():.2e}, "
              f"(exact phases) {(np.abs(exact - ref[k, bins]) / ref[k, bins]).max():.2e}")
        assert bins.size > 5 and rel.max() <= 1e-12, (k, rel.max())
    # the un-normalised standard power too
    if not gen:
        t, y, _ = objs[0]
        raw = oracle.single(t, y.copy(), None, synth.OMEGA_MIN, synth.OMEGA_MAX, nf, False, F64,
                            normalize=False)[0]
        bins = np.flatnonzero(raw >= 0.05 * raw.max())
        hp = hiprec.powers(False, t, y, None, hiprec.grid(synth.OMEGA_MIN, synth.OMEGA_MAX, nf, bins),
                           normalize=False, fp64_phase=True)
        assert (np.abs(hp - raw[bins]) / raw[bins]).max() <= 1e-12


def _lstsq_power(gen, t, y, dy, om):
    c, s = np.cos(om * t), np.sin(om * t)
    if not gen:
        X = np.stack([c, s], 1)
        coef = np.linalg.lstsq(X, y, rcond=None)[0]
        r = y - X @ coef
        return 0.5 * (y @ y - r @ r)
    sw = 1.0 / dy
    X = np.stack([np.ones_like(t), c, s], 1) * sw[:, None]
    b = y * sw
    coef = np.linalg.lstsq(X, b, rcond=None)[0]
    r = b - X @ coef
    w = sw * sw
    ybar = (w * y).sum() / w.sum()
    chi0 = (w * (y - ybar) ** 2).sum()
    return 1.0 - (r @ r) / chi0


@pytest.mark.parametrize("gen", [False, True], ids=["standard", "generalized"])
def test_matches_least_squares_definition(gen):
    rng = np.random.default_rng(3)
    n = 80
    t = np.sort(rng.uniform(0, 300, n))
    y = np.sin(2 * np.pi * t / 1.7) + rng.normal(0, 0.5, n) + (16.0 if gen else 0.0)
    dy = rng.uniform(0.05, 0.3, n) if gen else None
    om = np.linspace(0.5, 40.0, 97)
    hp = hiprec.powers(gen, t, y, dy, om, normalize=False)
    ls = np.array([_lstsq_power(gen, t, y, dy, o) for o in om])
    assert np.abs(hp - ls).max() <= 1e-9 * np.abs(ls).max()
    assert (np.abs(hp - ls) / np.abs(ls)).max() <= 1e-9


def _mp_power(gen, t, y, dy, om):
    mpmath.mp.dps = 40
    om = mpmath.mpf(float(om))
    ph = [om * mpmath.mpf(float(x)) for x in t]  # exact: 106 bits < 133
    c = [mpmath.cos(p) for p in ph]
    s = [mpmath.sin(p) for p in ph]
    y = [mpmath.mpf(float(v)) for v in y]
    if not gen:
        xc = mpmath.fsum(a * b for a, b in zip(y, c))
        xs = mpmath.fsum(a * b for a, b in zip(y, s))
        cc = mpmath.fsum(a * a for a in c)
        ss = mpmath.fsum(a * a for a in s)
        cs = mpmath.fsum(a * b for a, b in zip(c, s))
        return (ss * xc ** 2 + cc * xs ** 2 - 2 * cs * xc * xs) / (2 * (cc * ss - cs ** 2))
    w = [1 / mpmath.mpf(float(d)) ** 2 for d in dy]
    W = mpmath.fsum(w)
    w = [v / W for v in w]
    Y = mpmath.fsum(a * b for a, b in zip(w, y))
    yc = [v - Y for v in y]
    C = mpmath.fsum(a * b for a, b in zip(w, c))
    S = mpmath.fsum(a * b for a, b in zip(w, s))
    cc = mpmath.fsum(a * b * b for a, b in zip(w, c)) - C * C
    ss = mpmath.fsum(a * b * b for a, b in zip(w, s)) - S * S
    cs = mpmath.fsum(a * b * d for a, b, d in zip(w, c, s)) - C * S
    yc_ = mpmath.fsum(a * b * d for a, b, d in zip(w, yc, c))
    ys_ = mpmath.fsum(a * b * d for a, b, d in zip(w, yc, s))
    yy = mpmath.fsum(a * b * b for a, b in zip(w, yc))
    return (ss * yc_ ** 2 + cc * ys_ ** 2 - 2 * cs * yc_ * ys_) / ((cc * ss - cs ** 2) * yy)


@pytest.mark.parametrize("gen", [False, True], ids=["standard", "generalized"])
def test_matches_mpmath_on_ill_conditioned_bins(gen):
    """A 6-observation object over 1800 d: its 2x2 systems are near-singular
    at many frequencies, where FP64 evaluations scatter."""
    rng = np.random.default_rng(11 + gen)
    n = 6
    t = np.sort(rng.integers(0, 1801, n) + rng.uniform(0, 0.3, n))
    y = 17.0 + 0.3 * np.sin(2 * np.pi * t / 0.77) + rng.normal(0, 0.05, n)
    if not gen:
        y = y - 17.0
    dy = np.exp(rng.uniform(np.log(0.01), np.log(0.2), n)) if gen else None
    nf = 30011
    ref = _oracle_rows([(t, y, dy)], gen, nf, normalize=False)[0]
    allb = np.arange(nf)
    hp_all = hiprec.powers(gen, t, y, dy, hiprec.grid(synth.OMEGA_MIN, synth.OMEGA_MAX, nf, allb),
                           normalize=False)
    m = ref >= 1e-3 * ref.max()
    rel = np.where(m, np.abs(ref - hp_all) / hp_all, 0.0)
    worst = np.argsort(rel)[::-1][:12]
    print(f"oracle vs hiprec on P >= 1e-3 Pmax: max rel {rel.max():.2e}")
    om = hiprec.grid(synth.OMEGA_MIN, synth.OMEGA_MAX, nf, worst)
    eh, eo = [], []
    for b, o in zip(worst, om):
        exact = float(_mp_power(gen, t, y, dy, o))
        eh.append(abs(hp_all[b] - exact) / exact)
        eo.append(abs(ref[b] - exact) / exact)
    eh, eo = np.array(eh), np.array(eo)
    print(f"on the 12 worst bins: hiprec vs mpmath {eh.max():.2e}, oracle vs mpmath {eo.min():.2e}..{eo.max():.2e}")
    assert eh.max() <= 1e-13
    assert eo.max() > 1e-12 and (eh <= 1e-2 * eo).all()


@pytest.mark.parametrize("gen", [False, True], ids=["standard", "generalized"])
def test_invariant_under_time_shift(gen):
    data = synth.generalized_batch_1000() if gen else synth.standard_batch_1000()
    (t, y, dy), = _objects(data, 7, 8)
    t = np.round(t * 2.0 ** 30) / 2.0 ** 30  # on a 2^-30 grid: t + 1000.25 is exact
    om = hiprec.grid(synth.OMEGA_MIN, synth.OMEGA_MAX, 4096, np.arange(4096))
    a = hiprec.powers(gen, t, y, dy, om)
    b = hiprec.powers(gen, t + 1000.25, y, dy, om)
    m = a >= 1e-3 * a.max()
    assert (np.abs(a - b)[m] / a[m]).max() <= 1e-14


def _tau_form_gen(om, t, y, dy, nudge=0):
    """The reference's generalized tau-frame power in FP64 (numpy; the
    order of main.cu:1701-1777), with tau moved by `nudge` ulps."""
    w = 1.0 / dy
    w = w * w
    wsum = w.sum()
    ph = om[:, None] * t[None, :]
    so, co = np.sin(ph), np.cos(ph)
    S = (w * so).sum(1) / wsum
    C = (w * co).sum(1) / wsum
    S2 = (2.0 * w * so * co).sum(1) / wsum - 2.0 * S * C
    C2 = (w - 2.0 * w * so * so).sum(1) / wsum - (C * C - S * S)
    tau = 0.5 * np.arctan2(S2, C2) / om
    for _ in range(abs(nudge)):
        tau = np.nextafter(tau, np.inf if nudge > 0 else -np.inf)
    ph = om[:, None] * (t[None, :] - tau[:, None])
    so, co = np.sin(ph), np.cos(ph)
    Y = (w * y).sum() / wsum
    YY = (w * y * y).sum() / wsum - Y * Y
    Ct = (w * co).sum(1) / wsum
    St = (w * so).sum(1) / wsum
    YC = (w * y * co).sum(1) / wsum - Y * Ct
    YS = (w * y * so).sum(1) / wsum - Y * St
    CC = (w * co * co).sum(1) / wsum - Ct * Ct
    SS = (w * so * so).sum(1) / wsum - St * St
    return (YC * YC / CC + YS * YS / SS) / YY


def test_reference_formula_moves_with_one_ulp_of_tau():
    """Weak bins (P >= 1e-3 P_max) of generalized-batch objects (2000 d,
    per-observation dmag): one ulp of tau moves the FP64 tau-frame power by
    more than 5e-12 relative somewhere, and both results stay equally close
    to hiprec -- neither is the more correct one."""
    data = synth.generalized_batch_1000(dmag_per_obs=True)
    objs = _objects(data, 0, 3)
    nf = 40000
    om_all = hiprec.grid(synth.OMEGA_MIN, synth.OMEGA_MAX, nf, np.arange(nf))
    worst_shift, errs = 0.0, {0: [], 1: []}
    for t, y, dy in objs:
        yc = oracle.precenter(y, dy)  # the reference sweeps pre-centred y
        p0 = _tau_form_gen(om_all, t, yc, dy)
        p1 = _tau_form_gen(om_all, t, yc, dy, nudge=1)
        m = p0 >= 1e-3 * p0.max()
        rel = np.abs(p1 - p0)[m] / p0[m]
        worst_shift = max(worst_shift, rel.max())
        sel = np.flatnonzero(m)[np.argsort(rel)[::-1][:64]]
        hp = hiprec.powers(True, t, y, dy, om_all[sel])
        errs[0].append(np.abs(p0[sel] - hp) / hp)
        errs[1].append(np.abs(p1[sel] - hp) / hp)
    e0, e1 = np.concatenate(errs[0]), np.concatenate(errs[1])
    print(f"one ulp of tau: max rel shift {worst_shift:.2e}; error vs hiprec there: "
          f"{e0.max():.2e} (tau), {e1.max():.2e} (tau + 1 ulp)")
    assert worst_shift > 5e-12
    assert e1.max() <= 4 * e0.max() and e0.max() <= 4 * e1.max()
