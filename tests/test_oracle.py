"""Pinning the CPU oracle (oracle/ls_oracle.c).

The oracle is checked against:
  * the reference's own CPU functions compiled from its sources (oracle/_ref),
    bit for bit, when that build exists (this container);
  * golden fixtures produced by that build (tests/golden/), bit for bit;
  * the reference's recorded known answers (README run transcripts and the
    recorded output of source/lombscargle_test_example.py);
  * closed forms (least-squares projections) by brute force, invariances,
    special cases, and scipy's periodogram.
"""
import hashlib

import numpy as np
import pytest

import oracle
from conftest import load_data

TWO_PI = 2.0 * np.pi
F64, F32 = np.float64, np.float32


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


needs_ref = pytest.mark.skipif(not oracle.ref_available(False, F64),
                               reason="oracle/_ref not built (reference tree absent)")


# --------------------------------------------------------------------------
# against the reference's own compiled CPU path
# --------------------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("dtype", [F64, F32])
@pytest.mark.parametrize("error", [False, True])
@pytest.mark.parametrize("normalize", [True, False])
def test_batch_bitwise_vs_reference_build(dtype, error, normalize):
    if error and not normalize:
        pytest.skip("the generalized variant has no normalisation switch")
    o, t, y, dy = load_data("SDSS_stripe82_band_z.txt")
    sl = slice(0, 2500)
    a = oracle.batch(o[sl], t[sl], y[sl], dy[sl] if error else None, TWO_PI * 0.1, TWO_PI * 10.0,
                     1500, error, dtype, normalize=normalize, nthreads=4)
    b = oracle.ref_batch(o[sl], t[sl], y[sl], dy[sl] if error else None, TWO_PI * 0.1,
                         TWO_PI * 10.0, 1500, error, dtype, normalize=normalize)
    assert bits_equal(a[0], b[0])
    assert bits_equal(a[1], b[1]) and bits_equal(a[2], b[2])


@needs_ref
@pytest.mark.parametrize("dtype", [F64, F32])
@pytest.mark.parametrize("error", [False, True])
def test_single_bitwise_vs_reference_build(dtype, error):
    o, t, y, dy = load_data("8205_normalized_with_error.txt")
    a = oracle.single(t, y, dy if error else None, 3.14159, 150.79645, 1200, error, dtype,
                      nthreads=4)
    b = oracle.ref_single(t, y, dy if error else None, 3.14159, 150.79645, 1200, error, dtype)
    assert bits_equal(a[0], b[0])
    assert a[1] == b[1] and a[2] == b[2]


@needs_ref
@pytest.mark.parametrize("dtype", [F64, F32])
def test_helpers_bitwise_vs_reference_build(dtype):
    o, t, y, dy = load_data("SDSS_stripe82_band_z.txt")
    assert bits_equal(oracle.precenter(y[:300], dy[:300], dtype), oracle.ref_precenter(y[:300], dy[:300], dtype))
    assert oracle.stddev_sq(y[:300], dtype) == oracle.ref_stddev_sq(y[:300], dtype)
    ids = np.array([5, 5, 3, 3, 3, 5, 9, 9, 9, 9, 1], np.uint32)
    for got, ref in zip(oracle.object_ranges(ids), oracle.ref_object_ranges(ids)):
        assert np.array_equal(got, ref)
    for freq in (1.0, 7.3, 50.0):
        for err in (False, True):
            yy = oracle.precenter(y[:50], dy[:50], dtype) if err else y[:50]
            p = (oracle.gen_power(freq, t[:50], yy, dy[:50], dtype) if err
                 else oracle.std_power(freq, t[:50], yy, dtype))
            assert bits_equal(np.array(p, dtype),
                              np.array(oracle.ref_power(freq, t[:50], yy, dy[:50] if err else None,
                                                        err, dtype), dtype))


# --------------------------------------------------------------------------
# golden fixtures (generated from the reference build)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name,dtype", [("double", F64), ("float", F32)])
def test_golden_single_objects(golden, name, dtype):
    o, t, y, dy = load_data("8205_normalized_with_error.txt")
    fmin, fmax = TWO_PI * 0.1, TWO_PI * 10.0
    pg, per, _, _ = oracle.single(t, y, None, fmin, fmax, 10000, False, dtype)
    assert bits_equal(pg, golden[f"s8205_std_{name}"])
    assert per == golden[f"s8205_std_{name}_period"][0]
    pg, per, _, _ = oracle.single(t, y, dy, fmin, fmax, 10000, True, dtype)
    assert bits_equal(pg, golden[f"s8205_gen_{name}"])
    o, t, y, _ = load_data("243_normalized_ztf_filter2.txt")
    pg, _, _, _ = oracle.single(t, y, None, 3.142, 150.796, 10000, False, dtype)
    assert bits_equal(pg, golden[f"z243_std_{name}"])


def test_golden_unnormalized(golden):
    o, t, y, _ = load_data("8205_normalized.txt")
    pg, _, _, _ = oracle.single(t, y, None, TWO_PI * 0.1, TWO_PI * 10.0, 10000, False, F64,
                                normalize=False)
    assert bits_equal(pg, golden["s8205_std_double_raw"])


@pytest.mark.parametrize("name,dtype", [("double", F64), ("float", F32)])
@pytest.mark.parametrize("error", [False, True])
def test_golden_sdss_batch(golden, name, dtype, error):
    o, t, y, dy = load_data("SDSS_stripe82_band_z.txt")
    pg, per, pw, ix = oracle.batch(o, t, y, dy if error else None, TWO_PI * 0.1, TWO_PI * 10.0,
                                   20000, error, dtype, nthreads=8)
    key = f"sdss_{'gen' if error else 'std'}_{name}"
    assert hashlib.sha256(pg.tobytes()).hexdigest() == str(golden[key + "_sha256"])
    assert np.array_equal(ix, golden[key + "_argmax"])
    assert bits_equal(per, golden[key + "_period"])
    assert bits_equal(pw, golden[key + "_power"])


@pytest.mark.parametrize("name,dtype", [("double", F64), ("float", F32)])
@pytest.mark.parametrize("tag", ["std", "gen"])
def test_golden_synthetic_batch(golden, name, dtype, tag):
    err = tag == "gen"
    nf = int(golden["syn_nf"])
    pg, per, pw, ix = oracle.batch(golden[f"syn_{tag}_obj"], golden[f"syn_{tag}_t"],
                                   golden[f"syn_{tag}_y"], golden.get(f"syn_{tag}_dy"), 1.005,
                                   150.796, nf, err, dtype, nthreads=8)
    key = f"syn_{tag}_{name}"
    assert hashlib.sha256(pg.tobytes()).hexdigest() == str(golden[key + "_sha256"])
    assert np.array_equal(ix, golden[key + "_argmax"])


# --------------------------------------------------------------------------
# recorded known answers of the reference
# --------------------------------------------------------------------------
def test_kat_ztf243_single_standard_double():
    """README.md:82-96: period 0.096536, max power 1.080289 (un-normalised)."""
    o, t, y, _ = load_data("243_normalized_ztf_filter2.txt")
    pg, per, pw, ix = oracle.single(t, y, None, 3.142, 150.796, 1000000, False, F64,
                                    normalize=False, nthreads=8)
    assert round(per, 6) == 0.096536
    assert round(pw, 6) == 1.080289


def test_kat_8205_release_standard_double():
    """README.md:175-197: period 0.564541 at 1e4 bins."""
    o, t, y, _ = load_data("8205_normalized.txt")
    pg, per, pw, ix = oracle.single(t, y, None, 3.14159, 150.79645, 10000, False, F64,
                                    normalize=False)
    assert round(per, 6) == 0.564541


def _python_period(pg, fmin_c=0.1, fmax_c=10.0, nf=330000):
    return 1.0 / (fmin_c + (fmax_c - fmin_c) / nf * np.argmax(pg, axis=-1))


@pytest.mark.parametrize("error,expect", [(True, 0.564739), (False, 0.564729)])
def test_kat_python_example_8205_float(error, expect):
    """python_output_lomb_scargle_test_example.txt:1-32 (examples 1-4)."""
    o, t, y, dy = load_data("8205_normalized_with_error.txt")
    pg, _, _, _ = oracle.single(t, y, dy if error else None, TWO_PI * 0.1, TWO_PI * 10.0, 330000,
                                error, F32, nthreads=8)
    assert round(float(_python_period(pg)), 6) == expect


def test_kat_python_example_sdss_float_cpu():
    """python_output_lomb_scargle_test_example.txt:44-54 (example 6, CPU)."""
    o, t, y, dy = load_data("SDSS_stripe82_band_z.txt")
    pg, _, _, _ = oracle.batch(o, t, y, dy, TWO_PI * 0.1, TWO_PI * 10.0, 330000, True, F32,
                               nthreads=8)
    per = _python_period(pg)
    assert f"{per.sum():.6f}" == "96.527661"
    assert [round(float(p), 6) for p in per[:3]] == [0.529526, 0.554794, 0.570158]


@pytest.mark.slow
def test_kat_8205_single_generalized_double():
    """README.md:120-136: period 0.564736, power 0.712905 at 1e6 bins."""
    o, t, y, dy = load_data("8205_normalized_with_error.txt")
    pg, per, pw, ix = oracle.single(t, y, dy, 3.14159, 150.79645, 1000000, True, F64, nthreads=8)
    assert round(per, 6) == 0.564736
    assert round(pw, 6) == 0.712905


def test_auto_frequency_count_formula():
    """Example 9 of the recorded output: window 119.989583 d -> 11878 bins."""
    from gpu_lomb_scargle_amd.lombscarglegpu import computeNumFreqAuto
    ids = np.array([0, 0, 1, 1])
    t = np.array([0.0, 119.989583, 5.0, 6.0])
    assert computeNumFreqAuto(ids, t, 0.1, 10.0) == 11878


# --------------------------------------------------------------------------
# closed forms, invariances, special cases
# --------------------------------------------------------------------------
def _std_lstsq(omega, t, y):
    A = np.stack([np.cos(omega * t), np.sin(omega * t)], 1)
    coef, *_ = np.linalg.lstsq(A, y, rcond=None)
    return 0.5 * float(y @ (A @ coef))


def _gen_lstsq(omega, t, y, dy):
    w = 1.0 / dy ** 2
    sw = np.sqrt(w)
    A = np.stack([np.cos(omega * t), np.sin(omega * t), np.ones_like(t)], 1)
    coef, *_ = np.linalg.lstsq(A * sw[:, None], y * sw, rcond=None)
    chi2 = float(np.sum(w * (y - A @ coef) ** 2))
    ybar = np.sum(w * y) / np.sum(w)
    chi0 = float(np.sum(w * (y - ybar) ** 2))
    return 1.0 - chi2 / chi0


@pytest.mark.parametrize("seed", range(6))
def test_standard_power_is_least_squares_projection(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 64))
    t = np.sort(rng.uniform(0, 50, n))
    y = rng.normal(size=n) + np.sin(2.3 * t)
    for omega in rng.uniform(0.5, 40, 8):
        p = oracle.std_power(omega, t, y)
        assert p == pytest.approx(_std_lstsq(omega, t, y), rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_generalized_power_is_weighted_fit_reduction(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(6, 64))
    t = np.sort(rng.uniform(0, 80, n))
    dy = rng.uniform(0.05, 0.5, n)
    y = 17.0 + 0.4 * np.sin(1.7 * t + 0.3) + rng.normal(0, dy)
    yc = oracle.precenter(y, dy)
    for omega in rng.uniform(0.5, 40, 8):
        p = oracle.gen_power(omega, t, yc, dy)
        assert p == pytest.approx(_gen_lstsq(omega, t, y, dy), rel=1e-8, abs=1e-11)


def test_invariances():
    rng = np.random.default_rng(3)
    n = 200
    t = np.sort(rng.uniform(0, 300, n))
    dy = rng.uniform(0.05, 0.3, n)
    y = 16.0 + np.sin(0.9 * t) + rng.normal(0, dy)
    perm = rng.permutation(n)
    yc = oracle.precenter(y, dy)
    for omega in (0.4, 0.9, 3.3, 27.0):
        ps = oracle.std_power(omega, t, y - y.mean())
        pg = oracle.gen_power(omega, t, yc, dy)
        # time shift (tau absorbs it)
        assert oracle.std_power(omega, t + 1000.125, y - y.mean()) == pytest.approx(ps, rel=1e-9)
        assert oracle.gen_power(omega, t + 1000.125, yc, dy) == pytest.approx(pg, rel=1e-9)
        # permutation of observations
        assert oracle.std_power(omega, t[perm], (y - y.mean())[perm]) == pytest.approx(ps, rel=1e-12)
        assert oracle.gen_power(omega, t[perm], yc[perm], dy[perm]) == pytest.approx(pg, rel=1e-12)
        # scaling: standard scales by c^2, generalized is invariant
        assert oracle.std_power(omega, t, 3.0 * (y - y.mean())) == pytest.approx(9.0 * ps, rel=1e-12)
        assert oracle.gen_power(omega, t, 3.0 * yc, dy) == pytest.approx(pg, rel=1e-12)
        # generalized: a constant offset and a common dy scale do not matter
        assert oracle.gen_power(omega, t, yc + 5.0, dy) == pytest.approx(pg, rel=1e-9)
        assert oracle.gen_power(omega, t, yc, 2.5 * dy) == pytest.approx(pg, rel=1e-12)


def test_pure_sinusoid_special_case():
    n = 4000
    t = np.linspace(0, 400, n) + 0.01 * np.sin(np.arange(n))
    omega0, amp = 2.0, 1.5
    y = amp * np.cos(omega0 * t + 0.4)
    p = oracle.std_power(omega0, t, y)
    assert p == pytest.approx(n * amp ** 2 / 4.0, rel=2e-3)
    # normalised by 2/(N sigma^2) -> ~1
    pg, _, _, _ = oracle.single(t, y, None, omega0, omega0 + 1e-3, 1, False, F64)
    assert pg[0] == pytest.approx(1.0, rel=2e-3)
    pg, _, _, _ = oracle.single(t, y + 10.0, np.full(n, 0.1), omega0, omega0 + 1e-3, 1, True, F64)
    assert pg[0] == pytest.approx(1.0, rel=1e-6)


def test_scipy_cross_check():
    scipy_signal = pytest.importorskip("scipy.signal")
    o, t, y, dy = load_data("8205_normalized_with_error.txt")
    t, y, dy = t[:400], y[:400], dy[:400]
    omegas = oracle.freq_grid(TWO_PI * 0.1, TWO_PI * 10.0, 500)
    mine = np.array([oracle.std_power(w, t, y) for w in omegas])
    ref = scipy_signal.lombscargle(t, y, omegas, precenter=False, normalize=False)
    assert np.allclose(mine, ref, rtol=1e-9, atol=1e-9 * ref.max())
    yc = oracle.precenter(y, dy)
    mine = np.array([oracle.gen_power(w, t, yc, dy) for w in omegas])
    ref = scipy_signal.lombscargle(t, y, omegas, normalize=True, floating_mean=True,
                                   weights=1.0 / dy ** 2)
    assert np.allclose(mine, ref, rtol=1e-8, atol=1e-10)


def test_frequency_grid_excludes_max():
    g = oracle.freq_grid(1.005, 150.796, 1000, F64)
    step = (150.796 - 1.005) / 1000
    assert g[0] == 1.005 and g[-1] < 150.796
    assert g[-1] == pytest.approx(150.796 - step, rel=1e-14)


def test_object_ranges_are_runs():
    ids = np.array([7, 7, 2, 2, 2, 7, 7, 3], np.uint32)
    lo, hi, got = oracle.object_ranges(ids)
    assert lo.tolist() == [0, 2, 5, 7] and hi.tolist() == [1, 4, 6, 7]
    assert got.tolist() == [7, 2, 7, 3]
    lo, hi, got = oracle.object_ranges(np.array([4], np.uint32))
    assert lo.tolist() == [0] and hi.tolist() == [0]


def _top3_scans(p):
    """The reference's three scans written out (main.cu:1165-1219)."""
    mp, mi = [0.0] * 3, [0] * 3
    for r in range(3):
        for j, v in enumerate(p):
            if j in mi[:r]:
                continue
            if mp[r] < v:
                mp[r], mi[r] = v, j
    return mp, mi


def test_top3_semantics():
    rng = np.random.default_rng(11)
    rows = [rng.uniform(-1, 1, 40), np.array([0.5, 0.5, 0.2, 0.5, -1.0]),
            np.array([-1.0, -2.0, 0.0]), np.array([0.0, 3.0, 0.0, 0.0]), rng.integers(0, 3, 30) * 1.0]
    for row in rows:
        per, pw, ix = oracle.top3(row[None, :], 1.0, 2.0, F64)
        mp, mi = _top3_scans(row)
        assert ix[0].tolist() == mi and pw[0].tolist() == mp


def test_degenerate_two_points_not_finite():
    """No epsilon clamp: two points make CC*SS-CS^2 vanish (tau form -> inf/nan
    terms), as in the reference."""
    t = np.array([0.0, 1.0])
    y = np.array([1.0, -1.0])
    p = oracle.std_power(1.3, t, y)
    assert np.isfinite(p) or np.isnan(p) or np.isinf(p)
    if oracle.ref_available(False, F64):
        assert bits_equal(np.array(p), np.array(oracle.ref_power(1.3, t, y, None, False)))


def test_reference_float_generalized_is_libm_sensitive():
    """The reference's FP32 generalized algorithm depends on its libm: at
    MJD-scale times a 1-ulp change of tau flips the float rounding of t - tau
    (ulp 0.004 d at 5e4 d).  Swapping glibc's sinf/cosf/atan2f for correctly
    rounded ones moves a few bins by ~1e-2 while leaving almost all within
    1e-5; FP64 is unaffected.  This bounds what any other float libm (the
    GPU's included) can be expected to reproduce."""
    o, t, y, dy = load_data("SDSS_stripe82_band_z.txt")
    a = oracle.batch(o, t, y, dy, TWO_PI * 0.1, TWO_PI * 10.0, 3000, True, F32, nthreads=8)[0]
    with oracle.float_libm_cr():
        b = oracle.batch(o, t, y, dy, TWO_PI * 0.1, TWO_PI * 10.0, 3000, True, F32, nthreads=8)[0]
        d64 = oracle.batch(o, t, y, dy, TWO_PI * 0.1, TWO_PI * 10.0, 300, True, F64, nthreads=8)[0]
    a64 = oracle.batch(o, t, y, dy, TWO_PI * 0.1, TWO_PI * 10.0, 300, True, F64, nthreads=8)[0]
    a, b = a.astype(F64), b.astype(F64)
    mask = a >= 1e-2 * a.max(axis=1, keepdims=True)
    rel = np.abs(a - b)[mask] / a[mask]
    assert rel.max() > 1e-3            # some bins move a lot ...
    assert (rel > 1e-5).mean() < 1e-4  # ... but very few
    assert np.array_equal(a64, d64)    # the switch only touches the float build
