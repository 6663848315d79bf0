"""Pins the FP64 reference gradients (oracle/grad_ref.py) on the CPU.

The reference has no derivatives, so the gradients the GPU derivative kernel
is checked against are pinned here by what the mathematics fixes:
  * the value equals the C oracle's power (bit-equal to the reference CPU
    build) at the same frequency,
  * d/domega, d/dy, d/d(dy) equal central finite differences of the C
    oracle's power,
  * invariances: t -> t + t0 leaves every derivative unchanged; the
    generalized power is invariant under y -> y + k (sum_j dP/dy_j = 0), under
    dy -> k dy (sum_j dy_j dP/d(dy_j) = 0) and under y -> c y (sum_j y_j
    dP/dy_j = 0), as is the normalised standard power under y -> c y,
  * a planted mistake in the reference formula (a dropped cross term) is
    caught by the finite-difference check.
"""
import numpy as np
import pytest

import oracle
from oracle import grad_ref


def _obj(seed, n=80, span=300.0, gen=False):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.integers(0, int(span), n) + rng.uniform(0, 0.3, n))
    P = np.exp(rng.uniform(np.log(0.2), np.log(5.0)))
    y = np.sin(2 * np.pi * t / P + rng.uniform(0, 6.28)) + rng.normal(0, 0.3, n)
    dy = None
    if gen:
        y = y * 0.3 + rng.uniform(15, 20)
        dy = np.exp(rng.uniform(np.log(0.01), np.log(0.2), n))
    return t, y, dy, 2 * np.pi / P


def _oracle_power(variant, omega, t, y, dy, normalize=True):
    if variant == 1:
        return oracle.gen_power(omega, t, y, dy)
    p = oracle.std_power(omega, t, y)
    if normalize:
        p *= 2.0 / (len(y) * oracle.stddev_sq(y))
    return p


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_value_matches_c_oracle(variant, seed):
    t, y, dy, w0 = _obj(seed, gen=variant == 1)
    for om in (w0, 1.7, 23.456):
        d = grad_ref.derivatives(variant, om, t, y, dy)
        ref = _oracle_power(variant, om, t, y, dy)
        assert d["power"] == pytest.approx(ref, rel=1e-11, abs=1e-14)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("seed", [4, 5])
def test_derivatives_match_finite_differences(variant, seed):
    t, y, dy, w0 = _obj(seed, gen=variant == 1)
    tm = t - 0.5 * (t.min() + t.max())
    if variant == 1:
        # the generalized power is invariant under y -> y + k; centring keeps
        # the oracle's FP64 rounding noise (YY of 15..20 mag values) below
        # what a 1e-6 relative step in dy can resolve
        w = 1.0 / dy ** 2
        y = y - (w * y).sum() / w.sum()
    for om in (w0 + 1e-3, 3.21):
        d = grad_ref.derivatives(variant, om, tm, y, dy)
        # omega: Richardson-extrapolated central differences (error O(h^4))
        f = lambda w: _oracle_power(variant, w, tm, y, dy)
        h = 1e-5
        d1 = lambda h: (f(om + h) - f(om - h)) / (2 * h)
        fd = (4 * d1(h / 2) - d1(h)) / 3
        scale = abs(d["power"]) * np.abs(tm).max()
        assert abs(fd - d["dpower"]) <= 1e-6 * scale
        d2 = (f(om + 1e-4) - 2 * f(om) + f(om - 1e-4)) / 1e-8
        assert abs(d2 - d["d2power"]) <= 1e-3 * abs(d["power"]) * np.abs(tm).max() ** 2
        # y and dy: central differences with relative step 1e-6
        gy = np.empty(len(y))
        gd = np.empty(len(y))
        for j in range(len(y)):
            hy = 1e-6 * max(abs(y[j]), 1.0)
            yp, ym = y.copy(), y.copy()
            yp[j] += hy
            ym[j] -= hy
            gy[j] = (_oracle_power(variant, om, tm, yp, dy) - _oracle_power(variant, om, tm, ym, dy)) / (2 * hy)
            if variant == 1:
                hd = 1e-6 * dy[j]
                dp, dm = dy.copy(), dy.copy()
                dp[j] += hd
                dm[j] -= hd
                gd[j] = (_oracle_power(variant, om, tm, y, dp) - _oracle_power(variant, om, tm, y, dm)) / (2 * hd)
        gscale = np.abs(d["grad_y"]).max()
        assert np.abs(gy - d["grad_y"]).max() <= 1e-6 * gscale
        if variant == 1:
            assert np.abs(gd - d["grad_dy"]).max() <= 1e-6 * np.abs(d["grad_dy"]).max()


@pytest.mark.parametrize("variant", [0, 1])
def test_invariances(variant):
    t, y, dy, w0 = _obj(11, gen=variant == 1)
    a = grad_ref.derivatives(variant, w0, t, y, dy)
    b = grad_ref.derivatives(variant, w0, t + 2451234.5 - 2451000.0, y, dy)
    for k in ("power", "dpower", "d2power"):
        assert b[k] == pytest.approx(a[k], rel=1e-7, abs=1e-9 * abs(a["power"]))
    np.testing.assert_allclose(b["grad_y"], a["grad_y"], rtol=1e-6, atol=1e-7 * np.abs(a["grad_y"]).max())
    g = a["grad_y"]
    assert abs(np.dot(y, g)) <= 1e-10 * np.abs(y * g).sum()  # y -> c y
    if variant == 1:
        assert abs(g.sum()) <= 1e-10 * np.abs(g).sum()  # y -> y + k
        gd = a["grad_dy"]
        assert abs(np.dot(dy, gd)) <= 1e-10 * np.abs(dy * gd).sum()  # dy -> k dy


def test_planted_mistake_is_caught():
    """A reference with the cross term of the second denominator dropped
    disagrees with the finite differences of the C oracle."""
    import torch
    t, y, _, w0 = _obj(21)
    om = torch.tensor(w0 + 0.01, dtype=torch.float64, requires_grad=True)
    tt, yy = torch.as_tensor(t), torch.as_tensor(y)

    def broken(omega):
        arg = omega * tt
        c, s = torch.cos(arg), torch.sin(arg)
        xc, xs = (yy * c).sum(), (yy * s).sum()
        cc, ss, cs = (c * c).sum(), (s * s).sum(), (c * s).sum()
        tau = torch.atan2(2.0 * cs, cc - ss) / (2.0 * omega)
        ct, st = torch.cos(omega * tau), torch.sin(omega * tau)
        return 0.5 * ((ct * xc + st * xs) ** 2 / (ct * ct * cc + 2 * ct * st * cs + st * st * ss)
                      + (ct * xs - st * xc) ** 2 / (ct * ct * ss + st * st * cc))

    p = broken(om)
    (g,) = torch.autograd.grad(p, om)
    f = lambda w: oracle.std_power(w, t, y)
    fd = (f(float(om) + 1e-6) - f(float(om) - 1e-6)) / 2e-6
    assert abs(float(g) - fd) > 1e-3 * abs(fd)


def test_newton_reference_reaches_a_stationary_point():
    t, y, dy, w0 = _obj(31, gen=True)
    om, it = grad_ref.newton_refine(1, w0 + 2e-4, t, y, dy, max_step=1e-3)
    assert it > 0
    d = grad_ref.derivatives(1, om, t, y, dy)
    assert d["d2power"] < 0
    assert abs(d["dpower"]) <= 1e-6 * abs(d["power"]) * np.abs(t - t.mean()).max()
    assert d["power"] >= grad_ref.derivatives(1, w0 + 2e-4, t, y, dy)["power"]
