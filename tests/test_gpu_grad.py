"""GPU derivative path (lsamd_power_grad, lsamd_refine_peaks, the autograd
Function) against the FP64 CPU references.

Checked against: oracle/grad_ref.py (torch float64 autograd of the reference's tau
form, pinned in test_grad_ref.py) and central finite differences of the C
oracle's power.  Tolerances (stated per test):
  * power: relative 1e-10;
  * dP/domega: |d| <= 1e-9 * P * T/2, d2P/domega2: |d| <= 1e-8 * P * (T/2)^2,
    with T the object's time span (the derivative's natural scale);
  * dP/dy, dP/d(dy): |d| <= 1e-8 * max_j |grad_j| per object against the
    autograd reference; against finite differences with relative step 1e-6,
    |d_j| <= 1e-6 * max_j |grad_j| + 1e-11 * P / h_j, the second term being
    the finite difference's share of the C oracle's own FP64 rounding (its
    two-pass tau-form power carries up to ~4e-12 relative noise);
  * dP/domega against central differences with relative step 1e-6: twice
    the a-posteriori truncation estimate (4/3)|D(h) - D(h/2)|, plus
    1e-9 P T/2, plus the oracle rounding share 1e-10 P / h;
  * refined omega: within 1e-8 rad/day of the CPU Newton iteration from the
    same start.
"""
import numpy as np
import pytest

import oracle
from oracle import grad_ref
from gpu_lomb_scargle_amd import _native, device, synth

pytestmark = pytest.mark.gpu
STD, GEN = _native.STANDARD, _native.GENERALIZED


def _batch(K, seed, gen, span=1200.0, median=60):
    rng = np.random.default_rng(seed)
    counts = np.clip(np.rint(np.exp(np.log(median) + 0.5 * rng.standard_normal(K))), 8, 400).astype(int)
    ts, ys, dys, ws = [], [], [], []
    for n in counts:
        t = np.sort(rng.integers(0, int(span), n) + rng.uniform(0, 0.3, n))
        P = np.exp(rng.uniform(np.log(0.1), np.log(5.0)))
        y = np.sin(2 * np.pi * t / P + rng.uniform(0, 6.28)) + rng.normal(0, 0.3, n)
        dy = np.full(n, 1.0)
        if gen:
            y = 0.3 * y + rng.uniform(15, 20)
            dy = np.exp(rng.uniform(np.log(0.01), np.log(0.2), n))
        else:
            y = (y - y.mean()) / y.std()
        ts.append(t)
        ys.append(y)
        dys.append(dy)
        ws.append(2 * np.pi / P)
    off = np.zeros(K + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return np.concatenate(ts), np.concatenate(ys), np.concatenate(dys), off, np.array(ws)


def _dev(torch, t, y, dy, off, dtype):
    td = torch.float64 if dtype == np.float64 else torch.float32
    return (torch.tensor(t, dtype=td, device="cuda"), torch.tensor(y, dtype=td, device="cuda"),
            torch.tensor(dy, dtype=td, device="cuda"), torch.tensor(off, device="cuda"))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("variant", [STD, GEN])
def test_power_and_derivatives_match_reference(variant, dtype):
    import torch
    from gpu_lomb_scargle_amd import grad
    t, y, dy, off, w = _batch(24, 100 + variant, variant == GEN)
    # evaluate at the true frequency, near it and far from it
    rng = np.random.default_rng(7)
    omega = np.where(np.arange(24) % 3 == 0, w, np.where(np.arange(24) % 3 == 1, w + 3e-4,
                                                         rng.uniform(1.005, 150.0, 24)))
    T, Y, D, O = _dev(torch, t, y, dy, off, dtype)
    P, dP, d2P = grad.power_and_derivatives(torch.tensor(omega, dtype=torch.float64, device="cuda"),
                                            T, Y, D if variant == GEN else None, O, variant)
    P, dP, d2P = P.cpu().numpy(), dP.cpu().numpy(), d2P.cpu().numpy()
    tq, yq, dq = (np.asarray(a, dtype).astype(np.float64) for a in (t, y, dy))
    for k in range(24):
        a, b = off[k], off[k + 1]
        r = grad_ref.derivatives(variant, omega[k], tq[a:b], yq[a:b], dq[a:b] if variant == GEN else None)
        half = 0.5 * (tq[a:b].max() - tq[a:b].min())
        assert P[k] == pytest.approx(r["power"], rel=1e-10)
        assert abs(dP[k] - r["dpower"]) <= 1e-9 * r["power"] * half, (k, dP[k], r["dpower"])
        assert abs(d2P[k] - r["d2power"]) <= 1e-8 * r["power"] * half ** 2, (k, d2P[k], r["d2power"])


@pytest.mark.parametrize("variant", [STD, GEN])
def test_autograd_gradients_match_reference(variant):
    import torch
    from gpu_lomb_scargle_amd import grad
    K = 16
    t, y, dy, off, w = _batch(K, 200 + variant, variant == GEN)
    T, Y, D, O = _dev(torch, t, y, dy, off, np.float64)
    om = torch.tensor(w + 1e-4, dtype=torch.float64, device="cuda", requires_grad=True)
    Y.requires_grad_(True)
    D.requires_grad_(variant == GEN)
    P = grad.power(om, T, Y, D if variant == GEN else None, O, variant)
    up = torch.linspace(0.5, 1.5, K, dtype=torch.float64, device="cuda")
    (P * up).sum().backward()
    gom, gy = om.grad.cpu().numpy(), Y.grad.cpu().numpy()
    gd = D.grad.cpu().numpy() if variant == GEN else None
    upn = up.cpu().numpy()
    for k in range(K):
        a, b = off[k], off[k + 1]
        r = grad_ref.derivatives(variant, w[k] + 1e-4, t[a:b], y[a:b], dy[a:b] if variant == GEN else None)
        half = 0.5 * (t[a:b].max() - t[a:b].min())
        assert abs(gom[k] - upn[k] * r["dpower"]) <= 1e-9 * upn[k] * r["power"] * half
        ref_y = upn[k] * r["grad_y"]
        assert np.abs(gy[a:b] - ref_y).max() <= 1e-8 * np.abs(ref_y).max(), k
        if variant == GEN:
            ref_d = upn[k] * r["grad_dy"]
            assert np.abs(gd[a:b] - ref_d).max() <= 1e-8 * np.abs(ref_d).max(), k


@pytest.mark.parametrize("variant", [STD, GEN])
def test_gradients_match_finite_differences_64_objects(variant):
    """64 objects, FP64 central differences of the C oracle with relative
    step 1e-6 (omega, y, dy)."""
    import torch
    from gpu_lomb_scargle_amd import grad
    K = 64
    t, y, dy, off, w = _batch(K, 300 + variant, variant == GEN, median=30)
    if variant == GEN:  # generalized power is shift-invariant: centre (FD noise)
        for k in range(K):
            a, b = off[k], off[k + 1]
            u = 1.0 / dy[a:b] ** 2
            y[a:b] -= (u * y[a:b]).sum() / u.sum()
    omega = w + 2e-4
    T, Y, D, O = _dev(torch, t, y, dy, off, np.float64)
    a_ = _native.PointArgs()
    lib = _native.core()
    om = torch.tensor(omega, dtype=torch.float64, device="cuda")
    dP = torch.empty(K, dtype=torch.float64, device="cuda")
    Pw = torch.empty_like(dP)
    gy = torch.empty_like(Y)
    gd = torch.empty_like(D)
    a_.variant, a_.dtype, a_.normalize, a_.num_objects, a_.num_obs = variant, _native.F64, 1, K, len(t)
    a_.offsets, a_.t, a_.y = O.data_ptr(), T.data_ptr(), Y.data_ptr()
    a_.dy = D.data_ptr() if variant == GEN else None
    a_.omega, a_.power, a_.dpower = om.data_ptr(), Pw.data_ptr(), dP.data_ptr()
    a_.grad_y = gy.data_ptr()
    a_.grad_dy = gd.data_ptr() if variant == GEN else None
    from ctypes import byref
    _native.check(lib, lib.lsamd_power_grad(byref(a_), None))
    torch.cuda.synchronize()
    dP, Pw, gy, gd = dP.cpu().numpy(), Pw.cpu().numpy(), gy.cpu().numpy(), gd.cpu().numpy()

    def f(k, om, yy, dd):
        # times centred on the object's span: the power is invariant under a
        # time shift, and the oracle's phase rounding shrinks
        tt = t[off[k]:off[k + 1]]
        tt = tt - 0.5 * (tt.min() + tt.max())
        if variant == GEN:
            return oracle.gen_power(om, tt, yy, dd)
        return oracle.std_power(om, tt, yy) * 2.0 / (len(yy) * oracle.stddev_sq(yy))

    worst = 0.0
    for k in range(K):
        a, b = off[k], off[k + 1]
        yy, dd = y[a:b].copy(), dy[a:b].copy()
        half = 0.5 * (t[a:b].max() - t[a:b].min())
        h = 1e-6 * omega[k]
        D = lambda hh: (f(k, omega[k] + hh, yy, dd) - f(k, omega[k] - hh, yy, dd)) / (2 * hh)
        fd, fd2 = D(h), D(h / 2)
        bound = 2 * (4 / 3) * abs(fd - fd2) + 1e-9 * Pw[k] * half + 1e-10 * Pw[k] / h
        assert abs(fd - dP[k]) <= bound, (k, fd, dP[k], bound)
        gyf, gdf = np.empty(b - a), np.empty(b - a)
        hys, hds = np.empty(b - a), np.empty(b - a)
        for j in range(b - a):
            hys[j] = hy = 1e-6 * max(abs(yy[j]), 1.0)
            p, m = yy.copy(), yy.copy()
            p[j] += hy
            m[j] -= hy
            gyf[j] = (f(k, omega[k], p, dd) - f(k, omega[k], m, dd)) / (2 * hy)
            if variant == GEN:
                hds[j] = hd = 1e-6 * dd[j]
                p, m = dd.copy(), dd.copy()
                p[j] += hd
                m[j] -= hd
                gdf[j] = (f(k, omega[k], yy, p) - f(k, omega[k], yy, m)) / (2 * hd)
        g = gy[a:b]
        tol = 1e-6 * np.abs(g).max() + 1e-11 * Pw[k] / hys
        worst = max(worst, float((np.abs(gyf - g) / tol).max()))
        assert np.all(np.abs(gyf - g) <= tol), k
        if variant == GEN:
            g = gd[a:b]
            tol = 1e-6 * np.abs(g).max() + 1e-11 * Pw[k] / hds
            worst = max(worst, float((np.abs(gdf - g) / tol).max()))
            assert np.all(np.abs(gdf - g) <= tol), k
    print(f"worst gradient difference vs finite differences, in units of the tolerance: {worst:.2f}")


def test_torch_gradcheck():
    import torch
    from gpu_lomb_scargle_amd import grad
    t, y, dy, off, w = _batch(3, 400, True, median=12)
    T, Y, D, O = _dev(torch, t, y, dy, off, np.float64)
    om = torch.tensor(w + 1e-3, dtype=torch.float64, device="cuda", requires_grad=True)
    Y.requires_grad_(True)
    D.requires_grad_(True)
    # omega scale: the power varies on 2 pi / T ~ 5e-3 rad/day, so eps 1e-6 is fine
    assert torch.autograd.gradcheck(lambda o, yy, dd: grad.power(o, T, yy, dd, O, GEN), (om, Y, D),
                                    eps=1e-6, atol=1e-6, rtol=1e-5)
    Ys = torch.tensor(y, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda o, yy: grad.power(o, T, yy, None, O, STD), (om, Ys),
                                    eps=1e-6, atol=1e-6, rtol=1e-5)


@pytest.mark.parametrize("variant", [STD, GEN])
def test_refine_matches_cpu_newton(variant):
    import torch
    from gpu_lomb_scargle_amd import grad
    K = 24
    t, y, dy, off, w = _batch(K, 500 + variant, variant == GEN)
    step = 3e-4  # a grid spacing of 5e5 bins over 1.005..150.796
    rng = np.random.default_rng(9)
    om0 = w + rng.uniform(-0.5, 0.5, K) * step
    T, Y, D, O = _dev(torch, t, y, dy, off, np.float64)
    om, P, it = grad.refine_peaks(torch.tensor(om0, device="cuda"), T, Y, D if variant == GEN else None,
                                  O, variant, max_step=step, tol=1e-9, max_iter=30)
    om, P, it = om.cpu().numpy(), P.cpu().numpy(), it.cpu().numpy()
    for k in range(K):
        a, b = off[k], off[k + 1]
        dd = dy[a:b] if variant == GEN else None
        ref, rit = grad_ref.newton_refine(variant, om0[k], t[a:b], y[a:b], dd, max_step=step)
        assert rit > 0 and it[k] > 0, (k, rit, it[k])
        assert abs(om[k] - ref) <= 1e-8, (k, om[k], ref)
        r0 = grad_ref.derivatives(variant, om0[k], t[a:b], y[a:b], dd)
        r1 = grad_ref.derivatives(variant, om[k], t[a:b], y[a:b], dd)
        assert P[k] == pytest.approx(r1["power"], rel=1e-10)
        assert r1["power"] >= r0["power"] * (1 - 1e-12)
        assert r1["d2power"] < 0


@pytest.mark.parametrize("variant", [STD, GEN])
def test_grid_search_then_refine_recovers_periods(variant):
    """The refinement workload at reduced size: 256 objects of the 4096-object
    batch, FP32-split grid search over 5e5 bins (top-1 on the device), then
    Newton refinement.  Refined frequencies move below the grid spacing and
    never lower the power."""
    import torch
    from gpu_lomb_scargle_amd import grad
    data = synth.refinement_batch_4096(with_error=variant == GEN)
    K = 256
    rows = int(data["counts"][:K].sum())
    t, y = data["t"][:rows], data["y"][:rows]
    dy = data["dy"][:rows] if variant == GEN else None
    T, Y, D, O = device.numpy_to_device(data["obj_id"][:rows], t, y, dy, np.float64, "cuda")
    nf = 500000
    plan = device.SweepPlan(T, Y, D, O, min_freq=synth.OMEGA_MIN, max_freq=synth.OMEGA_MAX, num_freqs=nf,
                            variant=variant, form=_native.FORM_SPLIT32, pgram=False, top_k=1)
    plan.run()
    idx = plan.results()["index"][:, 0].to(torch.float64)
    dw = (synth.OMEGA_MAX - synth.OMEGA_MIN) / nf
    om0 = synth.OMEGA_MIN + idx * dw
    P0, _, _ = grad.power_and_derivatives(om0, T, Y, D, O, variant)
    om, P, it = grad.refine_peaks(om0, T, Y, D, O, variant, max_step=dw, tol=1e-9, max_iter=30)
    torch.cuda.synchronize()
    om0, om, P0, P, it = (x.cpu().numpy() for x in (om0, om, P0, P, it))
    assert np.all(P >= P0 * (1 - 1e-12))
    assert np.all(np.abs(om - om0) <= 2 * dw)
    assert (it > 0).mean() >= 0.99
    wt = 2 * np.pi / data["period"][:K]
    hit = np.abs(om0 - wt) < 2 * dw
    e0 = np.abs(om0 - wt)[hit]
    e1 = np.abs(om - wt)[hit]
    print(f"objects with the grid peak at the true frequency: {hit.sum()}/{K}; "
          f"median |omega - omega_true| grid {np.median(e0):.2e}, refined {np.median(e1):.2e} rad/day")
    assert hit.sum() > K // 3
    assert np.median(e1) < np.median(e0)
