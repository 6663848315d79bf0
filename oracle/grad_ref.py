"""FP64 reference gradients of the Lomb-Scargle power (test infrastructure).

Only the tests, ``__graft_entry__.smoke()`` and the benchmark's checks use
this module: it is what the GPU derivative path is checked against, never the
path itself.

The power is written here in the reference's own tau form and differentiated
by torch autograd in float64 on the CPU -- an algebraically different route
from the GPU kernel (tau-free closed form, hand-written jets and duals):

  standard     lombscarglecpuinnerloop, source/main.cu:1566-1599, with the
               Python builds' normalisation 2/(N sigma_pop^2),
               normalizepgram + computeStandardDevSquared, main.cu:1295-1339
  generalized  lombscarglecpuinnerloopAstroPy, main.cu:1701-1777 (two
               passes: tau from the weighted C, S, C2, S2 sums, then the
               centred tau-frame sums; weights dy^-2)

The generalized variant's in-place pre-centring (main.cu:1631-1662) does not
change the power (it is invariant under y -> y + k), so it is not applied.

Pinned by tests/test_grad_ref.py: the value against the C oracle (itself
bit-equal to the reference CPU build), the derivatives against central finite
differences of the C oracle and against invariances (shift in t; y -> y + k
and y -> c y; dy -> k dy).
"""
from __future__ import annotations

import numpy as np
import torch

_F = torch.float64


def std_power(omega, t, y, normalize=True):
    """Standard power at one angular frequency (0-d tensors / 1-d tensors)."""
    arg = omega * t
    c, s = torch.cos(arg), torch.sin(arg)
    xc, xs = (y * c).sum(), (y * s).sum()
    cc, ss, cs = (c * c).sum(), (s * s).sum(), (c * s).sum()
    tau = torch.atan2(2.0 * cs, cc - ss) / (2.0 * omega)
    ct, st = torch.cos(omega * tau), torch.sin(omega * tau)
    ct2, st2, cst = ct * ct, st * st, 2.0 * ct * st
    p = 0.5 * ((ct * xc + st * xs) ** 2 / (ct2 * cc + cst * cs + st2 * ss)
               + (ct * xs - st * xc) ** 2 / (ct2 * ss - cst * cs + st2 * cc))
    if normalize:
        n = y.numel()
        mean = y.sum() / n
        var = ((y - mean) ** 2).sum() / n
        p = p * 2.0 / (n * var)
    return p


def gen_power(omega, t, y, dy):
    """Generalized (floating-mean) power at one angular frequency."""
    w = 1.0 / (dy * dy)
    wsum = w.sum()
    arg = omega * t
    so, co = torch.sin(arg), torch.cos(arg)
    S = (w * so).sum() / wsum
    C = (w * co).sum() / wsum
    S2 = (2.0 * w * so * co).sum() / wsum - 2.0 * S * C
    C2 = (w - 2.0 * w * so * so).sum() / wsum - (C * C - S * S)
    tau = 0.5 * torch.atan2(S2, C2) / omega
    arg2 = omega * (t - tau)
    so, co = torch.sin(arg2), torch.cos(arg2)
    Y = (w * y).sum() / wsum
    YY = (w * y * y).sum() / wsum - Y * Y
    Ct = (w * co).sum() / wsum
    St = (w * so).sum() / wsum
    YC = (w * y * co).sum() / wsum - Y * Ct
    YS = (w * y * so).sum() / wsum - Y * St
    CC = (w * co * co).sum() / wsum - Ct * Ct
    SS = (w * so * so).sum() / wsum - St * St
    return (YC * YC / CC + YS * YS / SS) / YY


def power(variant, omega, t, y, dy=None, normalize=True):
    if variant == 1:
        return gen_power(omega, t, y, dy)
    return std_power(omega, t, y, normalize)


def derivatives(variant, omega, t, y, dy=None, normalize=True):
    """P, dP/domega, d2P/domega2, dP/dy, dP/d(dy) (numpy float64) of one
    object at one angular frequency."""
    om = torch.tensor(float(omega), dtype=_F, requires_grad=True)
    tt = torch.as_tensor(np.asarray(t, np.float64))
    yy = torch.as_tensor(np.asarray(y, np.float64)).clone().requires_grad_(True)
    dd = None
    if variant == 1:
        dd = torch.as_tensor(np.asarray(dy, np.float64)).clone().requires_grad_(True)
    p = power(variant, om, tt, yy, dd, normalize)
    inputs = [om, yy] + ([dd] if dd is not None else [])
    g = torch.autograd.grad(p, inputs, create_graph=True)
    d2 = torch.autograd.grad(g[0], om)[0]
    return {"power": float(p.detach()), "dpower": float(g[0].detach()), "d2power": float(d2),
            "grad_y": g[1].detach().numpy().copy(),
            "grad_dy": (g[2].detach().numpy().copy() if dd is not None else np.zeros(len(t)))}


def newton_refine(variant, omega0, t, y, dy=None, normalize=True, max_step=1e-3, tol=1e-9,
                  max_iter=30):
    """CPU Newton iteration on dP/domega with the same step rule as the GPU
    kernel (bounded step, uphill when the curvature is not negative)."""
    om = float(omega0)
    for it in range(1, max_iter + 1):
        d = derivatives(variant, om, t, y, dy, normalize)
        p1, p2 = d["dpower"], d["d2power"]
        step = -p1 / p2 if p2 < 0 else np.inf
        if not np.isfinite(step) or abs(step) > max_step:
            step = 0.0 if (p1 == 0 or not np.isfinite(p1)) else float(np.copysign(max_step, p1))
        om += step
        if abs(step) < tol:
            return om, it
    return om, -max_iter
