"""Differentiable Lomb-Scargle power and Newton peak refinement on the GPU.

``power(omega, t, y, dy, offsets, variant)`` evaluates each object's power
at its own angular frequency omega[k] (FP64, any omega, not just the grid)
through ``lsamd_power_grad`` and is differentiable with respect to omega, y
and dy through a ``torch.autograd.Function`` whose backward is the same HIP
kernel writing dP/domega, dP/dy_j and dP/d(dy_j) scaled by the upstream
gradient.  ``refine_peaks`` runs ``lsamd_refine_peaks``: Newton steps on omega
from a grid argmax to a stationary point of the power.

The power is the reference's (standard, normalised by 2/(N sigma^2) as its
Python builds do, or generalized with weights dy^-2); the reference itself
has no derivatives and no off-grid evaluation.  Nothing here has a CPU
fallback: without a GPU or the built library the calls raise.
"""
from __future__ import annotations

from ctypes import byref, c_void_p

import torch

from . import _native


def _dtype_code(tdtype):
    if tdtype == torch.float32:
        return _native.F32
    if tdtype == torch.float64:
        return _native.F64
    raise ValueError(f"unsupported dtype {tdtype}")


def _ptr(x):
    return None if x is None else c_void_p(x.data_ptr())


def _args(variant, omega, t, y, dy, offsets, normalize):
    if not t.is_cuda:
        raise RuntimeError("the Lomb-Scargle gradient path needs CUDA (HIP) tensors")
    if variant == _native.GENERALIZED and dy is None:
        raise ValueError("the generalized variant needs dy")
    for name, x in (("y", y), ("dy", dy)):
        if x is not None and (x.dtype != t.dtype or x.shape != t.shape):
            raise ValueError(f"{name} must match t in dtype and shape")
    if omega.dtype != torch.float64 or not omega.is_contiguous():
        raise ValueError("omega must be a contiguous float64 tensor")
    K = offsets.numel() - 1
    if omega.numel() != K:
        raise ValueError(f"omega has {omega.numel()} entries for {K} objects")
    a = _native.PointArgs()
    a.variant, a.dtype, a.normalize = int(variant), _dtype_code(t.dtype), int(bool(normalize))
    a.num_objects, a.num_obs = K, t.numel()
    a.offsets, a.t, a.y, a.dy = _ptr(offsets), _ptr(t), _ptr(y), _ptr(dy)
    a.omega = _ptr(omega)
    return a


def _offsets(offsets, device):
    return offsets.to(device=device, dtype=torch.int64).contiguous()


def power_and_derivatives(omega, t, y, dy, offsets, variant, normalize=True, stream=None):
    """P, dP/domega, d2P/domega2 at omega[k] (float64 tensors of K)."""
    lib = _native.core()
    offsets = _offsets(offsets, t.device)
    omega = omega.detach().to(torch.float64).contiguous()
    t, y = t.detach().contiguous(), y.detach().contiguous()
    dy = dy.detach().contiguous() if dy is not None else None
    a = _args(variant, omega, t, y, dy, offsets, normalize)
    K = omega.numel()
    P = torch.empty(K, dtype=torch.float64, device=t.device)
    dP = torch.empty_like(P)
    d2P = torch.empty_like(P)
    a.power, a.dpower, a.d2power = _ptr(P), _ptr(dP), _ptr(d2P)
    s = stream if stream is not None else torch.cuda.current_stream(t.device)
    _native.check(lib, lib.lsamd_power_grad(byref(a), c_void_p(s.cuda_stream)))
    return P, dP, d2P


class LombScarglePower(torch.autograd.Function):
    """P[k] = power of object k at omega[k]; gradients for omega, y, dy."""

    @staticmethod
    def forward(ctx, omega, t, y, dy, offsets, variant, normalize):
        lib = _native.core()
        offsets = _offsets(offsets, t.device)
        om = omega.detach().to(torch.float64).contiguous()
        tt, yy = t.detach().contiguous(), y.detach().contiguous()
        dd = dy.detach().contiguous() if dy is not None else None
        a = _args(variant, om, tt, yy, dd, offsets, normalize)
        K = om.numel()
        P = torch.empty(K, dtype=torch.float64, device=t.device)
        dP = torch.empty_like(P)
        a.power, a.dpower = _ptr(P), _ptr(dP)
        _native.check(lib, lib.lsamd_power_grad(byref(a), c_void_p(torch.cuda.current_stream(t.device).cuda_stream)))
        ctx.save_for_backward(om, tt, yy, dd if dd is not None else torch.empty(0, device=t.device),
                              offsets, dP)
        ctx.variant, ctx.normalize = int(variant), bool(normalize)
        ctx.has_dy = dd is not None
        ctx.omega_dtype = omega.dtype
        return P

    @staticmethod
    def backward(ctx, gP):
        om, tt, yy, dd, offsets, dP = ctx.saved_tensors
        dd = dd if ctx.has_dy else None
        need_omega, _, need_y, need_dy = ctx.needs_input_grad[:4]
        g_omega = g_y = g_dy = None
        gP = gP.detach().to(torch.float64).contiguous()
        if need_omega:
            g_omega = (gP * dP).to(ctx.omega_dtype)
        want_dy = need_dy and ctx.variant == _native.GENERALIZED
        if need_y or want_dy:
            lib = _native.core()
            a = _args(ctx.variant, om, tt, yy, dd, offsets, ctx.normalize)
            a.grad_scale = _ptr(gP)
            gy = torch.empty_like(yy) if need_y else None
            gd = torch.empty_like(dd) if want_dy else None
            a.grad_y, a.grad_dy = _ptr(gy), _ptr(gd)
            _native.check(lib, lib.lsamd_power_grad(byref(a), c_void_p(torch.cuda.current_stream(tt.device).cuda_stream)))
            g_y, g_dy = gy, gd
        if need_dy and g_dy is None and dd is not None:
            g_dy = torch.zeros_like(dd)  # the standard power does not depend on dy
        return g_omega, None, g_y, g_dy, None, None, None


def power(omega, t, y, dy, offsets, variant, normalize=True):
    """Differentiable power of each object at omega[k] (float64 tensor of K)."""
    return LombScarglePower.apply(omega, t, y, dy, offsets, variant, normalize)


def refine_peaks(omega0, t, y, dy, offsets, variant, *, max_step, tol=1e-9, max_iter=30,
                 normalize=True, stream=None):
    """Newton refinement of each object's peak from omega0[k] (float64, e.g.
    the grid argmax).  Returns (omega, power, iterations); iterations < 0
    marks an object that did not reach |step| < tol in max_iter steps."""
    lib = _native.core()
    offsets = _offsets(offsets, t.device)
    omega = omega0.detach().to(torch.float64).clone().contiguous()
    t, y = t.detach().contiguous(), y.detach().contiguous()
    dy = dy.detach().contiguous() if dy is not None else None
    a = _args(variant, omega, t, y, dy, offsets, normalize)
    K = omega.numel()
    P = torch.empty(K, dtype=torch.float64, device=t.device)
    iters = torch.empty(K, dtype=torch.int32, device=t.device)
    a.power, a.iterations = _ptr(P), _ptr(iters)
    a.max_step, a.tol, a.max_iter = float(max_step), float(tol), int(max_iter)
    s = stream if stream is not None else torch.cuda.current_stream(t.device)
    _native.check(lib, lib.lsamd_refine_peaks(byref(a), c_void_p(s.cuda_stream)))
    return omega, P, iters
