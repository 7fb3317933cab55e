"""Gradients of the PSF-matching chi^2 behind ``torch.autograd.Function``.

The model is the reference's subtraction model (src/subtract.cpp:466-639)
with a spatially constant kernel built from the Gauss-Hermite basis of
src/basis.cpp:14-59, whose scale beta and sub-pixel shift (dx, dy) become
free parameters:

    K(u)  = sum_n a_n g_nx(ux - dx) g_ny(uy - dy),   u in [-r, r]^2
    D     = S - b - K (x) R
    chi^2 = sum w D^2.

params = [a_0 .. a_{nc-1}, b, beta, dx, dy] as one float64 CUDA tensor.  The
reference fixes beta from the measured seeing (python/delta/pipeline.py:
267-282) and has no shift and no gradient code; here both are refined by
gradient.  Forward and backward run in libdeltagpu (``dg_kernel_chi2_device``,
``dg_kernel_difference_device``, ``dg_kernel_difference_vjp_device``): a
hand-written adjoint that recomputes each image tile instead of storing
anything image-sized.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _core


def n_params(n_max: int) -> int:
    return _core.n_components(n_max) + 4


class KernelChi2(torch.autograd.Function):
    """chi^2(params); one fused GPU pass gives chi^2 and, when params needs a
    gradient, the whole gradient (saved for backward)."""

    @staticmethod
    def forward(ctx, params, science, reference, weight, n_max, radius):
        need = bool(ctx.needs_input_grad[0])
        out = _core.kernel_chi2_device(science, reference, weight, params.detach().contiguous(),
                                       n_max, radius, grad=need)
        if need:
            ctx.save_for_backward(out[1:])
        return out[0].clone()

    @staticmethod
    def backward(ctx, grad_chi2):
        (grad,) = ctx.saved_tensors
        return grad_chi2 * grad, None, None, None, None, None


class KernelDifference(torch.autograd.Function):
    """D(params) as an image; backward is the vector-Jacobian product
    dL/dparams = -(sum_x G R(x - u) contracted with dK/dparams, sum G)."""

    @staticmethod
    def forward(ctx, params, science, reference, n_max, radius):
        p = params.detach().contiguous()
        d = _core.kernel_difference_device(science, reference, p, n_max, radius)
        ctx.save_for_backward(p, reference)
        ctx.n_max, ctx.radius = n_max, radius
        return d

    @staticmethod
    def backward(ctx, grad_d):
        p, reference = ctx.saved_tensors
        g = _core.kernel_difference_vjp_device(reference, grad_d.contiguous(), p, ctx.n_max,
                                               ctx.radius)
        return g[1:], None, None, None, None


def kernel_chi2(params, science, reference, weight, n_max, radius):
    """Differentiable chi^2 (0-d float64 tensor) of the constant-kernel model."""
    return KernelChi2.apply(params, science, reference, weight, n_max, radius)


def kernel_difference(params, science, reference, n_max, radius):
    """Differentiable difference image D (float32 tensor) of the model."""
    return KernelDifference.apply(params, science, reference, n_max, radius)


def shifted_kernel(params, n_max, radius):
    """K (ks x ks, row = uy + r) of params on the host in float64: the
    reference's sampled Gauss-Hermite functions (src/basis.cpp:14-59) at
    x = i - r - shift."""
    p = np.asarray(params, np.float64)
    nc = _core.n_components(n_max)
    beta, dx, dy = p[nc + 1], p[nc + 2], p[nc + 3]
    ks = 2 * radius + 1

    def table(shift):
        t = (np.arange(ks) - radius - shift) / beta
        e = np.exp(-0.5 * t * t)
        g = np.zeros((n_max + 1, ks))
        hm, hc = np.ones(ks), 2.0 * t
        for n in range(n_max + 1):
            if n == 0:
                hn = hm
            elif n == 1:
                hn = hc
            else:
                hn = 2.0 * t * hc - 2.0 * (n - 1) * hm
                hm, hc = hc, hn
            ln = -0.5 * (n * math.log(2.0) + math.lgamma(n + 1.0) + 0.5 * math.log(math.pi)
                         + math.log(beta))
            g[n] = math.exp(ln) * hn * e
        return g

    gx, gy = table(dx), table(dy)
    k = np.zeros((ks, ks))
    c = 0
    for tot in range(n_max + 1):
        for ny in range(tot + 1):
            k += p[c] * np.outer(gy[ny], gx[tot - ny])
            c += 1
    return k


def chi2_weight(science_var, reference_var, radius, kernel_power=1.0):
    """Inverse variance of D, 1 / (Var S + kernel_power Var R) with
    kernel_power = sum K^2, held fixed; zero within `radius` of the edge,
    where the model's zero padding of R does not describe S."""
    w = 1.0 / (science_var + kernel_power * reference_var)
    w = torch.where(torch.isfinite(w) & (w > 0), w, torch.zeros_like(w))
    if radius > 0:
        w[:radius, :] = 0
        w[-radius:, :] = 0
        w[:, :radius] = 0
        w[:, -radius:] = 0
    return w.contiguous()


def refine_kernel(science, reference, weight, params0, n_max, radius, free=None, max_iter=400,
                  history_size=30, tolerance_grad=1e-9, tolerance_change=1e-13):
    """Minimise chi^2 over the free entries of params with L-BFGS (strong
    Wolfe line search) on the adjoint gradient.  `free` is a boolean mask of
    params (default: all).  Returns {'params', 'chi2', 'evaluations',
    'history'}.  chi^2 is divided by the number of weighted pixels inside the
    optimiser so the tolerances are scale-free."""
    dev = science.device
    # the optimiser state (32-odd numbers) lives on the host; each evaluation
    # moves the parameter vector to the device and the gradient back
    base = torch.as_tensor(np.asarray(params0, np.float64) if not torch.is_tensor(params0)
                           else params0.detach().cpu(), dtype=torch.float64).clone()
    n = base.numel()
    mask = (torch.ones(n, dtype=torch.bool) if free is None
            else torch.as_tensor(np.asarray(free, bool) if not torch.is_tensor(free)
                                 else free.cpu(), dtype=torch.bool))
    idx = torch.nonzero(mask).flatten()
    z = base[idx].clone().requires_grad_(True)
    npix = float((weight > 0).sum()) if weight is not None else float(science.numel())
    opt = torch.optim.LBFGS([z], lr=1.0, max_iter=max_iter, max_eval=int(max_iter * 1.5),
                            history_size=history_size, tolerance_grad=tolerance_grad,
                            tolerance_change=tolerance_change, line_search_fn="strong_wolfe")
    history = []

    def full():
        return base.index_put((idx,), z)

    def closure():
        opt.zero_grad()
        loss = kernel_chi2(full().to(dev), science, reference, weight, n_max, radius) / npix
        loss = loss.cpu()
        loss.backward()
        history.append(float(loss.detach()))
        return loss

    opt.step(closure)
    with torch.no_grad():
        p = full().detach().to(dev)
        chi2 = float(kernel_chi2(p, science, reference, weight, n_max, radius))
    return {"params": p, "chi2": chi2, "evaluations": len(history), "history": history}
