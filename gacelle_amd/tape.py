"""Plain per-operation GPU tape: the baseline the fused path is measured
against.

This restates askadam.model_gradient (utils/askadam.m:23-117) for the
mono-exponential model the way MATLAB's dlarray/dlgradient executes it:
every elementwise operation is its own GPU kernel with a materialised
intermediate, the gradient comes from a recorded reverse tape (PyTorch
autograd here), and the Adam update is applied with the same formula as
the fused kernel.  No entry point of the package uses it; bench.py times
it beside the fused path.
"""
from __future__ import annotations

import math

import torch


def askadam_r2star_tape(y, te, lb, ub, z0, iterations, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, loss="l1"):
    """Returns (z_final, last loss).  y [M, Nv], te [M], z0 [2, Nv] on the GPU."""
    z = z0.clone()
    m = torch.zeros_like(z)
    v = torch.zeros_like(z)
    te = te.reshape(-1, 1)
    last = None
    for epoch in range(1, iterations + 1):
        zz = z.detach().requires_grad_(True)
        m0 = lb[0] + (ub[0] - lb[0]) * (1.0 / (1.0 + torch.exp(-zz[0])))
        r2 = lb[1] + (ub[1] - lb[1]) * (1.0 / (1.0 + torch.exp(-zz[1])))
        s = m0.unsqueeze(0) * torch.exp(-te * r2.unsqueeze(0))
        s = torch.where(torch.isfinite(s), s, torch.zeros_like(s))
        d = s - y
        res = torch.abs(d) if loss == "l1" else d * d
        lval = res.mean()
        (g,) = torch.autograd.grad(lval, zz)
        last = lval.detach()
        if epoch < iterations:
            with torch.no_grad():
                m.mul_(beta1).add_(g, alpha=1 - beta1)
                v.mul_(beta2).addcmul_(g, g, value=1 - beta2)
                lr_eff = lr * math.sqrt(1 - beta2 ** epoch) / (1 - beta1 ** epoch)
                z.sub_(lr_eff * m / (torch.sqrt(v) + eps))
    return z, float(last)
