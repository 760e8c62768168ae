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


def mh_r2star_tape(y, te, x0, lb, ub, steps, iterations, reps, thinning, n_burnin, generator):
    """Metropolis-Hastings (utils/mcmc.m:82-190) for the mono-exponential
    model as the reference executes it on the GPU: per iteration a normal
    draw for every variable, the proposal, the out-of-bounds flag and clamp,
    the forward signal and its sum of squares, logP_Gaussian
    (utils/logP_Gaussian.m:18-22), a uniform draw and the acceptance, each
    its own kernel over every voxel; repetitions run one after the other.

    y [M, Nv], te [M], x0 [3, Nv] (M0, R2*, noise) on the GPU.  Returns the
    posterior mean [3, Nv] over the retained samples of every repetition."""
    M, nv = y.shape
    dev = y.device
    te = te.reshape(-1, 1)
    lbt = torch.tensor(lb, dtype=torch.float32, device=dev).reshape(3, 1)
    ubt = torch.tensor(ub, dtype=torch.float32, device=dev).reshape(3, 1)
    st = torch.tensor(steps, dtype=torch.float32, device=dev).reshape(3, 1)

    def logp(x):
        s = x[0:1] * torch.exp(-te * x[1:2])
        sos = ((s - y) ** 2).sum(0)
        n2 = x[2] * x[2]
        return -sos / (2.0 * n2) - 0.5 * M * torch.log(2.0 * math.pi * n2)

    sums = torch.zeros((3, nv), dtype=torch.float64, device=dev)
    n = 0
    for _ in range(reps):
        x = torch.minimum(torch.maximum(x0, lbt), ubt)
        lp = logp(x)
        for k in range(1, iterations + 1):
            xp = x + st * torch.randn((3, nv), generator=generator, device=dev)
            oob = ((xp < lbt) | (xp > ubt)).any(0)
            xp = torch.minimum(torch.maximum(xp, lbt), ubt)
            lpp = logp(xp)
            ratio = torch.nan_to_num(torch.exp(lpp - lp), nan=1.0).clamp(max=1.0)
            acc = (ratio > torch.rand(nv, generator=generator, device=dev)) & ~oob
            x = torch.where(acc, xp, x)
            lp = torch.where(acc, lpp, lp)
            if k > n_burnin and (k - n_burnin + 1) % thinning == 0:
                sums += x.double()
                n += 1
    return (sums / max(n, 1)).float()
