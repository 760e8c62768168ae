"""Shared builders for the tests: synthetic problems and C-ABI descriptors."""
from __future__ import annotations

import numpy as np

from gacelle_amd import _abi, models
from gacelle_amd.askadam import askadam
from gacelle_amd.mcmc import mcmc

TE12 = np.arange(2, 47, 4) * 1e-3          # 2:4:46 ms
TE_JOINT = np.arange(2, 25, 2) * 1e-3      # 2:2:24 ms
FA_JOINT = np.array([3, 8, 15, 25], float)  # degrees
TR_JOINT = 30e-3


def no_stop(**kw):
    f = {"tol": 0, "convergenceValue": -1, "convergenceStepTol": 0, "convergenceGradTol": 0, "verbose": False}
    f.update(kw)
    return f


def r2s_problem(nv, seed=0, snr=100):
    rng = np.random.default_rng(seed)
    s0 = rng.uniform(0.5, 1.5, nv)
    r2 = rng.uniform(10, 60, nv)
    s = s0 * np.exp(-np.outer(TE12, r2))
    sigma = s0.mean() / snr
    y = (s + sigma * rng.standard_normal(s.shape)).astype(np.float32)
    return y, np.stack([s0, r2]), sigma


def r2s_setup(fitting, solver="askadam"):
    m = models.gpuR2starMapping(TE12)
    if solver == "askadam":
        f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200], **fitting))
    else:
        f = mcmc.check_set_default_basic(dict(modelParams=["M0", "R2star", "noise"], lb=[0, 0.1, 0.001],
                                              ub=[2, 200, 0.1], **fitting))
    md = m.gc_model(f, solver, TE12.size)
    return m, f, md


def logit_start(theta0, md):
    P = theta0.shape[0]
    z = np.zeros_like(theta0, dtype=np.float64)
    for p in range(P):
        lb, ub = md.lb[p], md.ub[p]
        eb = 1e-4 * (ub - lb)
        c = np.clip(theta0[p], lb + eb, ub - eb)
        t = (c - lb) / (ub - lb)
        z[p] = np.log(t / (1 - t))
    return z.astype(np.float32)


def mcmc_cfg(algo, iteration, thinning, burnin, rep=1, W=1, seed=48463, replay=False, steps=(), redblack=False,
             a=2.0, start_range=1e-3, Q=3):
    c = _abi.gc_mcmc_cfg()
    c.algorithm = algo
    c.iteration = iteration
    c.thinning = thinning
    c.n_burnin = burnin
    c.repetition = rep
    c.n_walker = W
    c.ensemble_update = _abi.GC_ENSEMBLE_REDBLACK if redblack else _abi.GC_ENSEMBLE_SIMULTANEOUS
    c.rng_mode = _abi.GC_RNG_REPLAY if replay else _abi.GC_RNG_NATIVE
    c.noise_slot = Q - 1
    c.n_var = Q
    c.seed = seed
    c.step_size = a
    c.start_range = start_range
    for q, s in enumerate(steps):
        c.x_step[q] = s
    return c


def replay_loss_stopping(trace, f):
    """utils/askadam.m:512-671 restated on a loss trace (loss signals only):
    the iteration at which the reference loop would stop.  `f` is a fitting
    dict, or a gc_askadam_cfg."""
    if isinstance(f, _abi.gc_askadam_cfg):
        f = dict(convergenceWindow=f.convergence_window, emaDecay=f.ema_decay, convergenceValue=f.convergence_value,
                 patienceConvergence=f.patience_conv, tol=f.tol,
                 convergenceModel="ema" if f.convergence_model == _abi.GC_CONV_EMA else "linear")
    ema = trace[0]
    buf = np.ones(int(f["convergenceWindow"]))
    min_loss = trace[0]
    pc = 0
    for e in range(1, len(trace)):
        loss = trace[e]
        min_loss = min(min_loss, loss)
        if f["convergenceModel"] == "ema":
            en = f["emaDecay"] * ema + (1 - f["emaDecay"]) * loss
            conv = abs(en - ema) / (abs(ema) + 1e-8)
            ema = en
        else:
            buf = np.append(buf[1:], loss)
            x = np.arange(1, buf.size + 1)
            conv = -np.polyfit(x, buf, 1)[0]
        if conv > f["convergenceValue"] or e <= f["convergenceWindow"]:
            pc = 0
        elif (min_loss - loss) > f["convergenceValue"]:
            pc = 0
        else:
            pc += 1
        if pc > f["patienceConvergence"]:
            return e, _abi.GC_STOP_CONVERGENCE
        if loss < f["tol"]:
            return e, _abi.GC_STOP_TOL
    return len(trace) - 1, _abi.GC_STOP_NONE


def matlab_iqr(a):
    """iqr along rows with MATLAB's prctile convention (host reference)."""
    from gacelle_amd import utils

    out = np.empty(a.shape[0])
    for i in range(a.shape[0]):
        out[i] = utils.prctile(a[i], 75) - utils.prctile(a[i], 25)
    return out


def matlab_mode_bins(a, lb, ub, nbin):
    """mode(discretize(x, edges)) with edges = linspace(lb-1e-8, ub+1e-8, Nbin)
    and the bin centre returned (utils/mcmc.m:845-857), host reference."""
    edges = np.linspace(lb - 1e-8, ub + 1e-8, nbin)
    idx = np.clip(np.searchsorted(edges, a, side="right") - 1, 0, nbin - 2)
    out = np.empty(a.shape[0])
    for i in range(a.shape[0]):
        counts = np.bincount(idx[i], minlength=nbin - 1)
        b = int(np.argmax(counts))  # smallest bin on ties, as MATLAB mode
        out[i] = (edges[b] + edges[b + 1]) / 2
    return out


def wglobal_draws(rng, R, it, warm, W, nv, NND, NG, redblack):
    """Replay draws for goodman_weare_wglobal_constant in the C-ABI layouts:
    (normals, uniforms, partners, g_uniforms, g_partners).  Partners are
    derangements (simultaneous) or anchors from the other half (red-black)."""
    T = warm + it
    h, hm = W // 2, W - W // 2
    normals = rng.standard_normal((W, nv, NND)).astype(np.float32)
    uni = rng.uniform(0, 1, (R, it, 2, 2, hm, nv) if redblack else (R, it, 2, W, nv)).astype(np.float32)
    gu = np.concatenate([rng.uniform(0, 1, W * NG),
                         rng.uniform(0, 1, R * T * 2 * (2 * hm if redblack else W))]).astype(np.float32)

    def table(steps):
        if redblack:
            return np.concatenate([rng.integers(h, W, (R, steps, h)), rng.integers(0, h, (R, steps, W - h))],
                                  axis=2).astype(np.int32)
        return np.stack([[np.roll(np.arange(W), 1 + rng.integers(0, W - 1)) for _ in range(steps)]
                         for _ in range(R)]).astype(np.int32)

    return normals, uni, table(it), gu, table(T)
