"""goodman_weare_wglobal_constant and gw_global_block (utils/mcmc.m:385-626)
written over whole arrays in the reference's own shapes -- locals
[NND, Nv, W], globals [NG, 1, W], logP [1, Nv, W] -- replaying caller
draws.  A second restatement, structured like the reference's array program
and independent of the oracle's per-walker loops, used to pin the oracle's
block order, draw layouts and bookkeeping.  Weights are ones."""
from __future__ import annotations

import numpy as np


def _min1(x):
    # MATLAB min(x, 1) ignores NaN
    return np.where(np.isnan(x), 1.0, np.minimum(x, 1.0))


def _exp(x):
    with np.errstate(over="ignore"):
        return np.exp(x)


def gw_wglobal(fwd, y, x0, isg, lb, ub, noise_q, iteration, thinning, burnin, rep, W, a, start_range, warmup,
               redblack, normals, uniforms, partners, g_uniforms, g_partners):
    """fwd(theta [Q, n]) -> s [M, n].  Draw arrays as in the C ABI:
    normals [W][nv][NND]; uniforms [rep][iter][2][W][nv] (red-black
    [rep][iter][2][2][hm][nv]); partners [rep][iter][W]; g_uniforms
    [W*NG] then [rep][T][2][W] (red-black [rep][T][2][2][hm]); g_partners
    [rep][T][W].  Returns (xPosterior_ND [NND,Nv,W,Ns,rep],
    xPosterior_global [NG,1,W,Ns,rep], local trace, global trace)."""
    y = np.asarray(y, np.float64)
    M, nv = y.shape
    isg = np.asarray(isg, bool)
    Q = isg.size
    loc, glo = np.where(~isg)[0], np.where(isg)[0]
    NND, NG = loc.size, glo.size
    T = warmup + iteration
    h, hm = W // 2, W - W // 2
    Ns = 0 if iteration < burnin + 1 else (iteration - burnin - 1) // thinning + 1
    lb32, ub32 = np.asarray(lb, np.float32), np.asarray(ub, np.float32)
    lbl, ubl = lb32[loc][:, None, None].astype(np.float64), ub32[loc][:, None, None].astype(np.float64)
    lbg, ubg = lb32[glo][:, None, None].astype(np.float64), ub32[glo][:, None, None].astype(np.float64)
    # ub - lb of the single-precision bound arrays (:473-476)
    wl = (ub32 - lb32)[loc][:, None, None].astype(np.float64)
    wg = (ub32 - lb32)[glo][:, None, None].astype(np.float64)

    def logp(xl, g):
        wn = xl.shape[2]
        th = np.empty((Q, nv, wn))
        th[loc] = xl
        th[glo] = np.broadcast_to(g, (NG, nv, wn))
        s = fwd(th.reshape(Q, nv * wn)).reshape(M, nv, wn)
        sos = np.sum((s - y[:, :, None]) ** 2, axis=0)[None]
        noise = th[noise_q][None]
        return -sos / (2 * noise ** 2) - M / 2 * np.log(2 * np.pi * noise ** 2)

    # initial ensemble (:484-492): randn([NND,Nv,W]) and rand([NG,1,W])
    n = np.asarray(normals, np.float64).reshape(W, nv, NND).transpose(2, 1, 0)
    xl0 = np.clip(x0[loc][:, :, None] + wl * start_range * n, lbl, ubl)
    gu = np.asarray(g_uniforms, np.float64)
    g0 = lbg + wg * gu[: W * NG].reshape(W, NG).T[:, None, :]
    gsteps = gu[W * NG:].reshape((rep, T, 2, 2, hm) if redblack else (rep, T, 2, W))
    lu = np.asarray(uniforms, np.float64).reshape((rep, iteration, 2, 2, hm, nv) if redblack
                                                   else (rep, iteration, 2, W, nv))
    lp0 = logp(xl0, g0)

    post = np.zeros((NND, nv, W, Ns, rep))
    gpost = np.zeros((NG, 1, W, Ns, rep))
    ltrace = np.zeros((rep, iteration, W, nv), np.uint8)
    gtrace = np.zeros((rep, T, W), np.uint8)
    halves = [np.arange(h), np.arange(h, W)]
    for r in range(rep):
        lp, x, g = lp0.copy(), xl0.copy(), g0.copy()

        def global_block(t):
            nonlocal g, lp
            p = np.asarray(g_partners[r, t - 1])
            if not redblack:
                u = gsteps[r, t - 1]
                zz = (((a - 1) * u[0] + 1) ** 2 / a)[None, None, :]
                gp = g[:, :, p] + (g - g[:, :, p]) * zz
                oob = np.any((gp < lbg) | (gp > ubg), axis=0)
                gp = np.clip(gp, lbg, ubg)
                lpg = logp(x, gp)
                dlogp = np.sum(lpg - lp, axis=1, keepdims=True)
                acc = (_min1(zz ** (NG - 1) * _exp(dlogp)) > u[1][None, None, :]) & ~oob
                gtrace[r, t - 1] = acc[0, 0]
                g[:, :, acc[0, 0]] = gp[:, :, acc[0, 0]]
                lp[:, :, acc[0, 0]] = lpg[:, :, acc[0, 0]]
            else:
                for s in range(2):
                    act = halves[s]
                    na = act.size
                    pidx = p[(0 if s == 0 else h):][:na]
                    u = gsteps[r, t - 1, s][:, :na]
                    zz = (((a - 1) * u[0] + 1) ** 2 / a)[None, None, :]
                    gp = g[:, :, pidx] + (g[:, :, act] - g[:, :, pidx]) * zz
                    oob = np.any((gp < lbg) | (gp > ubg), axis=0)
                    gp = np.clip(gp, lbg, ubg)
                    lpg = logp(x[:, :, act], gp)
                    dlogp = np.sum(lpg - lp[:, :, act], axis=1, keepdims=True)
                    acc = ((_min1(zz ** (NG - 1) * _exp(dlogp)) > u[1][None, None, :]) & ~oob)[0, 0]
                    gtrace[r, t - 1, act] = acc
                    g[:, :, act[acc]] = gp[:, :, acc]
                    lp[:, :, act[acc]] = lpg[:, :, acc]

        for t in range(1, warmup + 1):
            global_block(t)
        counter = 0
        for k in range(1, iteration + 1):
            p = np.asarray(partners[r, k - 1])
            if not redblack:
                u = lu[r, k - 1]
                zz = (((a - 1) * u[0] + 1) ** 2 / a).T[None]
                xp = x[:, :, p] + (x - x[:, :, p]) * zz
                oob = np.any((xp < lbl) | (xp > ubl), axis=0)
                xp = np.clip(xp, lbl, ubl)
                lpl = logp(xp, g)
                acc = ((_min1(zz ** (NND - 1) * _exp(lpl - lp)) > u[1].T[None]) & ~oob[None])[0]
                ltrace[r, k - 1] = acc.T
                lp[0][acc] = lpl[0][acc]
                x[:, acc] = xp[:, acc]
            else:
                for s in range(2):
                    act = halves[s]
                    na = act.size
                    pidx = p[(0 if s == 0 else h):][:na]
                    u = lu[r, k - 1, s][:, :na]
                    zz = (((a - 1) * u[0] + 1) ** 2 / a).T[None]
                    xa = x[:, :, pidx]
                    xp = xa + (x[:, :, act] - xa) * zz
                    oob = np.any((xp < lbl) | (xp > ubl), axis=0)
                    xp = np.clip(xp, lbl, ubl)
                    lpl = logp(xp, g[:, :, act])
                    lpa = lp[:, :, act]
                    acc = ((_min1(zz ** (NND - 1) * _exp(lpl - lpa)) > u[1].T[None]) & ~oob[None])[0]
                    ltrace[r, k - 1, act] = acc.T
                    xact = x[:, :, act]
                    xact[:, acc] = xp[:, acc]
                    lpa[0][acc] = lpl[0][acc]
                    x[:, :, act] = xact
                    lp[:, :, act] = lpa
            global_block(warmup + k)
            if k > burnin and (k - burnin + 1) % thinning == 0:
                if counter < Ns:
                    post[:, :, :, counter, r] = x
                    gpost[:, :, :, counter, r] = g
                counter += 1
    return post, gpost, ltrace, gtrace
