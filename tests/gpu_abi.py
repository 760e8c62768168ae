"""Direct C-ABI drivers used by the GPU parity tests (device buffers from
torch, every computation in libgacelle_hip.so)."""
from __future__ import annotations

import ctypes

import numpy as np

from gacelle_amd import _abi, runtime


def ctx():
    return runtime.context()


def fwd(md, theta, extra=None):
    c = ctx()
    theta = np.ascontiguousarray(theta, np.float32)
    nv = theta.shape[1]
    th = c.to_device(theta)
    ex = c.to_device(extra) if extra is not None else None
    s = c.empty((md.n_meas, nv))
    c.check(c.lib.gc_fwd(c.handle, ctypes.byref(md), runtime.ptr(th), runtime.ptr(ex), runtime.ptr(s), nv))
    return runtime.to_host(s)


def fwd_adj(md, loss, transform, x, y, w=None, extra=None, scale=1.0, residual=False):
    c = ctx()
    x = np.ascontiguousarray(x, np.float32)
    P, nv = x.shape
    xd = c.to_device(x)
    yd = c.to_device(y)
    wd = c.to_device(w) if w is not None else None
    ex = c.to_device(extra) if extra is not None else None
    g = c.empty((P, nv))
    ls = c.zeros((1,), np.float64)
    rs = c.empty((md.n_meas, nv)) if residual else None
    c.check(c.lib.gc_fwd_adj(c.handle, ctypes.byref(md), loss, transform, runtime.ptr(xd), runtime.ptr(yd),
                             runtime.ptr(wd), runtime.ptr(ex), nv, scale, runtime.ptr(g), runtime.ptr(ls),
                             runtime.ptr(rs)))
    out = dict(grad=runtime.to_host(g), loss_sum=float(runtime.to_host(ls)[0]))
    if residual:
        out["residual"] = runtime.to_host(rs)
    return out


def dict_search(y, d, w=None, with_sse=True):
    """gc_dict_search: y [m, nv], d [m, nd] (float32), w [m] host or None."""
    c = ctx()
    y = np.ascontiguousarray(y, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    m, nv = y.shape
    yd, dd = c.to_device(y), c.to_device(d)
    best = c.empty((max(nv, 1),), np.int32)
    sse = c.empty((max(nv, 1),), np.float64) if with_sse else None
    wp = None if w is None else (ctypes.c_float * m)(*np.asarray(w, np.float32))
    c.check(c.lib.gc_dict_search(c.handle, runtime.ptr(yd), nv, m, runtime.ptr(dd), d.shape[1], wp,
                                 runtime.ptr(best), runtime.ptr(sse)))
    return runtime.to_host(best)[:nv], (runtime.to_host(sse)[:nv] if with_sse else None)


def askadam(md, cfg, vols, trace=False, signs=False):
    """vols: list of dict(y, z0, w=None, extra=None, vox=None, dims=None,
    prior=None, reg_offset=0, surf_nbr=None, surf_len=None, n_mean=0).
    Returns per-volume dicts; signs=True adds each volume's residual sign
    trace ("signs", uint32 [iteration, ceil(M/16), nv])."""
    c = ctx()
    n = len(vols)
    arr = (_abi.gc_volume * n)()
    keep, metas = [], []
    for i, v in enumerate(vols):
        y = np.ascontiguousarray(v["y"], np.float32)
        M, nv = y.shape
        z0 = np.ascontiguousarray(v["z0"], np.float32)
        P = z0.shape[0]
        yd = c.to_device(y)
        wd = c.to_device(v["w"]) if v.get("w") is not None else None
        ex = c.to_device(v["extra"]) if v.get("extra") is not None else None
        zd = c.to_device(z0)
        zm = c.empty((P, nv))
        rl = c.empty((nv,))
        rlm = c.empty((nv,))
        rs = c.empty((M, nv))
        rsm = c.empty((M, nv))
        tr = (ctypes.c_double * (int(cfg.iteration) + 1))() if trace else None
        sg = c.zeros((int(cfg.iteration), (M + 15) // 16, nv), np.int32) if signs else None
        a = arr[i]
        a.n_mean = int(v.get("n_mean", 0))
        a.sign_trace_dev = runtime.ptr(sg)
        a.nv = nv
        a.y_dev = runtime.ptr(yd)
        a.w_dev = runtime.ptr(wd)
        a.extra_dev = runtime.ptr(ex)
        a.z_dev = runtime.ptr(zd)
        a.zmin_dev = runtime.ptr(zm)
        a.resloss_dev = runtime.ptr(rl)
        a.resloss_min_dev = runtime.ptr(rlm)
        a.residual_dev = runtime.ptr(rs)
        a.residual_min_dev = runtime.ptr(rsm)
        a.loss_trace_host = ctypes.cast(tr, ctypes.POINTER(ctypes.c_double)) if tr is not None else None
        if v.get("vox") is not None:
            vd = c.to_device(np.asarray(v["vox"], np.int64), np.int64)
            a.vox_dev = runtime.ptr(vd)
            for k in range(3):
                a.dims[k] = int(v["dims"][k])
            keep.append(vd)
        if v.get("prior") is not None:   # [2*n_reg, nv] mu rows then sigma rows, and the out-of-mask constant
            pd = c.to_device(np.asarray(v["prior"], np.float32))
            a.prior_dev = runtime.ptr(pd)
            a.reg_offset = float(v.get("reg_offset", 0.0))
            keep.append(pd)
        if v.get("surf_nbr") is not None:   # [max_nb, ngrid] global vertex indices and edge lengths
            sn = c.to_device(np.asarray(v["surf_nbr"], np.int32), np.int32)
            sl = c.to_device(np.asarray(v["surf_len"], np.float32))
            a.surf_nbr_dev, a.surf_len_dev = runtime.ptr(sn), runtime.ptr(sl)
            a.surf_max_nb = int(np.shape(v["surf_nbr"])[0])
            keep += [sn, sl]
        keep += [yd, wd, ex, zd, zm, rl, rlm, rs, rsm, sg]
        metas.append((zd, zm, rl, rlm, rs, rsm, tr, sg))
    c.check(c.lib.gc_askadam_run(c.handle, ctypes.byref(md), ctypes.byref(cfg), arr, n))
    outs = []
    for i, (zd, zm, rl, rlm, rs, rsm, tr, sg) in enumerate(metas):
        a = arr[i]
        o = dict(z_final=runtime.to_host(zd), z_min=runtime.to_host(zm), resloss=runtime.to_host(rl),
                 resloss_min=runtime.to_host(rlm), residual=runtime.to_host(rs), residual_min=runtime.to_host(rsm),
                 loss=a.loss, loss_fidelity=a.loss_fidelity, loss_reg=a.loss_reg, min_loss=a.min_loss,
                 min_iteration=a.min_iteration, n_iteration=a.n_iteration, stop_reason=a.stop_reason,
                 convergence=a.convergence, n_launch=a.n_launch, block_threads=a.block_threads,
                 iters_per_launch=a.iters_per_launch)
        if tr is not None:
            o["trace"] = np.array(tr[: a.n_iteration + 1])
        if sg is not None:
            o["signs"] = runtime.to_host(sg).view(np.uint32)
        outs.append(o)
    return outs


def _prefilled(c, shape, fill):
    t = c.empty(shape)
    if fill is not None:
        t.fill_(fill)
    return t


def mcmc(md, cfg, y, x0, w=None, extra=None, normals=None, uniforms=None, partners=None, store=True, trace=False,
         post_fill=None):
    c = ctx()
    y = np.ascontiguousarray(y, np.float32)
    M, nv = y.shape
    Q, R = cfg.n_var, cfg.repetition
    W = cfg.n_walker if cfg.algorithm == _abi.GC_MCMC_GW else 1
    it = cfg.iteration
    ns = 0 if it < cfg.n_burnin + 1 else (it - cfg.n_burnin - 1) // cfg.thinning + 1
    io = _abi.gc_mcmc_io()
    yd = c.to_device(y)
    x0d = c.to_device(np.ascontiguousarray(x0, np.float32))
    wd = c.to_device(w) if w is not None else None
    ex = c.to_device(extra) if extra is not None else None
    nd = c.to_device(normals) if normals is not None else None
    ud = c.to_device(uniforms) if uniforms is not None else None
    pd = c.to_device(np.asarray(partners, np.int32), np.int32) if partners is not None else None
    post = _prefilled(c, (Q, R, ns, W, nv), post_fill) if store and ns > 0 else None
    mean = c.empty((Q, nv))
    std = c.empty((Q, nv))
    acc = c.empty((nv,))
    xf = c.empty((R, W, Q, nv))
    tr = c.zeros((R, it, W, nv), np.uint8) if trace else None
    io.nv = nv
    for name, t in (("y_dev", yd), ("w_dev", wd), ("extra_dev", ex), ("x0_dev", x0d), ("normals_dev", nd),
                    ("uniforms_dev", ud), ("partners_dev", pd), ("posterior_dev", post), ("mean_dev", mean),
                    ("std_dev", std), ("accept_dev", acc), ("xfinal_dev", xf), ("trace_dev", tr)):
        setattr(io, name, runtime.ptr(t))
    c.check(c.lib.gc_mcmc_run(c.handle, ctypes.byref(md), ctypes.byref(cfg), ctypes.byref(io)))
    return dict(posterior=None if post is None else runtime.to_host(post), mean=runtime.to_host(mean),
                std=runtime.to_host(std), accept=runtime.to_host(acc), xfinal=runtime.to_host(xf),
                trace=None if tr is None else runtime.to_host(tr), n_sample=io.n_sample, n_stored=io.n_stored,
                block_threads=io.block_threads, kernel_ms=io.kernel_ms)


def mcmc_global(md, cfg, y, x0, w=None, extra=None, normals=None, uniforms=None, partners=None, g_uniforms=None,
                g_partners=None, store=True, trace=False, post_fill=None):
    """gc_mcmc_run with cfg.global_mask set (goodman_weare_wglobal_constant)."""
    c = ctx()
    y = np.ascontiguousarray(y, np.float32)
    M, nv = y.shape
    Q, R, W = cfg.n_var, cfg.repetition, cfg.n_walker
    NG = bin(cfg.global_mask & ((1 << Q) - 1)).count("1")
    NL = Q - NG
    it = cfg.iteration
    T = cfg.global_warmup + it
    ns = 0 if it < cfg.n_burnin + 1 else (it - cfg.n_burnin - 1) // cfg.thinning + 1
    io = _abi.gc_mcmc_io()
    dev = lambda a, dt=np.float32: None if a is None else c.to_device(np.ascontiguousarray(a, dt), dt)
    bufs = dict(y_dev=dev(y), x0_dev=dev(x0), w_dev=dev(w), extra_dev=dev(extra), normals_dev=dev(normals),
                uniforms_dev=dev(uniforms), partners_dev=dev(partners, np.int32), g_uniforms_dev=dev(g_uniforms),
                g_partners_dev=dev(g_partners, np.int32),
                posterior_dev=_prefilled(c, (NL, R, ns, W, nv), post_fill) if store and ns > 0 else None,
                g_posterior_dev=_prefilled(c, (NG, R, ns, W), post_fill) if store and ns > 0 else None,
                mean_dev=c.empty((Q, nv)), std_dev=c.empty((Q, nv)), accept_dev=c.empty((nv,)),
                xfinal_dev=c.empty((R, W, NL, nv)), g_xfinal_dev=c.empty((R, W, NG)),
                trace_dev=c.zeros((R, it, W, nv), np.uint8) if trace else None,
                g_trace_dev=c.zeros((R, T, W), np.uint8) if trace else None)
    io.nv = nv
    for name, t in bufs.items():
        setattr(io, name, runtime.ptr(t))
    c.check(c.lib.gc_mcmc_run(c.handle, ctypes.byref(md), ctypes.byref(cfg), ctypes.byref(io)))
    h = lambda k: None if bufs[k] is None else runtime.to_host(bufs[k])
    return dict(posterior=h("posterior_dev"), g_posterior=h("g_posterior_dev"), mean=h("mean_dev"), std=h("std_dev"),
                accept=h("accept_dev"), xfinal=h("xfinal_dev"), g_xfinal=h("g_xfinal_dev"), trace=h("trace_dev"),
                g_trace=h("g_trace_dev"), n_sample=io.n_sample, n_stored=io.n_stored, kernel_ms=io.kernel_ms)


def tv(img, mask, mode, vsize, lam, nmasked):
    """gc_tv_fwd_adj on a column-major [x, y, z] grid: (loss, grad)."""
    c = ctx()
    img = np.asfortranarray(img, np.float32)
    dims3 = (list(img.shape) + [1, 1, 1])[:3]
    imgd = c.to_device(img.ravel(order="F"))
    maskd = c.to_device(np.asfortranarray(mask, np.uint8).ravel(order="F"), np.uint8)
    g = c.zeros((img.size,))
    dims = (ctypes.c_int32 * 3)(*[int(d) for d in dims3])
    vs = (ctypes.c_double * 3)(*[float(v) for v in vsize])
    loss = ctypes.c_double()
    c.check(c.lib.gc_tv_fwd_adj(c.handle, runtime.ptr(imgd), runtime.ptr(maskd), dims, mode, vs, float(lam),
                                int(nmasked), ctypes.byref(loss), runtime.ptr(g)))
    return loss.value, runtime.to_host(g).reshape(dims3, order="F")


def rescale(md, transform, theta):
    """gc_rescale_parameters: theta [P, nv] -> z."""
    c = ctx()
    th = c.to_device(np.ascontiguousarray(theta, np.float32))
    z = c.empty(th.shape)
    c.check(c.lib.gc_rescale_parameters(c.handle, ctypes.byref(md), transform, runtime.ptr(th), runtime.ptr(z),
                                        th.shape[1]))
    return runtime.to_host(z)


def unscale(md, transform, z):
    """gc_unscale_parameters: z [P, nv] -> theta."""
    c = ctx()
    zd = c.to_device(np.ascontiguousarray(z, np.float32))
    th = c.empty(zd.shape)
    c.check(c.lib.gc_unscale_parameters(c.handle, ctypes.byref(md), transform, runtime.ptr(zd), runtime.ptr(th),
                                        zd.shape[1]))
    return runtime.to_host(th)
