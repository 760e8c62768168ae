"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP library's parity tests
and the CPU-path baseline of bench.py.  The product (gacelle_amd) never
imports this module.  Struct layouts are those of include/gacelle.h.

Functions here and in oracle.cpp / models.hpp restate the reference
(GACELLE, MATLAB) line by line; each C++ function cites its source.
Parity notes:
  * pinned by closed forms / known answers / brute force / finite
    differences in tests/test_oracle.py: every forward model, the dual-
    number gradients, TV and its adjoint, logP_Gaussian, Philox4x32-10,
    the Adam/SGDM/RMSProp steps, the stopping signals, MH/GW invariants
    and posterior statistics of an analytic target.
  * parity unpinned (MATLAB built-ins that cannot run here): adamupdate's
    epsilon placement, l1loss/l2loss/huber/mse weighting, dlgradient
    subgradients at kinks, MATLAB's rng/gpurng streams, prctile/iqr/mode.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int32, c_int64, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_philox4x32_10.argtypes = [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)]
        L.orc_native_partners.argtypes = [c_uint64, c_int32, c_int32, c_int32, c_int32, POINTER(c_int32)]
        L.orc_native_normal.argtypes = [c_uint64, c_uint32, c_uint64, c_uint32, c_int32]
        L.orc_native_normal.restype = c_double
        L.orc_logp_gaussian.argtypes = [c_double, c_double, c_int32]
        L.orc_logp_gaussian.restype = c_double
        L.orc_fwd.argtypes = [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]
        L.orc_fwd_value.argtypes = [c_void_p, c_int64, c_void_p, c_void_p, c_void_p]
        L.orc_tv.argtypes = [c_void_p, c_void_p, POINTER(c_int32), c_int32, POINTER(c_double), c_double, c_int64,
                             POINTER(c_double), c_void_p]
        L.orc_askadam.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                  POINTER(c_int32), c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_double), c_int32, c_int32,
                                  c_void_p, c_int64]
        L.orc_epgx_sample.argtypes = [c_void_p, c_double, c_double, c_double, c_double, c_double,
                                      POINTER(c_double)]
        L.orc_surface_tv.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int32, c_double, c_int64,
                                     POINTER(c_double), c_void_p]
        L.orc_mcmc.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_int32, c_int32]
        L.orc_native_partners_global.argtypes = [c_uint64, c_int32, c_int32, c_int32, c_int32, POINTER(c_int32)]
        L.orc_norm_inv.argtypes = [c_double]
        L.orc_norm_inv.restype = c_double
        L.orc_native_mh_draws.argtypes = [c_uint64, c_uint32, c_uint64, c_int32, POINTER(c_double), POINTER(c_double)]
        L.orc_rescale_parameters.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, c_int64]
        L.orc_unscale_parameters.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, c_int64]
        L.orc_native_gw_uniforms.argtypes = [c_uint64, c_uint32, c_uint64, c_int32, POINTER(c_double),
                                             POINTER(c_double)]
        L.orc_dict_search.argtypes = [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                      c_int32]
        L.orc_mcmc_global.argtypes = [c_void_p, c_void_p, c_int64] + [c_void_p] * 20 + [c_int32, c_int32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def philox(ctr, key):
    c = (c_uint32 * 4)(*ctr)
    k = (c_uint32 * 2)(*key)
    o = (c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def native_partners(seed, rep, it, W, mode):
    out = (c_int32 * W)()
    lib().orc_native_partners(seed, rep, it, W, mode, out)
    return np.array(out, dtype=np.int32)


def native_partners_global(seed, rep, it, W, mode):
    out = (c_int32 * W)()
    lib().orc_native_partners_global(seed, rep, it, W, mode, out)
    return np.array(out, dtype=np.int32)


def norm_inv(u):
    return lib().orc_norm_inv(float(u))


def dict_search(y, dict_, w=None, threads=1):
    """gpuNEXI.likelihood's search (NEXI/gpuNEXI.m:822-849) in FP64:
    y [m, nv], dict_ [m, nd] (float32 values), w [m] or None.  Returns
    (best [nv] int32, 0-based, -1 when no entry scores; sse [nv])."""
    y = np.ascontiguousarray(y, np.float32)
    d = np.ascontiguousarray(dict_, np.float32)
    m, nv = y.shape
    assert d.shape[0] == m and m <= 64
    best = np.empty(nv, np.int32)
    sse = np.empty(nv, np.float64)
    wv = None if w is None else np.ascontiguousarray(w, np.float64)
    rc = lib().orc_dict_search(_p(y), nv, m, _p(d), d.shape[1], _p(wv), _p(best), _p(sse), int(threads))
    assert rc == 0
    return best, sse


def native_mh_draws(seed, it, chain, Q):
    """The MH native stream at (iteration, chain): (normals [Q], acceptance uniform)."""
    n = (c_double * Q)()
    u = c_double()
    assert lib().orc_native_mh_draws(seed, it, chain, Q, n, ctypes.byref(u)) == 0
    return np.array(n), u.value


def rescale_parameters(md, transform, theta):
    """askadam.rescale_parameters (utils/askadam.m:1077-1089), FP64: theta [P, nv] -> z."""
    th = np.ascontiguousarray(theta, np.float64)
    z = np.zeros_like(th)
    assert lib().orc_rescale_parameters(ctypes.byref(md), transform, _p(th), _p(z), th.shape[1]) == 0
    return z


def unscale_parameters(md, transform, z):
    """askadam.unscale_parameters (utils/askadam.m:1063-1073), FP64: z [P, nv] -> theta."""
    zz = np.ascontiguousarray(z, np.float64)
    th = np.zeros_like(zz)
    assert lib().orc_unscale_parameters(ctypes.byref(md), transform, _p(zz), _p(th), zz.shape[1]) == 0
    return th


def native_gw_uniforms(seed, it, chain, sub=0):
    """The Goodman-Weare native stream at (iteration, walker chain, half):
    (stretch uniform, acceptance uniform)."""
    uz, ua = c_double(), c_double()
    assert lib().orc_native_gw_uniforms(seed, it, chain, sub, ctypes.byref(uz), ctypes.byref(ua)) == 0
    return uz.value, ua.value


def native_normal(seed, it, chain, base, q):
    return lib().orc_native_normal(seed, it, chain, base, q)


def logp_gaussian(sos, noise, nm):
    return lib().orc_logp_gaussian(sos, noise, nm)


def fwd(md, theta, extra=None, jac=True):
    """s [M, nv] (and ds/dtheta [M, P, nv]) in FP64."""
    theta = _f64(theta)
    P, nv = theta.shape
    extra = _f64(extra) if extra is not None else np.zeros((1, nv))
    s = np.zeros((md.n_meas, nv))
    j = np.zeros((md.n_meas, P, nv)) if jac else None
    rc = lib().orc_fwd(ctypes.byref(md), nv, _p(theta), _p(extra), _p(s), _p(j))
    assert rc == 0, rc
    return (s, j) if jac else s


def fwd_value(md, theta, extra=None):
    theta = _f64(theta)
    P, nv = theta.shape
    extra = _f64(extra) if extra is not None else np.zeros((1, nv))
    s = np.zeros((md.n_meas, nv))
    rc = lib().orc_fwd_value(ctypes.byref(md), nv, _p(theta), _p(extra), _p(s))
    assert rc == 0, rc
    return s


def tv(img, mask, mode, vsize, lam, nmasked, grad=True):
    img = np.asfortranarray(img, dtype=np.float64)
    mask = np.asfortranarray(mask, dtype=np.uint8)
    dims = (c_int32 * 3)(*[int(s) for s in (list(img.shape) + [1, 1, 1])[:3]])
    vs = (c_double * 3)(*(list(vsize) + [1, 1, 1])[:3])
    loss = c_double()
    g = np.zeros(img.shape, dtype=np.float64, order="F") if grad else None
    lib().orc_tv(img.ctypes.data_as(c_void_p), mask.ctypes.data_as(c_void_p), dims, mode, vs, lam, nmasked,
                 ctypes.byref(loss), None if g is None else g.ctypes.data_as(c_void_p))
    return (loss.value, g) if grad else loss.value


def epgx_sample(md, mvf, r1iew, kfr, r2siw, fa):
    """The EPG-X outputs of one sample and their derivatives (15 values)."""
    out = (c_double * 15)()
    lib().orc_epgx_sample(ctypes.byref(md), mvf, r1iew, kfr, r2siw, fa, out)
    return np.array(out[:])


def surface_tv(img, mask, nbr, elen, lam, nmasked, grad=True):
    """surface_total_variation on a vertex image [ngrid] with global 0-based
    neighbours nbr [max_nb, ngrid] and edge lengths elen [max_nb, ngrid]."""
    img = np.ascontiguousarray(img, dtype=np.float64).ravel()
    mask = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    elen = np.ascontiguousarray(elen, dtype=np.float32)
    loss = c_double()
    g = np.zeros(img.size) if grad else None
    rc = lib().orc_surface_tv(_p(img), _p(mask), img.size, _p(nbr), _p(elen), nbr.shape[0], lam, nmasked,
                              ctypes.byref(loss), _p(g))
    assert rc == 0, rc
    return (loss.value, g) if grad else loss.value


def askadam(md, cfg, y, z0, w=None, extra=None, vox=None, dims=None, precision=0, nthreads=1, trace=False,
            prior_grid=None, surf_nbr=None, surf_len=None, sign_trace=None, n_mean=0):
    """prior_grid: GC_REG_PRIOR maps on the whole grid (column-major),
    [2*n_reg, ngrid]: the mu rows, then the sigma rows; with vox and dims.
    surf_nbr / surf_len: GC_REG_SURFACE_TV neighbours (global 0-based vertex
    indices) and edge lengths, [max_nb, ngrid], with vox and dims.
    sign_trace: uint32 [iteration, ceil(M/16), nv] residual sign codes (two
    bits per measurement: 1 = +, 2 = -, 0 = no gradient) that the L1
    derivative follows instead of the oracle's own signs.
    n_mean: N of the loss mean (0 = nv)."""
    y = _f32(y)
    M, nv = y.shape
    z0 = _f32(z0)
    P = z0.shape[0]
    w = _f32(w)
    extra = _f32(extra)
    vox = None if vox is None else np.ascontiguousarray(vox, dtype=np.int64)
    d = (c_int32 * 3)(*(list(dims) if dims is not None else [1, 1, 1]))
    zf = np.zeros((P, nv))
    zm = np.zeros((P, nv))
    rf = np.zeros((M, nv))
    rm = np.zeros((M, nv))
    tr = np.zeros(int(cfg.iteration) + 1) if trace else None
    sc = (c_double * 10)()
    pg = None if prior_grid is None else np.ascontiguousarray(prior_grid, np.float32)
    sn = None if surf_nbr is None else np.ascontiguousarray(surf_nbr, np.int32)
    sl = None if surf_len is None else np.ascontiguousarray(surf_len, np.float32)
    st = None
    if sign_trace is not None:
        st = np.ascontiguousarray(sign_trace, np.uint32)
        assert st.shape == (int(cfg.iteration), (M + 15) // 16, nv), st.shape
    rc = lib().orc_askadam(ctypes.byref(md), ctypes.byref(cfg), nv, _p(y), _p(w), _p(extra), _p(vox), d, _p(pg),
                           _p(sn), _p(sl), 0 if sn is None else sn.shape[0], _p(z0),
                           _p(zf), _p(zm), _p(rf), _p(rm), _p(tr), sc, precision, nthreads,
                           _p(st), int(n_mean))
    assert rc == 0, rc
    out = dict(z_final=zf, z_min=zm, residual=rf, residual_min=rm, loss=sc[0], loss_fidelity=sc[1],
               loss_reg=sc[2], min_loss=sc[3], min_loss_fidelity=sc[4], min_loss_reg=sc[5], convergence=sc[6],
               n_iteration=int(sc[7]), min_iteration=int(sc[8]), stop_reason=int(sc[9]))
    if trace:
        out["trace"] = tr[: out["n_iteration"] + 1]
    return out


def mcmc(md, cfg, y, x0, w=None, extra=None, normals=None, uniforms=None, partners=None, store=True,
         trace=False, margin=False, precision=0, nthreads=1):
    y = _f32(y)
    M, nv = y.shape
    x0 = _f32(x0)
    Q = cfg.n_var
    R = cfg.repetition
    W = cfg.n_walker if cfg.algorithm == 1 else 1
    it = cfg.iteration
    ns = 0 if it < cfg.n_burnin + 1 else (it - cfg.n_burnin - 1) // cfg.thinning + 1
    post = np.zeros((Q, R, ns, W, nv)) if store else None
    mean = np.zeros((Q, nv))
    std = np.zeros((Q, nv))
    acc = np.zeros(nv)
    xf = np.zeros((R, W, Q, nv))
    tr = np.zeros((R, it, W, nv), dtype=np.uint8) if trace else None
    mg = np.zeros((R, it, W, nv)) if margin else None
    rc = lib().orc_mcmc(ctypes.byref(md), ctypes.byref(cfg), nv, _p(y), _p(_f32(w)), _p(_f32(extra)), _p(x0),
                        _p(_f32(normals)), _p(_f32(uniforms)),
                        None if partners is None else np.ascontiguousarray(partners, np.int32).ctypes.data_as(c_void_p),
                        _p(post), _p(mean), _p(std), _p(acc), _p(xf), _p(tr), _p(mg), precision, nthreads)
    assert rc == 0, rc
    return dict(posterior=post, mean=mean, std=std, accept=acc, xfinal=xf, trace=tr, margin=mg, n_sample=ns)


def mcmc_global(md, cfg, y, x0, w=None, extra=None, normals=None, uniforms=None, partners=None, g_uniforms=None,
                g_partners=None, store=True, trace=False, margin=False, precision=0, nthreads=1):
    """goodman_weare_wglobal_constant; cfg.global_mask marks the global
    variables.  Local outputs cover the NND local variables in model order."""
    y = _f32(y)
    M, nv = y.shape
    x0 = _f32(x0)
    Q = cfg.n_var
    R = cfg.repetition
    W = cfg.n_walker
    it = cfg.iteration
    T = cfg.global_warmup + it
    NG = bin(cfg.global_mask & ((1 << Q) - 1)).count("1")
    NND = Q - NG
    ns = 0 if it < cfg.n_burnin + 1 else (it - cfg.n_burnin - 1) // cfg.thinning + 1
    post = np.zeros((NND, R, ns, W, nv)) if store else None
    gpost = np.zeros((NG, R, ns, W)) if store else None
    mean = np.zeros((Q, nv))
    std = np.zeros((Q, nv))
    acc = np.zeros(nv)
    xf = np.zeros((R, W, NND, nv))
    gxf = np.zeros((R, W, NG))
    tr = np.zeros((R, it, W, nv), dtype=np.uint8) if trace else None
    gtr = np.zeros((R, T, W), dtype=np.uint8) if trace else None
    mg = np.zeros((R, it, W, nv)) if margin else None
    gmg = np.zeros((R, T, W)) if margin else None

    def ip(a):
        return None if a is None else np.ascontiguousarray(a, np.int32).ctypes.data_as(c_void_p)

    rc = lib().orc_mcmc_global(ctypes.byref(md), ctypes.byref(cfg), nv, _p(y), _p(_f32(w)), _p(_f32(extra)), _p(x0),
                               _p(_f32(normals)), _p(_f32(uniforms)), ip(partners), _p(_f32(g_uniforms)),
                               ip(g_partners), _p(post), _p(gpost), _p(mean), _p(std), _p(acc), _p(xf), _p(gxf),
                               _p(tr), _p(gtr), _p(mg), _p(gmg), precision, nthreads)
    assert rc == 0, rc
    return dict(posterior=post, g_posterior=gpost, mean=mean, std=std, accept=acc, xfinal=xf, g_xfinal=gxf,
                trace=tr, g_trace=gtr, margin=mg, g_margin=gmg, n_sample=ns)
