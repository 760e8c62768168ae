"""goodman_weare_wglobal_constant on the device (gc_mcmc_run with a global
mask) against the oracle, through the C ABI: replayed draws and the native
stream, simultaneous and red-black updates, warm-up, repetitions, and a
grid-integrated posterior; and the mcmc mirror end to end with a one-element
pars0 field."""
import numpy as np
import pytest

from gacelle_amd import _abi
from oracle import oracle as orc

from .helpers import TE12, mcmc_cfg, r2s_setup, wglobal_draws

pytestmark = pytest.mark.gpu


def _problem(nv, seed):
    md = r2s_setup({}, solver="mcmc")[2]
    rng = np.random.default_rng(seed)
    th = np.stack([rng.uniform(0.8, 1.2, nv), rng.uniform(20, 40, nv), np.full(nv, 0.01)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    return md, rng, th.astype(np.float32), y


def _compare(g, o, isg, R, rtol):
    """Decisions identical except after a marginal first divergence (global
    decisions sum FP32 logP over every voxel, so their margin carries that
    sum's rounding); identical chains agree to FP32 rounding."""
    loc, glo = np.where(~isg)[0], np.where(isg)[0]
    nv = g["trace"].shape[-1]
    same_rep = []
    for r in range(R):
        d = g["g_trace"][r] != o["g_trace"][r]
        if d.any():
            t, w = np.argwhere(d)[0]
            assert abs(o["g_margin"][r, t, w]) < 1e-2, (r, t, w, o["g_margin"][r, t, w])
            same_rep.append(False)
        else:
            same_rep.append(True)
    assert sum(same_rep) >= 1
    n_cmp = 0
    for r in np.where(same_rep)[0]:
        dl = (g["trace"][r] != o["trace"][r]).any(axis=(0, 1))  # [nv]
        for v in np.where(dl)[0]:
            dd = g["trace"][r, :, :, v] != o["trace"][r, :, :, v]
            k, w = np.argwhere(dd)[0]
            assert abs(o["margin"][r, k, w, v]) < 1e-3, (r, k, w, v, o["margin"][r, k, w, v])
        ok = ~dl
        assert ok.mean() > 0.9
        np.testing.assert_allclose(g["posterior"][:, r][..., ok], o["posterior"][:, r][..., ok], rtol=rtol, atol=1e-6)
        np.testing.assert_allclose(g["xfinal"][r][..., ok], o["xfinal"][r][..., ok], rtol=rtol, atol=1e-6)
        np.testing.assert_allclose(g["g_posterior"][:, r], o["g_posterior"][:, r], rtol=rtol, atol=1e-7)
        np.testing.assert_allclose(g["g_xfinal"][r], o["g_xfinal"][r], rtol=rtol, atol=1e-7)
        n_cmp += ok.sum()
    assert n_cmp > 0
    if all(same_rep):
        dl = (g["trace"] != o["trace"]).any(axis=(0, 1, 2))
        for q in loc:
            np.testing.assert_allclose(g["mean"][q][~dl], o["mean"][q][~dl], rtol=rtol)
        for q in glo:
            np.testing.assert_allclose(g["mean"][q], o["mean"][q], rtol=rtol)
            np.testing.assert_allclose(g["std"][q], o["std"][q], rtol=5e-3)
        np.testing.assert_allclose(g["accept"][~dl], o["accept"][~dl], rtol=1e-6)


@pytest.mark.parametrize("redblack", [False, True])
@pytest.mark.parametrize("mask", [0b100, 0b110])
def test_replay_matches_oracle(redblack, mask):
    nv, W, R, it, warm = 3000, 6, 2, 30, 3
    md, rng, x0, y = _problem(nv, 100 + mask + redblack)
    cfg = mcmc_cfg(_abi.GC_MCMC_GW, it, 3, 5, rep=R, W=W, replay=True, redblack=redblack, start_range=0.05)
    cfg.global_mask, cfg.global_warmup = mask, warm
    NG = bin(mask).count("1")
    draws = wglobal_draws(rng, R, it, warm, W, nv, 3 - NG, NG, redblack)
    kw = dict(normals=draws[0], uniforms=draws[1], partners=draws[2], g_uniforms=draws[3], g_partners=draws[4])
    print(f"wglobal replay mask={mask:03b} redblack={redblack} nv={nv}", flush=True)
    from . import gpu_abi

    g = gpu_abi.mcmc_global(md, cfg, y, x0, trace=True, post_fill=np.nan, **kw)
    o = orc.mcmc_global(md, cfg, y, x0, trace=True, margin=True, nthreads=8, **kw)
    assert g["n_sample"] == o["n_sample"] and g["n_stored"] == g["n_sample"] - 1
    # the slot thinning leaves unfilled holds zeros (utils/mcmc.m:481-482)
    assert np.all(g["posterior"][:, :, -1] == 0) and np.all(g["g_posterior"][:, :, -1] == 0)
    isg = np.array([(mask >> q) & 1 for q in range(3)], bool)
    # stretch moves x_p + z (x - x_p) carry the partner's FP32 rounding into
    # every walker, compounding over the 33 steps (as for the plain GW fixtures)
    _compare(g, o, isg, R, rtol=5e-4)


@pytest.mark.parametrize("redblack", [False, True])
def test_native_stream_matches_oracle(redblack):
    nv, W, R, it, warm = 500, 8, 2, 40, 4
    md, rng, x0, y = _problem(nv, 7 + redblack)
    cfg = mcmc_cfg(_abi.GC_MCMC_GW, it, 2, 4, rep=R, W=W, redblack=redblack, start_range=0.05, seed=1234)
    cfg.global_mask, cfg.global_warmup = 0b100, warm
    print(f"wglobal native redblack={redblack}", flush=True)
    from . import gpu_abi

    g = gpu_abi.mcmc_global(md, cfg, y, x0, trace=True)
    o = orc.mcmc_global(md, cfg, y, x0, trace=True, margin=True, nthreads=8)
    _compare(g, o, np.array([False, False, True]), R, rtol=5e-4)


def test_two_globals_match_grid_posterior():
    """The oracle test's grid posterior, sampled on the device."""
    from gacelle_amd import models
    from gacelle_amd.mcmc import mcmc as M

    from . import gpu_abi

    rng = np.random.default_rng(5)
    nv, sig = 6, 0.02
    y = (np.exp(-np.outer(TE12, np.full(nv, 30.0))) + sig * rng.standard_normal((12, nv))).astype(np.float32)
    f = M.check_set_default_basic(dict(modelParams=["M0", "R2star", "noise"], lb=[0.9995, 10, 0.005],
                                       ub=[1.0005, 50, 0.08]))
    md = models.gpuR2starMapping(TE12).gc_model(f, "mcmc", 12)
    r2 = np.linspace(10, 50, 1601)
    s = np.linspace(0.005, 0.08, 1501)
    S = ((y.astype(np.float64)[None] - np.exp(-TE12[None, :, None] * r2[:, None, None])) ** 2).sum(axis=(1, 2))
    lp = -S[:, None] / (2 * s[None] ** 2) - 12 * nv * np.log(s[None])
    p = np.exp(lp - lp.max())
    p /= p.sum()
    cfg = mcmc_cfg(_abi.GC_MCMC_GW, 6000, 1, 1000, rep=1, W=24, seed=3)
    cfg.global_mask = 0b110
    x0 = np.stack([np.ones(nv), np.full(nv, 30.0), np.full(nv, sig)]).astype(np.float32)
    print("wglobal grid posterior", flush=True)
    g = gpu_abi.mcmc_global(md, cfg, y, x0)
    for i, (marg, v) in enumerate(((p.sum(1), r2), (p.sum(0), s))):
        mu = (marg * v).sum()
        sd = np.sqrt((marg * (v - mu) ** 2).sum())
        smp = g["g_posterior"][i]
        assert abs(smp.mean() - mu) < 0.08 * sd, (i, smp.mean(), mu, sd)
        assert smp.std() == pytest.approx(sd, rel=0.06), i


def test_mirror_optimisation_with_a_shared_noise():
    """mcmc().optimisation with pars0.noise a scalar dispatches to
    goodman_weare_wglobal_constant (utils/mcmc.m:54-72); global metrics are
    scalars, local ones images."""
    from gacelle_amd import models
    from gacelle_amd.mcmc import mcmc as M

    rng = np.random.default_rng(4)
    mask = np.zeros((8, 6, 2), bool)
    mask[1:7, 1:5, :] = True
    nv = int(mask.sum())
    s0 = rng.uniform(0.8, 1.2, mask.shape)
    r2 = rng.uniform(20, 40, mask.shape)
    img = s0[..., None] * np.exp(-r2[..., None] * TE12) + 0.01 * rng.standard_normal(mask.shape + (12,))
    fitting = dict(modelParams=["M0", "R2star", "noise"], lb=[0, 0.1, 0.001], ub=[2, 200, 0.1], algorithm="GW",
                   iteration=400, thinning=2, burnin=0.5, Nwalker=10, metric=["mean", "std", "median", "iqr", "mode"],
                   globalWarmup=5, verbose=False)
    pars0 = dict(M0=s0.astype(np.float32), R2star=r2.astype(np.float32), noise=np.float32(0.02))
    print("wglobal mirror", flush=True)
    out = M().optimisation(img.astype(np.float32), mask, None, pars0, fitting,
                           models.gpuR2starMapping(TE12))
    assert out["posterior"]["M0"].shape == (nv, 10, 100, 1)
    assert out["posterior"]["noise"].shape == (1, 10, 100, 1)
    for m in ("mean", "std", "median", "iqr", "mode"):
        assert np.ndim(out[m]["noise"]) == 0 and out[m]["M0"].shape == mask.shape
    post = out["posterior"]["noise"].reshape(-1).astype(np.float64)
    assert float(out["mean"]["noise"]) == pytest.approx(post.mean(), rel=1e-5)
    assert float(out["median"]["noise"]) == pytest.approx(np.median(post), rel=1e-3)
    assert 0.005 < float(out["mean"]["noise"]) < 0.02
    m0 = out["mean"]["M0"][mask]
    np.testing.assert_allclose(m0, s0[mask], rtol=0.05)
