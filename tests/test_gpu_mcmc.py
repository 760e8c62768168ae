"""GPU parity of the MCMC samplers against the oracle (C ABI, gc_mcmc_run)."""
import numpy as np
import pytest

from gacelle_amd import _abi
from oracle import oracle as orc

from . import gpu_abi
from .helpers import mcmc_cfg, r2s_problem, r2s_setup

pytestmark = pytest.mark.gpu


def _first_divergence_ok(g_trace, o_trace, margin, tol=1e-3):
    """Chains whose accept decisions differ: the first differing decision
    must be a marginal one (|log ratio - log u| < tol in the oracle)."""
    diff = g_trace != o_trace
    flat = diff.reshape(-1, diff.shape[-1]) if diff.ndim > 2 else diff
    R, it = g_trace.shape[0], g_trace.shape[1]
    bad = 0
    n_div = 0
    for r in range(R):
        d = diff[r].reshape(it, -1)
        m = margin[r].reshape(it, -1)
        cols = np.where(d.any(axis=0))[0]
        for c in cols:
            k = np.argmax(d[:, c])
            n_div += 1
            if not abs(m[k, c]) < tol:
                bad += 1
    return n_div, bad


def test_mh_replay_matches_oracle():
    nv = 2000
    y, th, sigma = r2s_problem(nv, seed=31)
    _, f, md = r2s_setup(dict(iteration=500, thinning=5, burnin=0.1), solver="mcmc")
    it, R, Q = 500, 2, 3
    cfg = mcmc_cfg(_abi.GC_MCMC_MH, it, 5, 50, rep=R, replay=True, steps=(0.01, 1, 0.005))
    rng = np.random.default_rng(5)
    normals = rng.standard_normal((R, it, nv, Q)).astype(np.float32)
    uniforms = rng.uniform(0, 1, (R, it, nv)).astype(np.float32)
    x0 = np.stack([th[0], th[1], np.full(nv, sigma)]).astype(np.float32)
    g = gpu_abi.mcmc(md, cfg, y, x0, normals=normals, uniforms=uniforms, trace=True)
    o = orc.mcmc(md, cfg, y, x0, normals=normals, uniforms=uniforms, trace=True, margin=True)
    same = np.all(g["trace"] == o["trace"], axis=(1, 2)).reshape(R, -1)
    frac = same.mean()
    assert frac > 0.97, frac
    n_div, bad = _first_divergence_ok(g["trace"], o["trace"], o["margin"])
    assert bad == 0, (n_div, bad)
    # identical decision sequences -> identical chains within FP32 tolerance
    xs_g = g["xfinal"][:, 0]  # [R, Q, nv]
    xs_o = o["xfinal"][:, 0]
    for r in range(R):
        ok = same[r]
        np.testing.assert_allclose(xs_g[r][:, ok], xs_o[r][:, ok], rtol=1e-5, atol=1e-6)
    post_ok = np.all(same, axis=0)
    np.testing.assert_allclose(g["posterior"][..., post_ok], o["posterior"][..., post_ok], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g["mean"][:, post_ok], o["mean"][:, post_ok], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g["std"][:, post_ok], o["std"][:, post_ok], rtol=1e-3, atol=1e-6)


def test_mh_native_stream_matches_oracle():
    nv = 1000
    y, th, sigma = r2s_problem(nv, seed=32)
    _, f, md = r2s_setup({}, solver="mcmc")
    cfg = mcmc_cfg(_abi.GC_MCMC_MH, 400, 20, 40, rep=4, steps=(0.01, 1, 0.005))
    x0 = np.stack([th[0], th[1], np.full(nv, sigma)]).astype(np.float32)
    g = gpu_abi.mcmc(md, cfg, y, x0, trace=True)
    o = orc.mcmc(md, cfg, y, x0, trace=True, margin=True)
    same = np.all(g["trace"] == o["trace"], axis=(1, 2))
    assert same.mean() > 0.97
    n_div, bad = _first_divergence_ok(g["trace"], o["trace"], o["margin"])
    assert bad == 0, (n_div, bad)


@pytest.mark.parametrize("redblack", [False, True])
def test_gw_replay_matches_oracle(redblack):
    """Same draw stream on device and oracle.  A voxel's walkers interact,
    so one marginal FP32-vs-FP64 decision changes the whole ensemble from
    there on, and stretch moves amplify rounding differences over many
    iterations: over a 20-iteration horizon every first divergence must be
    marginal, and voxels without any divergence must match to FP32
    tolerance.  (Long horizons are compared statistically.)"""
    nv, W, R, it, Q = 300, 16, 2, 20, 3
    y, th, sigma = r2s_problem(nv, seed=33)
    _, f, md = r2s_setup({}, solver="mcmc")
    cfg = mcmc_cfg(_abi.GC_MCMC_GW, it, 2, 4, rep=R, W=W, replay=True, redblack=redblack)
    rng = np.random.default_rng(8)
    normals = rng.standard_normal((W, nv, Q)).astype(np.float32)
    if redblack:
        hm = (W + 1) // 2
        uniforms = rng.uniform(0, 1, (R, it, 2, 2, hm, nv)).astype(np.float32)
        h = W // 2
        partners = np.zeros((R, it, W), np.int32)
        partners[..., :h] = rng.integers(h, W, (R, it, h))
        partners[..., h:] = rng.integers(0, h, (R, it, W - h))
    else:
        uniforms = rng.uniform(0, 1, (R, it, 2, W, nv)).astype(np.float32)
        partners = np.zeros((R, it, W), np.int32)
        for r in range(R):
            for k in range(it):
                while True:
                    p = rng.permutation(W)
                    if np.all(p != np.arange(W)):
                        break
                partners[r, k] = p
    x0 = np.stack([th[0], th[1], np.full(nv, sigma)]).astype(np.float32)
    g = gpu_abi.mcmc(md, cfg, y, x0, normals=normals, uniforms=uniforms, partners=partners, trace=True,
                     post_fill=np.nan)
    o = orc.mcmc(md, cfg, y, x0, normals=normals, uniforms=uniforms, partners=partners, trace=True, margin=True)
    d = g["trace"] != o["trace"]  # [R, it, W, nv]
    for v in np.where(d.any(axis=(0, 1, 2)))[0]:
        for r in range(R):
            dv = d[r, :, :, v]
            if not dv.any():
                continue
            k = int(np.argmax(dv.any(axis=1)))
            ws = np.where(dv[k])[0]
            assert np.all(np.abs(o["margin"][r, k, ws, v]) < 1e-3), (v, r, k, o["margin"][r, k, ws, v])
    same_vox = ~d.any(axis=(0, 1, 2))
    assert same_vox.mean() > 0.5, same_vox.mean()
    np.testing.assert_allclose(g["mean"][:, same_vox], o["mean"][:, same_vox], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g["std"][:, same_vox], o["std"][:, same_vox], rtol=1e-3, atol=1e-6)
    # stretch moves x_p + z (x - x_p) compound float32 rounding along the
    # chain; the per-parameter bar is 1e-4 relative
    np.testing.assert_allclose(g["posterior"][..., same_vox], o["posterior"][..., same_vox], rtol=1e-4, atol=1e-6)


def test_gw_native_stream_matches_oracle():
    """Native Philox stream and host partner tables: same decisions as the
    oracle restatement (up to marginal FP32 flips)."""
    nv, W, R, it = 200, 16, 2, 40
    y, th, sigma = r2s_problem(nv, seed=34)
    _, f, md = r2s_setup({}, solver="mcmc")
    cfg = mcmc_cfg(_abi.GC_MCMC_GW, it, 5, 10, rep=R, W=W)
    x0 = np.stack([th[0], th[1], np.full(nv, sigma)]).astype(np.float32)
    g = gpu_abi.mcmc(md, cfg, y, x0, trace=True)
    o = orc.mcmc(md, cfg, y, x0, trace=True, margin=True)
    same_vox = np.all(g["trace"] == o["trace"], axis=(0, 1, 2))
    assert same_vox.mean() > 0.5, same_vox.mean()
    np.testing.assert_allclose(g["mean"][:, same_vox], o["mean"][:, same_vox], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("algo", ["mh", "gw"])
def test_long_run_posterior_statistics_match_oracle(algo):
    """Long native-stream runs: device and FP64 oracle posteriors agree
    statistically (per-voxel mean differences are within the Monte-Carlo
    error and unbiased; standard deviations agree)."""
    nv = 200
    y, th, sigma = r2s_problem(nv, seed=35)
    _, f, md = r2s_setup({}, solver="mcmc")
    if algo == "mh":
        cfg = mcmc_cfg(_abi.GC_MCMC_MH, 4000, 10, 400, rep=4, steps=(0.01, 1, 0.005))
        n_eff = 4 * 360 / 10.0
    else:
        cfg = mcmc_cfg(_abi.GC_MCMC_GW, 1000, 10, 100, rep=1, W=16)
        n_eff = 16 * 90 / 5.0
    x0 = np.stack([th[0], th[1], np.full(nv, sigma)]).astype(np.float32)
    g = gpu_abi.mcmc(md, cfg, y, x0, store=False)
    o = orc.mcmc(md, cfg, y, x0, store=False)
    for q in range(2):
        se = np.sqrt(g["std"][q] ** 2 + o["std"][q] ** 2) / np.sqrt(n_eff)
        z = (g["mean"][q] - o["mean"][q]) / se
        assert abs(np.median(z)) < 0.5 and np.mean(np.abs(z) < 4) > 0.95, (q, np.median(z), np.mean(np.abs(z) < 4))
        ratio = g["std"][q] / o["std"][q]
        assert 0.8 < np.median(ratio) < 1.25, (q, np.median(ratio))


def test_mh_native_statistics_analytic_target():
    """TE ~ 0 makes s = M0: with the noise fixed the M0 posterior is
    N(mean(y), sigma^2/M); native-stream chains must reproduce it."""
    from gacelle_amd import models
    from gacelle_amd.mcmc import mcmc as M

    te = np.full(12, 1e-9)
    m = models.gpuR2starMapping(te)
    nv, sigma = 400, 0.02
    rng = np.random.default_rng(3)
    mu = rng.uniform(0.8, 1.2, nv)
    y = (mu + sigma * rng.standard_normal((12, nv))).astype(np.float32)
    f = M.check_set_default_basic(dict(modelParams=["M0", "R2star", "noise"], lb=[0, 0.1, sigma],
                                       ub=[2, 200, sigma]))
    md = m.gc_model(f, "mcmc", 12)
    cfg = mcmc_cfg(_abi.GC_MCMC_MH, 20000, 10, 2000, rep=4, steps=(0.01, 1, 0.0))
    x0 = np.stack([y.mean(axis=0), np.full(nv, 30.0), np.full(nv, sigma)]).astype(np.float32)
    g = gpu_abi.mcmc(md, cfg, y, x0, store=False)
    post_mean = y.astype(np.float64).mean(axis=0)
    post_sd = sigma / np.sqrt(12)
    z = (g["mean"][0] - post_mean) / post_sd
    assert abs(np.mean(z)) < 0.05 and np.std(z) < 0.2, (np.mean(z), np.std(z))
    np.testing.assert_allclose(np.median(g["std"][0]), post_sd, rtol=0.05)


@pytest.mark.parametrize("algo", [_abi.GC_MCMC_MH, _abi.GC_MCMC_GW])
def test_unfilled_posterior_slots_are_zero(algo):
    """iteration 30, burn-in 5, thinning 3: Ns = numel(6:3:30) = 9 slots but
    only 8 iterations pass the thinning test (utils/mcmc.m:170); the last
    slot keeps the zeros of the preallocated posterior and enters the mean."""
    nv = 200
    y, th, sigma = r2s_problem(nv, seed=3)
    md = r2s_setup({}, solver="mcmc")[2]
    W = 8 if algo == _abi.GC_MCMC_GW else 1
    cfg = mcmc_cfg(algo, 30, 3, 5, rep=2, W=W, steps=(0.01, 1, 0.001))
    x0 = np.stack([th[0], th[1], np.full(nv, sigma)]).astype(np.float32)
    g = gpu_abi.mcmc(md, cfg, y, x0, post_fill=np.nan)
    assert (g["n_sample"], g["n_stored"]) == (9, 8)
    assert np.all(g["posterior"][:, :, 8] == 0) and np.all(np.isfinite(g["posterior"]))
    p = g["posterior"].astype(np.float64)  # [Q, R, Ns, W, nv]
    np.testing.assert_allclose(g["mean"], p.transpose(0, 4, 1, 2, 3).reshape(3, nv, -1).mean(-1), rtol=1e-5)


@pytest.mark.parametrize("redblack", [False, True])
def test_nexi_gl64_rician_gw_statistics_match_oracle(redblack):
    """Goodman-Weare at the NEXI benchmark settings (64-point Gauss-Legendre
    quadrature, Rician mean with its sigma row, 16 walkers; utils/mcmc.m:
    192-383, NEXI/gpuNEXIrice.m:635-660), native streams, against the FP64
    oracle's chains.  The exchange rate is barely identified at this SNR, so
    posteriors are broad and the Monte-Carlo error is calibrated, not
    assumed: the oracle run twice with different seeds gives the spread of
    posterior means that chance alone produces, and the device's departure
    from the oracle must be of that size (median |d| at most 1.5x), unbiased
    (the median departure within 4 standard errors of a sample median,
    1.2533 x 1.4826 MAD / sqrt(n); oracle-vs-oracle and device-vs-oracle
    pairs over three seeds put the median at up to 0.6 of the chance
    spread), with posterior SDs that agree."""
    from gacelle_amd import models
    from gacelle_amd.mcmc import mcmc as M

    b = np.repeat(np.arange(1, 7, dtype=float), 2)
    D = np.tile([20.0, 30.0], 6)
    m = models.gpuNEXI(b, D, quad="gl64", rician=True)
    names = ["fa", "Da", "De", "ra", "noise"]
    f = M.check_set_default_basic(dict(modelParams=names, lb=list(m.lb[:4]) + [0.01], ub=list(m.ub[:4]) + [0.1]))
    md = m.gc_model(f, "mcmc", None)
    md0 = models.gpuNEXI(b, D, quad="gl64").gc_model(
        M.check_set_default_basic(dict(modelParams=names[:4], lb=list(m.lb[:4]), ub=list(m.ub[:4]))), "mcmc", None)
    md0.n_par = 5
    rng = np.random.default_rng(44)
    nv = 96
    fa = rng.uniform(0.2, 0.8, nv)
    th = np.stack([fa, rng.uniform(1.5, 3, nv), rng.uniform(0.5, 1.5, nv), (1 - fa) / rng.uniform(2, 50, nv),
                   np.full(nv, 0.02)])
    s = orc.fwd_value(md0, th)
    y = np.sqrt((s + 0.02 * rng.standard_normal(s.shape)) ** 2 + (0.02 * rng.standard_normal(s.shape)) ** 2)
    y = y.astype(np.float32)
    sig = np.full((1, nv), 0.02, np.float32)
    x0 = th.astype(np.float32)
    cfg = mcmc_cfg(_abi.GC_MCMC_GW, 2000, 10, 200, rep=1, W=16, redblack=redblack, Q=5)
    cfg.noise_slot = 4
    g = gpu_abi.mcmc(md, cfg, y, x0, extra=sig, store=False)
    o = orc.mcmc(md, cfg, y, x0, extra=sig, store=False, nthreads=16)
    cfg2 = mcmc_cfg(_abi.GC_MCMC_GW, 2000, 10, 200, rep=1, W=16, redblack=redblack, Q=5, seed=777)
    cfg2.noise_slot = 4
    o2 = orc.mcmc(md, cfg2, y, x0, extra=sig, store=False, nthreads=16)
    for q in range(5):
        chance = np.median(np.abs(o2["mean"][q] - o["mean"][q]))
        dev = g["mean"][q] - o["mean"][q]
        assert np.median(np.abs(dev)) <= 1.5 * chance, (names[q], np.median(np.abs(dev)), chance)
        se_med = 1.2533 * 1.4826 * np.median(np.abs(dev - np.median(dev))) / np.sqrt(nv)
        assert abs(np.median(dev)) <= 4.0 * se_med, (names[q], np.median(dev), se_med)
        ratio = g["std"][q] / o["std"][q]
        assert 0.8 < np.median(ratio) < 1.25, (names[q], np.median(ratio))
