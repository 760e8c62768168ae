"""GPU tests of launch choices that must not change results: the step-form
iteration blocks replayed from captured hipGraphs (against one launch per
iteration), the threads per block of the MH / GW samplers and of the blocked
askadam kernel (fixed sizes against each other and against the autotuner's
pick), and the autotuner's iterations per launch."""
import numpy as np
import pytest

from gacelle_amd import _abi, models, utils
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import TE12, logit_start, mcmc_cfg, r2s_problem, r2s_setup

pytestmark = pytest.mark.gpu


def _with_option(option, value, fn):
    c = gpu_abi.ctx()
    c.set_option(option, value)
    try:
        return fn()
    finally:
        c.set_option(_abi.GC_OPT_GRAPHS, 1)
        c.set_option(_abi.GC_OPT_MCMC_BLOCK, 0)
        c.set_option(_abi.GC_OPT_ASK_BLOCK, 0)


def _tv_problem(iteration, stopping, block_iters):
    dims = (20, 18, 5)
    rng = np.random.default_rng(5)
    mask = rng.random(dims) < 0.75
    vox = utils.find_mask(mask)
    nv = vox.size
    stop = {} if stopping else dict(tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0)
    _, f, md = r2s_setup(dict(iteration=iteration, lossFunction="L1", initialLearnRate=0.02, regmap=["R2star"],
                              TVmode="3D", voxelSize=[1, 1, 2], iterationsPerLaunch=block_iters,
                              **{"lambda": [0.002]}, **stop))
    cfg = askadam().make_cfg(f, models.gpuR2starMapping(TE12))
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    return md, cfg, [dict(y=y, z0=z0, vox=vox, dims=dims)]


@pytest.mark.parametrize("iteration,stopping,block_iters", [(150, False, 0), (300, True, 100), (97, False, 33)])
def test_step_graphs_equal_one_launch_per_iteration(iteration, stopping, block_iters):
    """Graph replay (chunks of up to 64 launches, descriptor rows moved to the
    front between chunks) gives bit-identical fits, traces and stops."""
    md, cfg, vols = _tv_problem(iteration, stopping, block_iters)
    g = _with_option(_abi.GC_OPT_GRAPHS, 1, lambda: gpu_abi.askadam(md, cfg, vols, trace=True)[0])
    d = _with_option(_abi.GC_OPT_GRAPHS, 0, lambda: gpu_abi.askadam(md, cfg, vols, trace=True)[0])
    assert g["n_iteration"] == d["n_iteration"] and g["stop_reason"] == d["stop_reason"]
    assert g["min_iteration"] == d["min_iteration"]
    np.testing.assert_array_equal(g["trace"], d["trace"])
    for k in ("z_final", "z_min", "resloss", "residual"):
        np.testing.assert_array_equal(g[k], d[k], err_msg=k)
    assert g["loss_reg"] == d["loss_reg"]


def _r2s_mcmc(nv, algo, rep, W, replay=False, redblack=False, iteration=60):
    y, th, _ = r2s_problem(nv, seed=3)
    _, f, md = r2s_setup(dict(iteration=iteration), solver="mcmc")
    cfg = mcmc_cfg(algo, iteration, 5, 10, rep=rep, W=W, steps=(0.01, 1.0, 0.005), redblack=redblack,
                   replay=replay)
    x0 = np.vstack([th, np.full((1, nv), 0.01)]).astype(np.float32)
    draws = {}
    if replay:
        rng = np.random.default_rng(9)
        Q = 3
        if algo == _abi.GC_MCMC_MH:
            draws = dict(normals=rng.standard_normal((rep * iteration, nv, Q)).astype(np.float32),
                         uniforms=rng.random((rep * iteration, nv)).astype(np.float32))
        else:
            h = W // 2
            part = np.empty((rep * iteration, W), np.int32)
            for i in range(rep * iteration):
                for w in range(W):
                    if redblack:
                        lo, hi = (h, W) if w < h else (0, h)
                        part[i, w] = rng.integers(lo, hi)
                    else:
                        p = rng.integers(0, W - 1)
                        part[i, w] = p + (p >= w)
            nu = rep * iteration * (2 * 2 * max(h, W - h) if redblack else 2 * W)
            draws = dict(normals=rng.standard_normal((W, nv, Q)).astype(np.float32),
                         uniforms=rng.random((nu, nv)).astype(np.float32), partners=part)
    return md, cfg, y, x0, draws


CASES = [
    ("mh_rep4", _abi.GC_MCMC_MH, 4, 1, False, False),
    ("gw_w16_wave_local", _abi.GC_MCMC_GW, 1, 16, False, False),
    ("gw_w8_rep2_redblack", _abi.GC_MCMC_GW, 2, 8, False, True),
    ("gw_w50_rep2_across_waves", _abi.GC_MCMC_GW, 2, 50, False, False),
    ("gw_w16_replay", _abi.GC_MCMC_GW, 1, 16, True, False),
    ("mh_replay", _abi.GC_MCMC_MH, 2, 1, True, False),
]


@pytest.mark.parametrize("name,algo,rep,W,replay,redblack", CASES)
def test_sampler_block_sizes_give_identical_chains(name, algo, rep, W, replay, redblack):
    """Every block size (64, 128, 256 threads, at least one voxel's chains
    per block) gives bit-identical posteriors, statistics and acceptance."""
    nv = 1537  # ragged against every block's voxel count
    md, cfg, y, x0, draws = _r2s_mcmc(nv, algo, rep, W, replay, redblack)
    gsz = rep * W
    outs = {}
    for bs in (64, 128, 256):
        if bs < gsz:
            continue
        outs[bs] = _with_option(_abi.GC_OPT_MCMC_BLOCK, bs,
                                lambda: gpu_abi.mcmc(md, cfg, y, x0, trace=True, **draws))
        assert outs[bs]["block_threads"] == bs
    assert len(outs) >= 2
    ref = outs[max(outs)]
    for bs, o in outs.items():
        for k in ("posterior", "mean", "std", "accept", "xfinal", "trace"):
            np.testing.assert_array_equal(o[k], ref[k], err_msg=f"{name} bs={bs} {k}")


def test_block_size_below_the_chain_group_is_rejected():
    md, cfg, y, x0, draws = _r2s_mcmc(64, _abi.GC_MCMC_GW, 2, 50)
    with pytest.raises(_abi.GacelleError, match="below Nwalker"):
        _with_option(_abi.GC_OPT_MCMC_BLOCK, 64, lambda: gpu_abi.mcmc(md, cfg, y, x0))


@pytest.mark.parametrize("algo,rep,W", [(_abi.GC_MCMC_MH, 4, 1), (_abi.GC_MCMC_GW, 1, 16)])
def test_autotuned_block_size_is_a_candidate_and_changes_nothing(algo, rep, W):
    """The autotuner's trial runs write no outputs: the tuned call equals a
    fixed-size call of the same block size, and the pick is cached."""
    nv = 40000
    md, cfg, y, x0, draws = _r2s_mcmc(nv, algo, rep, W, iteration=200)
    tuned = gpu_abi.mcmc(md, cfg, y, x0)
    bs = tuned["block_threads"]
    assert bs in (64, 128, 256) and bs >= rep * W
    again = gpu_abi.mcmc(md, cfg, y, x0)
    assert again["block_threads"] == bs
    fixed = _with_option(_abi.GC_OPT_MCMC_BLOCK, 256, lambda: gpu_abi.mcmc(md, cfg, y, x0))
    for k in ("posterior", "mean", "std", "accept", "xfinal"):
        np.testing.assert_array_equal(tuned[k], fixed[k], err_msg=k)
        np.testing.assert_array_equal(again[k], fixed[k], err_msg=k)


def test_option_values_are_checked():
    c = gpu_abi.ctx()
    for opt, bad in ((_abi.GC_OPT_GRAPHS, 2), (_abi.GC_OPT_MCMC_BLOCK, 96), (99, 0)):
        with pytest.raises(_abi.GacelleError):
            c.set_option(opt, bad)


def _blocked_problem(loss, norms):
    """Two volumes of uneven size (the second not a multiple of 64 voxels)
    for the blocked kernel, with a loss-convergence stop inside a block."""
    vols = []
    for nv, seed in ((1500, 21), (333, 22)):
        y, _, _ = r2s_problem(nv, seed=seed)
        vols.append(dict(y=y, z0=None))
    extra = dict(iteration=400, verbose=False, lossFunction=loss, initialLearnRate=2e-2, convergenceValue=1e-4,
                 convergenceStepTol=1e-9 if norms else 0, convergenceGradTol=0)
    _, f, md = r2s_setup(extra)
    cfg = askadam().make_cfg(f, None)
    for v in vols:
        nv = v["y"].shape[1]
        v["z0"] = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    return md, cfg, vols


@pytest.mark.parametrize("loss,norms", [("L1", False), ("L2", True)])
def test_askadam_block_threads_do_not_change_results(loss, norms):
    """The blocked kernel at 64, 128 and 256 threads per block and at the
    autotuner's pick: bit-identical fits, loss traces, stops and residuals
    (every wavefront sums its own 64 voxels; the loss-only path through its
    LDS ring, the norms path directly)."""
    md, cfg, vols = _blocked_problem(loss, norms)
    runs = {}
    for bs in (64, 128, 256, 0):
        runs[bs] = _with_option(_abi.GC_OPT_ASK_BLOCK, bs, lambda: gpu_abi.askadam(md, cfg, vols, trace=True))
        for o in runs[bs]:
            assert o["block_threads"] == (bs if bs else o["block_threads"]) and o["block_threads"] in (64, 128, 256)
    ref = runs[256]
    if loss == "L1":
        assert ref[0]["stop_reason"] != _abi.GC_STOP_NONE
    for bs in (64, 128, 0):
        for i, (g, r) in enumerate(zip(runs[bs], ref)):
            tag = f"block {bs}, volume {i}"
            assert (g["n_iteration"], g["stop_reason"], g["min_iteration"]) == \
                (r["n_iteration"], r["stop_reason"], r["min_iteration"]), tag
            np.testing.assert_array_equal(g["trace"], r["trace"], err_msg=tag)
            for k in ("z_final", "z_min", "resloss", "residual"):
                np.testing.assert_array_equal(g[k], r[k], err_msg=f"{tag} {k}")


def test_askadam_autotuned_launch_shape():
    """With stopping criteria on and no iterationsPerLaunch, the autotuner
    sets K from its measured launch and iteration costs (16 .. 256), keeps
    its pick for the next call of the same model variant, and the fit
    equals one with a fixed K of 1 (the block replay is exact)."""
    md, cfg, vols = _blocked_problem("L1", False)
    a = gpu_abi.askadam(md, cfg, vols, trace=True)
    b = gpu_abi.askadam(md, cfg, vols, trace=True)
    assert 16 <= a[0]["iters_per_launch"] <= 256
    assert (b[0]["block_threads"], b[0]["iters_per_launch"]) == (a[0]["block_threads"], a[0]["iters_per_launch"])
    cfg1 = _abi.gc_askadam_cfg.from_buffer_copy(cfg)
    cfg1.block_iters = 1
    c = gpu_abi.askadam(md, cfg1, vols, trace=True)
    assert c[0]["iters_per_launch"] == 1
    for g, r in zip(a, c):
        assert (g["n_iteration"], g["stop_reason"], g["min_iteration"]) == \
            (r["n_iteration"], r["stop_reason"], r["min_iteration"])
        np.testing.assert_array_equal(g["trace"], r["trace"])
        for k in ("z_final", "z_min", "resloss", "residual"):
            np.testing.assert_array_equal(g[k], r[k], err_msg=k)
