"""The oracle's normal-prior regulariser (gpuNEXI.prior_distribution_normal,
NEXI/gpuNEXI.m:1256-1303) pinned on the CPU: its gradient against finite
differences of the starting loss (one SGDM step is z0 - lr g), the loss at
theta = mu (only the grid voxels outside the mask remain), linearity in
lambda, and the exclusion of non-finite z-scores."""
import numpy as np

from gacelle_amd import _abi, models, utils
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from .helpers import logit_start, r2s_setup

TE12 = np.arange(2, 47, 4) * 1e-3


def _problem(lam=0.05, optimiser="sgdm", iteration=2, lr=1.0, seed=3):
    dims = (5, 4, 3)
    rng = np.random.default_rng(seed)
    mask = rng.random(dims) < 0.7
    vox = utils.find_mask(mask)
    nv = vox.size
    _, f, md = r2s_setup(dict(iteration=iteration, lossFunction="L2", optimiser=optimiser, initialLearnRate=lr,
                              tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0,
                              regmap=["R2star"], regularisationType="prior", **{"lambda": [lam]}))
    cfg = askadam().make_cfg(f, models.gpuR2starMapping(TE12))
    assert cfg.reg_type == _abi.GC_REG_PRIOR and cfg.n_reg == 1
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    ngrid = int(np.prod(dims))
    mu = rng.uniform(20, 40, ngrid).astype(np.float32)
    sd = rng.uniform(2, 8, ngrid).astype(np.float32)
    z0 = logit_start(np.stack([rng.uniform(0.8, 1.2, nv), rng.uniform(15, 45, nv)]), md)
    return md, cfg, y, z0, vox, dims, mu, sd


def _loss0(md, cfg, y, z0, vox, dims, grid):
    return orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, prior_grid=grid, trace=True)["trace"][0]


def test_prior_gradient_matches_finite_differences():
    md, cfg, y, z0, vox, dims, mu, sd = _problem()
    grid = np.stack([mu, sd])
    out = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, prior_grid=grid)
    g = -(out["z_final"] - z0.astype(np.float64)) / cfg.lr0       # SGDM: z1 = z0 - lr g (epoch 2 evaluates z1)
    h = 1e-3
    for v in (0, 3, 7):
        zp, zm = z0.copy(), z0.copy()
        zp[1, v] += h
        zm[1, v] -= h
        fd = (_loss0(md, cfg, y, zp, vox, dims, grid) - _loss0(md, cfg, y, zm, vox, dims, grid)) / (2 * h)
        assert abs(fd - g[1, v]) <= 2e-3 * abs(g[1, v]) + 1e-7, (v, fd, g[1, v])


def test_prior_closed_forms():
    md, cfg, y, z0, vox, dims, mu, sd = _problem()
    nv = vox.size
    th = md.lb[1] + (md.ub[1] - md.lb[1]) / (1 + np.exp(-z0[1].astype(np.float64)))
    mu2 = mu.copy()
    mu2[vox] = th.astype(np.float32)                              # theta = mu inside the mask
    outside = np.ones(mu.size, bool)
    outside[vox] = False
    expect = cfg.lambda_[0] / nv * np.sum(np.abs(mu[outside] / sd[outside]))
    r0 = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, prior_grid=np.stack([mu2, sd]), trace=True)
    fid0 = orc.askadam(md, askadam().make_cfg(dict(askadam.check_set_default_basic(
        dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200], iteration=1, lossFunction="L2",
             optimiser="sgdm", initialLearnRate=1.0, tol=0, convergenceValue=-1, convergenceStepTol=0,
             convergenceGradTol=0))), models.gpuR2starMapping(TE12)), y, z0, trace=True)["trace"][0]
    np.testing.assert_allclose(r0["trace"][0] - fid0, expect, rtol=1e-5)
    # linear in lambda
    grid = np.stack([mu, sd])
    l1 = _loss0(md, cfg, y, z0, vox, dims, grid) - fid0
    cfg2 = type(cfg).from_buffer_copy(cfg)
    cfg2.lambda_[0] = 2 * cfg.lambda_[0]
    l2 = _loss0(md, cfg2, y, z0, vox, dims, grid) - fid0
    np.testing.assert_allclose(l2, 2 * l1, rtol=1e-9)
    # a zero sigma at a masked voxel (z-score inf) drops that voxel's term,
    # as does a zero sigma outside the mask
    sd3 = sd.copy()
    sd3[vox[2]] = 0
    sd3[np.flatnonzero(outside)[0]] = 0
    l3 = _loss0(md, cfg, y, z0, vox, dims, np.stack([mu, sd3])) - fid0
    zin = np.abs((th[2] - mu[vox[2]]) / sd[vox[2]])
    zout = np.abs(mu[np.flatnonzero(outside)[0]] / sd[np.flatnonzero(outside)[0]])
    np.testing.assert_allclose(l3, l1 - cfg.lambda_[0] / nv * (zin + zout), rtol=1e-6)
