"""GPU parity of the normal-prior regulariser (GC_REG_PRIOR;
gpuNEXI.prior_distribution_normal, NEXI/gpuNEXI.m:1256-1303) inside the
askadam loop, against the oracle's whole-grid restatement, through the C
ABI; and gpuNEXI.fit with regularisationType 'prior' end to end."""
import numpy as np
import pytest

from gacelle_amd import models, utils
from gacelle_amd.askadam import _prior_rows
from oracle import oracle as orc

from . import gpu_abi
from .test_oracle_prior import _problem

pytestmark = pytest.mark.gpu


def _device_prior(cfg, vox, dims, mu, sd):
    f = dict(regularisationType="prior", regmap=["R2star"], **{"lambda": [cfg.lambda_[0]]})
    mask = np.zeros(int(np.prod(dims)), bool)
    mask[vox] = True
    mask = mask.reshape(dims, order="F")
    ex = dict(mu={"R2star": mu.reshape(dims, order="F")}, sigma={"R2star": sd.reshape(dims, order="F")})
    return _prior_rows(f, ex, mask, vox)


def test_prior_step_matches_oracle():
    """One SGDM step: the prior's value (in-mask sum on the device plus the
    host's out-of-mask constant) and its gradient match the FP64 oracle."""
    print("prior sgdm step", flush=True)
    md, cfg, y, z0, vox, dims, mu, sd = _problem(lr=1.0)
    rows, off = _device_prior(cfg, vox, dims, mu, sd)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, prior=rows, reg_offset=off)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, prior_grid=np.stack([mu, sd]), trace=True)
    np.testing.assert_allclose(got["trace"], ref["trace"], rtol=1e-5)
    assert got["loss_reg"] == pytest.approx(ref["loss_reg"], rel=1e-5)
    dz_g = got["z_final"].astype(np.float64) - z0
    dz_o = ref["z_final"] - z0
    for p in range(2):
        err = np.abs(dz_g[p] - dz_o[p]) / np.abs(dz_o[p]).max()
        assert err.max() < 1e-4, (p, err.max())


def test_prior_trajectory_within_fp32_spread():
    print("prior adam trajectory", flush=True)
    md, cfg, y, z0, vox, dims, mu, sd = _problem(lam=0.02, optimiser="adam", iteration=150, lr=0.02)
    rows, off = _device_prior(cfg, vox, dims, mu, sd)
    grid = np.stack([mu, sd])
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, prior=rows, reg_offset=off)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, prior_grid=grid, trace=True)
    f32 = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, prior_grid=grid, trace=True, precision=1)
    assert got["n_iteration"] == ref["n_iteration"] == 150
    spread = np.abs(f32["trace"] - ref["trace"]).max() / ref["trace"].max()
    assert np.abs(got["trace"] - ref["trace"]).max() / ref["trace"].max() <= 3 * spread + 1e-6


def test_nexi_fit_with_prior_pulls_to_mu():
    """A strong prior on fa pulls the fitted map towards mu; missing maps
    raise the reference's errors."""
    print("nexi prior fit", flush=True)
    b = np.array([1.0, 2.5, 4.0, 6.0] * 2)
    D = np.array([20.0] * 4 + [30.0] * 4)
    m = models.gpuNEXI(b, D)
    f = m.check_set_default(dict(lmax=0))
    m.updateProperty(f)
    dims = (6, 5, 2)
    nv = int(np.prod(dims))
    rng = np.random.default_rng(4)
    fa = rng.uniform(0.3, 0.7, nv)
    th = np.stack([fa, rng.uniform(1.8, 2.8, nv), rng.uniform(0.6, 1.2, nv), (1 - fa) / rng.uniform(5, 40, nv)])
    md = models.gpuNEXI(b, D).gc_model(dict(modelParams=m.modelParams, lb=m.lb, ub=m.ub), "askadam", None)
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((8, nv))).T.reshape(dims + (8,), order="F")
    mask = np.ones(dims, bool)
    mu = np.full(dims, 0.5)
    ex = dict(mu=dict(fa=mu), sigma=dict(fa=np.full(dims, 0.05)))
    common = dict(lmax=0, verbose=False, tol=0, convergenceValue=-1, lossFunction="L2", initialLearnRate=0.01,
                  iteration=400, regmap=["fa"])
    free = models.gpuNEXI(b, D).fit(y, mask, dict(common), None, None)
    pri = models.gpuNEXI(b, D).fit(y, mask, dict(common, regularisationType="prior", **{"lambda": [0.5]}),
                                   None, ex)
    d_free = np.abs(free["final"]["fa"] - 0.5).mean()
    d_pri = np.abs(pri["final"]["fa"] - 0.5).mean()
    assert d_pri < 0.5 * d_free, (d_pri, d_free)
    assert pri["final"]["loss"] > 0
    from gacelle_amd._abi import GacelleError
    with pytest.raises(GacelleError, match="missingPriorDistribution"):
        models.gpuNEXI(b, D).fit(y, mask, dict(common, regularisationType="prior", **{"lambda": [0.5]}), None, None)
    with pytest.raises(GacelleError, match="incompletePriorDistribution"):
        models.gpuNEXI(b, D).fit(y, mask, dict(common, regularisationType="prior", **{"lambda": [0.5]}), None,
                                 dict(mu=dict(fa=mu), sigma=dict()))
