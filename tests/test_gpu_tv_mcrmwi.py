"""GPU parity of the per-iteration askadam path: spatial TV inside the loop
(any model) and the MCR-MWI model (fused forward + adjoint, streamed
data), against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

from gacelle_amd import _abi, models, utils
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import logit_start, r2s_setup

pytestmark = pytest.mark.gpu

TE24 = np.arange(2, 49, 2) * 1e-3
FA4 = np.array([5.0, 10.0, 20.0, 50.0])


def _mcrmwi(cplx=True, exch=True, te=TE24, lam=0.0, tv="3D"):
    m = models.gpuMCRMWI(te, 38e-3, FA4)
    f = m.check_set_default(dict(isComplex=cplx, isFitExchange=exch, isEPG=False))
    m.updateProperty(f)
    f["modelParams"] = list(m.modelParams)
    f["lb"], f["ub"] = list(m.lb), list(m.ub)
    if lam > 0:
        f["lambda"] = [lam]
        f["TVmode"] = tv
        f["voxelSize"] = [1, 1, 2]
    return m, f, m.gc_model(f, "askadam", None)


def _mcrmwi_theta(m, f, nv, rng):
    th = []
    for name, sl in m.param_slots(f):
        k = m.modelParams.index(name)
        lo, hi = m.lb[k], m.ub[k]
        n = sl.stop - sl.start
        a = lo + (hi - lo) * rng.uniform(0.3, 0.7, (n, nv))
        th.append(a)
    return np.concatenate(th, axis=0).astype(np.float32)


def _mcrmwi_extra(m, nv, rng):
    ex = np.zeros((3 + m.fa.size, nv), np.float32)
    ex[0] = rng.uniform(0.9, 1.1, nv)
    ex[1] = rng.uniform(-0.5, 0.5, nv)
    ex[2] = 1.0
    ex[3:] = rng.uniform(-0.05, 0.05, (m.fa.size, nv))
    return ex


@pytest.mark.parametrize("cplx,exch,uniform", [(True, True, True), (True, False, True), (False, True, True),
                                               (False, False, False), (True, True, False)])
def test_mcrmwi_forward_matches_oracle(cplx, exch, uniform):
    te = TE24 if uniform else np.sort(np.random.default_rng(3).uniform(2e-3, 48e-3, 24))
    m, f, md = _mcrmwi(cplx, exch, te)
    assert bool(md.flags & _abi.GC_FLAG_UNIFORM_TE) == uniform
    rng = np.random.default_rng(4)
    nv = 300
    th = _mcrmwi_theta(m, f, nv, rng)
    ex = _mcrmwi_extra(m, nv, rng)
    g = gpu_abi.fwd(md, th, ex)
    o = orc.fwd_value(md, th, ex)
    scale = np.abs(o).max(axis=0, keepdims=True)
    assert np.abs(g - o).max() / scale.max() < 2e-5
    np.testing.assert_allclose(g, o, rtol=1e-3, atol=2e-5 * scale.max())


@pytest.mark.parametrize("cplx,exch", [(True, True), (False, True), (True, False)])
@pytest.mark.parametrize("loss", [_abi.GC_LOSS_L1, _abi.GC_LOSS_L2])
def test_mcrmwi_adjoint_matches_oracle_gradient(cplx, exch, loss):
    m, f, md = _mcrmwi(cplx, exch)
    rng = np.random.default_rng(5)
    nv = 200
    th = _mcrmwi_theta(m, f, nv, rng)
    ex = _mcrmwi_extra(m, nv, rng)
    y = (orc.fwd_value(md, _mcrmwi_theta(m, f, nv, rng), ex) + 0.01 * rng.standard_normal((md.n_meas, nv)))
    y = y.astype(np.float32)
    z = logit_start(th.astype(np.float64), md)
    g = gpu_abi.fwd_adj(md, loss, _abi.GC_TRANSFORM_SIGMOID, z, y, extra=ex, scale=1.0)
    # reference gradient: the oracle's dual Jacobian, the loss derivative and the sigmoid chain
    P = md.n_par
    lb = np.array([md.lb[p] for p in range(P)])[:, None]
    ub = np.array([md.ub[p] for p in range(P)])[:, None]
    sg = 1 / (1 + np.exp(-z.astype(np.float64)))
    thz = lb + (ub - lb) * sg
    s, jac = orc.fwd(md, thz, ex)
    r = s - y
    dl = np.sign(r) if loss == _abi.GC_LOSS_L1 else 2 * r
    ref = np.einsum("mv,mpv->pv", dl, jac) * (ub - lb) * sg * (1 - sg)
    for p in range(P):
        err = np.abs(g["grad"][p] - ref[p]) / max(np.abs(ref[p]).max(), 1e-12)
        assert np.quantile(err, 0.99) < 1e-4 and err.max() < 5e-3, (p, err.max())


@pytest.mark.parametrize("loss", [_abi.GC_LOSS_L1, _abi.GC_LOSS_L2])
def test_mcrmwi_pini_per_flip_angle_matches_oracle(loss):
    """One initial phase per flip angle (extra rows 3+nfa..3+2nfa): forward
    and adjoint against the oracle."""
    m, f, _ = _mcrmwi(True, True)
    f["piniPerFA"] = True
    md = m.gc_model(f, "askadam", None)
    rng = np.random.default_rng(15)
    nv = 200
    th = _mcrmwi_theta(m, f, nv, rng)
    ex = np.vstack([_mcrmwi_extra(m, nv, rng), rng.uniform(-3, 3, (m.fa.size, nv)).astype(np.float32)])
    g = gpu_abi.fwd(md, th, ex)
    o = orc.fwd_value(md, th, ex)
    assert np.abs(g - o).max() / np.abs(o).max() < 2e-5
    y = (orc.fwd_value(md, _mcrmwi_theta(m, f, nv, rng), ex) + 0.01 * rng.standard_normal((md.n_meas, nv)))
    y = y.astype(np.float32)
    z = logit_start(th.astype(np.float64), md)
    ga = gpu_abi.fwd_adj(md, loss, _abi.GC_TRANSFORM_SIGMOID, z, y, extra=ex, scale=1.0)
    P = md.n_par
    lb = np.array([md.lb[p] for p in range(P)])[:, None]
    ub = np.array([md.ub[p] for p in range(P)])[:, None]
    sg = 1 / (1 + np.exp(-z.astype(np.float64)))
    s, jac = orc.fwd(md, lb + (ub - lb) * sg, ex)
    r = s - y
    dl = np.sign(r) if loss == _abi.GC_LOSS_L1 else 2 * r
    ref = np.einsum("mv,mpv->pv", dl, jac) * (ub - lb) * sg * (1 - sg)
    for p in range(P):
        err = np.abs(ga["grad"][p] - ref[p]) / max(np.abs(ref[p]).max(), 1e-12)
        assert np.quantile(err, 0.99) < 1e-4 and err.max() < 5e-3, (p, err.max())


def _tv_problem(tv, iteration, optimiser, lr, noisy_start):
    dims = (14, 12, 6)
    rng = np.random.default_rng(11)
    mask = rng.random(dims) < 0.7
    vox = utils.find_mask(mask)
    nv = vox.size
    _, f, md = r2s_setup(dict(iteration=iteration, lossFunction="L2", optimiser=optimiser, initialLearnRate=lr,
                              tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0,
                              regmap=["R2star"], TVmode=tv, voxelSize=[1, 1, 2], **{"lambda": [0.002]}))
    cfg = askadam().make_cfg(f, models.gpuR2starMapping(np.arange(2, 47, 4) * 1e-3))
    assert cfg.n_reg == 1 and cfg.regmap[0] == 1
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    if noisy_start:
        z0 = logit_start(np.stack([rng.uniform(0.8, 1.2, nv), rng.uniform(20, 40, nv)]), md)
    else:
        z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    return md, cfg, y, z0, vox, dims


@pytest.mark.parametrize("tv", ["2D", "3D"])
def test_askadam_tv_step_matches_oracle(tv):
    """One SGDM step (z1 = z0 - lr g) from a start with no tied neighbours:
    the TV value and its stencil adjoint, through the loop, match FP64."""
    md, cfg, y, z0, vox, dims = _tv_problem(tv, 2, "sgdm", 50.0, True)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, vox=vox, dims=dims)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, trace=True)
    np.testing.assert_allclose(got["trace"], ref["trace"], rtol=1e-5)
    assert got["loss_reg"] == pytest.approx(ref["loss_reg"], rel=1e-5)
    dz_g = got["z_final"].astype(np.float64) - z0
    dz_o = ref["z_final"] - z0
    for p in range(2):
        err = np.abs(dz_g[p] - dz_o[p]) / np.abs(dz_o[p]).max()
        assert err.max() < 1e-4, (p, err.max())


@pytest.mark.parametrize("tv", ["2D", "3D"])
def test_askadam_tv_trajectory_within_fp32_spread(tv):
    """150 Adam iterations from a constant start.  The TV subgradient
    sign(theta_j - theta_i) flips as neighbours cross, so any FP32 run
    parts from FP64: the bar is the spread of the oracle's own FP32
    restatement of the same loop (the device within 3x of it)."""
    md, cfg, y, z0, vox, dims = _tv_problem(tv, 150, "adam", 0.01, False)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, vox=vox, dims=dims)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, trace=True)
    f32 = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, trace=True, precision=1)
    assert got["n_iteration"] == ref["n_iteration"] == 150
    spread = np.abs(f32["trace"] - ref["trace"]).max() / ref["trace"].max()
    assert np.abs(got["trace"] - ref["trace"]).max() / ref["trace"].max() <= 3 * spread + 1e-6
    lb = np.array([md.lb[0], md.lb[1]])[:, None]
    ub = np.array([md.ub[0], md.ub[1]])[:, None]

    def theta(z):
        return lb + (ub - lb) / (1 + np.exp(-np.asarray(z, np.float64)))

    t_ref = theta(ref["z_final"])
    rel_g = np.abs(theta(got["z_final"]) - t_ref) / np.abs(t_ref)
    rel_32 = np.abs(theta(f32["z_final"]) - t_ref) / np.abs(t_ref)
    for q in (0.5, 0.99):
        assert np.all(np.quantile(rel_g, q, axis=1) <= 3 * np.quantile(rel_32, q, axis=1) + 1e-5), q


def test_askadam_tv_gradient_pulls_neighbours_together():
    """With a large lambda the TV term dominates: the fitted R2* map is
    smoother than the unregularised fit (sanity of the sign of the term)."""
    dims = (10, 10, 4)
    rng = np.random.default_rng(12)
    mask = np.ones(dims, bool)
    vox = utils.find_mask(mask)
    nv = vox.size
    common = dict(iteration=300, lossFunction="L2", initialLearnRate=0.05, tol=0, convergenceValue=-1,
                  convergenceStepTol=0, convergenceGradTol=0)
    th = np.stack([np.ones(nv), rng.uniform(20, 40, nv)])
    _, f0, md = r2s_setup(common)
    y = orc.fwd_value(md, th).astype(np.float32) + 0.02 * rng.standard_normal((12, nv)).astype(np.float32)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    out = {}
    for lam in (0.0, 0.05):
        extra = dict(regmap=["R2star"], TVmode="3D", voxelSize=[1, 1, 1], **{"lambda": [lam]})
        _, f, md = r2s_setup(dict(common, **extra))
        cfg = askadam().make_cfg(f, models.gpuR2starMapping(np.arange(2, 47, 4) * 1e-3))
        got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, vox=vox, dims=dims)])[0]
        r2 = 0.1 + 199.9 / (1 + np.exp(-got["z_final"][1].astype(np.float64)))
        img = r2.reshape(dims, order="F")
        out[lam] = np.abs(np.diff(img, axis=0)).mean() + np.abs(np.diff(img, axis=1)).mean()
    assert out[0.05] < 0.5 * out[0.0], out


@pytest.mark.parametrize("lam", [0.0, 1e-3])
def test_askadam_mcrmwi_matches_oracle(lam):
    m, f, md = _mcrmwi(True, True, lam=lam, tv="3D")
    dims = (8, 7, 4)
    rng = np.random.default_rng(21)
    mask = rng.random(dims) < 0.8
    vox = utils.find_mask(mask)
    nv = vox.size
    th = _mcrmwi_theta(m, f, nv, rng)
    ex = _mcrmwi_extra(m, nv, rng)
    y = (orc.fwd_value(md, th, ex) + 0.005 * rng.standard_normal((md.n_meas, nv))).astype(np.float32)
    f2 = askadam.check_set_default_basic(dict(f, iteration=60, lossFunction="L2", initialLearnRate=0.01, tol=0,
                                              convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0))
    cfg = askadam().make_cfg(f2, m)
    assert (cfg.n_reg == 1 and cfg.regmap[0] == 1) if lam > 0 else cfg.n_reg == 0
    z0 = logit_start(_mcrmwi_theta(m, f, nv, np.random.default_rng(22)).astype(np.float64), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, extra=ex, vox=vox, dims=dims)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0, extra=ex, vox=vox, dims=dims, trace=True)
    assert got["n_iteration"] == ref["n_iteration"] == 60
    np.testing.assert_allclose(got["trace"], ref["trace"], rtol=1e-4)
    P = md.n_par
    lb = np.array([md.lb[p] for p in range(P)])[:, None]
    ub = np.array([md.ub[p] for p in range(P)])[:, None]
    tg = lb + (ub - lb) / (1 + np.exp(-got["z_final"].astype(np.float64)))
    to = lb + (ub - lb) / (1 + np.exp(-ref["z_final"]))
    rel = np.abs(tg - to) / (ub - lb)
    assert rel.max() < 1e-4, rel.max(axis=1)


def test_tv_needs_voxel_indices():
    _, f, md = r2s_setup(dict(iteration=10, regmap=["R2star"], **{"lambda": [0.01]}))
    cfg = askadam().make_cfg(f, models.gpuR2starMapping(np.arange(2, 47, 4) * 1e-3))
    y = np.ones((12, 50), np.float32)
    with pytest.raises(_abi.GacelleError, match="TV needs vox_dev"):
        gpu_abi.askadam(md, cfg, [dict(y=y, z0=np.zeros((2, 50), np.float32))])


def test_estimate_end_to_end_with_per_flip_angle_field():
    """gpuMCRMWI.estimate (MCRMWI/gpuMCRMWI.m:217-317) on a synthetic complex
    volume with a per-flip-angle background field and no pini: scaling,
    validation, prior, weights, per-flip-angle pini and the askadam fit run
    end to end; the fit lowers the loss below the starting point's and S0
    comes back in data units.  (The default pini takes the first flip
    angle's first-echo phase, pool mixture included, so dpini and dfreqBKG
    absorb an offset: this checks the pipeline, not the model's accuracy,
    which the parity tests pin.)"""
    m = models.gpuMCRMWI(TE24, 38e-3, FA4)
    f = m.check_set_default(dict(isComplex=True, isFitExchange=True))
    mt = models.gpuMCRMWI(TE24, 38e-3, FA4)
    mt.updateProperty(f)
    f["modelParams"], f["lb"], f["ub"] = list(mt.modelParams), list(mt.lb), list(mt.ub)
    f["piniPerFA"] = True
    md = mt.gc_model(f, "askadam", None)
    dims = (6, 5, 2)
    nv = int(np.prod(dims))
    rng = np.random.default_rng(31)
    th = _mcrmwi_theta(mt, f, nv, rng).astype(np.float64)
    k = mt.modelParams.index("S0")
    th[k] = rng.uniform(0.8, 1.2, nv) * 1000.0
    fb = rng.uniform(-0.05, 0.05, dims + (4,))
    phase0 = rng.uniform(-1, 1, dims)
    B0g = 3.0 * 42.57747892
    pini_fa = np.angle(np.exp(1j * (phase0[..., None] - 2 * np.pi * fb * B0g * TE24[0])))
    ex = np.zeros((3 + 8, nv))
    ex[0], ex[2] = 1.0, 1.0
    ex[3:7] = fb.reshape(nv, 4, order="F").T
    ex[7:11] = pini_fa.reshape(nv, 4, order="F").T
    md.lb[k], md.ub[k] = 0.0, 4000.0
    s = orc.fwd_value(md, th, ex)                                    # [24*4*2, nv]
    nte, nfa = 24, 4
    z = (s[: nte * nfa] + 1j * s[nte * nfa:]).reshape(nte, nfa, nv)
    img = np.transpose(z, (2, 0, 1)).reshape(dims + (nte, nfa), order="F")
    img = img + 2.0 * (rng.standard_normal(img.shape) + 1j * rng.standard_normal(img.shape))
    mask = np.ones(dims, bool)
    common = dict(isComplex=True, isFitExchange=True, initialLearnRate=0.01, lossFunction="L2", verbose=False,
                  tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0)
    o1 = m.estimate(img, mask, dict(freqBKG=fb), dict(common, iteration=1))
    m2 = models.gpuMCRMWI(TE24, 38e-3, FA4)
    o2 = m2.estimate(img, mask, dict(freqBKG=fb), dict(common, iteration=800))
    assert o2["final"]["S0"].shape == dims and o2["final"]["dfreqBKG"].shape == dims + (4,)
    assert o2["final"]["loss"] < 0.8 * o1["final"]["loss"], (o1["final"]["loss"], o2["final"]["loss"])
    assert o2["min"]["loss"] <= o2["final"]["loss"]
    keep = o2["mask"]
    assert keep.sum() > 0.8 * nv
    ratio = o2["final"]["S0"][keep] / th[k].reshape(dims, order="F")[keep]
    assert 0.5 < np.median(ratio) < 2.0
