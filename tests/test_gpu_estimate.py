"""estimate() pipelines end to end on the GPU: the Rician NEXI model
(NEXI/gpuNEXIrice.m:104-447: sigma normalisation, b0-normalised
invariants, masks, likelihood start at the median sigma) and the joint
R1-R2* model with the MCMC solver (R1R2s/gpuJointR1R2starMapping.m:105-192:
M0 of every metric and of the posterior back to the data scale)."""
import numpy as np
import pytest

from gacelle_amd import models
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _dirs(n, seed):
    g = np.random.default_rng(seed).normal(size=(n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def test_nexi_rician_estimate_end_to_end():
    print("nexi rician estimate", flush=True)
    b = np.array([1.0, 2.5, 4.0, 6.0] * 2)
    D = np.array([20.0] * 4 + [30.0] * 4)
    m0 = models.gpuNEXI(b, D, rician=True)
    f = m0.check_set_default(dict(lmax=0))
    m0.updateProperty(f)
    dims = (5, 4, 2)
    nv = int(np.prod(dims))
    rng = np.random.default_rng(12)
    fa = rng.uniform(0.3, 0.7, nv)
    th = np.stack([fa, rng.uniform(1.8, 2.8, nv), rng.uniform(0.6, 1.2, nv), (1 - fa) / rng.uniform(5, 40, nv)])
    md = models.gpuNEXI(m0.b, m0.Delta).gc_model(dict(modelParams=m0.modelParams, lb=m0.lb, ub=m0.ub), "askadam", None)
    sl0 = orc.fwd_value(md, th)                                    # [8, nv], b0-normalised
    s0 = rng.uniform(800, 1200, nv)
    ndir = 6
    bvals, bdel, raw = [], [], []
    for delta in (20.0, 30.0):                                     # a b=0 volume in each diffusion-time group
        bvals.append(0.0)
        bdel.append(delta)
        raw.append(s0)
        for k, bk in enumerate(m0.b):
            if m0.Delta[k] != delta:
                continue
            for _ in range(ndir):
                bvals.append(float(bk))
                bdel.append(delta)
                raw.append(sl0[k] * s0)
    raw = np.stack(raw, axis=1) + rng.normal(0, 2.0, (nv, len(bvals)))
    raw = raw.reshape(dims + (-1,), order="F")
    g = _dirs(len(bvals), 3)
    sigma = np.full(dims, 10.0)
    ex = dict(bval=np.array(bvals), bvec=g.T, BDELTA=np.array(bdel), ldelta=1.0, sigma=sigma)
    mask = np.ones(dims, bool)
    common = dict(lmax=0, verbose=False, tol=0, convergenceValue=-1, lossFunction="L2", initialLearnRate=0.005,
                  start="likelihood")
    o1 = models.gpuNEXI(b, D, rician=True).estimate(raw, mask, ex, dict(common, iteration=1))
    o2 = models.gpuNEXI(b, D, rician=True).estimate(raw, mask, ex, dict(common, iteration=300))
    assert o2["mask"].sum() >= 0.9 * nv
    assert o2["final"]["loss"] < o1["final"]["loss"]
    # the prepared sigma is sigma / b0: the Rician floor of a 1% noise level
    m = models.gpuNEXI(b, D, rician=True)
    _, _, exn = m.prepare_dwi_data_rice(raw, mask, ex, 0)
    np.testing.assert_allclose(exn["sigma"], sigma / raw[..., np.array(bvals) == 0].mean(axis=3), rtol=1e-12)


TE = np.arange(2, 25, 3) * 1e-3
FA = np.array([4.0, 10.0, 20.0, 35.0])
TR = 30e-3


def test_joint_estimate_mcmc_rescales_m0():
    print("joint estimate mcmc", flush=True)
    rng = np.random.default_rng(6)
    dims = (4, 4, 2)
    m0 = rng.uniform(800, 1200, dims)
    r1 = rng.uniform(0.6, 1.5, dims)
    r2 = rng.uniform(15, 40, dims)
    e1 = np.exp(-TR * r1)
    sig = []
    for fa in np.float32(FA):
        a = np.deg2rad(fa)
        amp = m0 * np.sin(a) * (1 - e1) / (1 - np.cos(a) * e1)
        sig.append(amp[..., None] * np.exp(-r2[..., None] * TE))
    data = np.stack(sig, axis=4)
    data = data + rng.normal(0, 1.0, data.shape)
    m = models.gpuJointR1R2starMapping(TE, TR, FA)
    out = m.estimate(data, np.ones(dims, bool), None, dict(solver="mcmc", iteration=2000, thinning=10,
                                                          metric=["mean", "std", "median"]))
    for k in ("mean", "median"):
        rel = np.abs(out[k]["M0"] - m0) / m0
        assert np.median(rel) < 0.05, (k, np.median(rel))
    post = np.asarray(out["posterior"]["M0"])
    assert np.nanmedian(post[post > 0]) > 100          # data units, not the unit scale
    rel_r1 = np.abs(out["mean"]["R1"] - r1) / r1
    assert np.median(rel_r1) < 0.1


def test_segmented_estimate_matches_one_piece():
    """gpuR2starMapping.estimate split into slabs (fitting.NSegmentUser, or a
    memory budget below the footprint) equals the one-piece fit: the voxels
    are independent and, without stopping criteria, every slab runs the same
    iterations (Adam normalises the 1/Nv gradient scale), with an L2 loss."""
    print("segmented estimate", flush=True)
    from gacelle_amd import segment

    rng = np.random.default_rng(9)
    dims = (12, 10, 8)
    te = np.arange(2, 47, 4) * 1e-3
    s0 = rng.uniform(500, 1500, dims)
    r2 = rng.uniform(10, 60, dims)
    data = s0[..., None] * np.exp(-r2[..., None] * te) + rng.normal(0, 5.0, dims + (12,))
    mask = rng.random(dims) < 0.8
    # L2: with L1 the sign of near-zero residuals makes trajectories
    # sensitive to the FP32 rounding of the per-slab gradient scale
    common = dict(iteration=300, verbose=False, tol=0, convergenceValue=-1, convergenceStepTol=0,
                  convergenceGradTol=0, initialLearnRate=0.01, lossFunction="L2")
    one = models.gpuR2starMapping(te).estimate(data, mask, dict(common))
    two = models.gpuR2starMapping(te).estimate(data, mask, dict(common, NSegmentUser=2))
    assert isinstance(two["final"]["loss"], list) and len(two["final"]["loss"]) == 2
    m = mask
    for name in ("M0", "R2star"):
        np.testing.assert_allclose(two["final"][name][m], one["final"][name][m], rtol=2e-3)
    bpv = models.gpuR2starMapping(te)._segment_bytes(data, models.gpuR2starMapping.check_set_default({}))
    budget = segment.FIXED_BYTES + bpv * int(mask.sum()) // 3
    three = models.gpuR2starMapping(te).estimate(data, mask, dict(common, memoryBudget=budget))
    assert len(three["final"]["loss"]) >= 3
    np.testing.assert_allclose(three["final"]["R2star"][m], one["final"]["R2star"][m], rtol=2e-3)
