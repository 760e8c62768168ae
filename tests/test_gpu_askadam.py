"""GPU parity of the fused askadam path against the FP64 oracle, through
the C ABI (gc_fwd, gc_fwd_adj, gc_askadam_run)."""
import numpy as np
import pytest

from gacelle_amd import _abi, models
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import (FA_JOINT, TE12, TE_JOINT, TR_JOINT, logit_start, no_stop, r2s_problem, r2s_setup,
                      replay_loss_stopping)

pytestmark = pytest.mark.gpu


def _joint_md(fitting_extra=None):
    m = models.gpuJointR1R2starMapping(TE_JOINT, TR_JOINT, FA_JOINT)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1],
                                             ub=[2, 10, 200], **(fitting_extra or {})))
    return m, f, m.gc_model(f, "askadam", TE_JOINT.size * FA_JOINT.size)


def _nexi_md(quad="trapz14", lmax=0, rician=False):
    b = np.repeat(np.arange(1, 7, dtype=float), 2)
    D = np.tile([20.0, 30.0], 6)
    m = models.gpuNEXI(b, D, quad=quad, rician=rician)
    names = ["fa", "Da", "De", "ra"] + (["p2"] if lmax == 2 else [])
    f = askadam.check_set_default_basic(dict(modelParams=names, lb=m.lb[: len(names)], ub=m.ub[: len(names)]))
    return m, f, m.gc_model(f, "askadam", None)


def _nexi_theta(nv, rng, lmax=0):
    fa = rng.uniform(0.2, 0.8, nv)
    Da = rng.uniform(1.5, 3, nv)
    De = rng.uniform(0.5, 1.5, nv)
    tex = rng.uniform(2, 50, nv)
    th = [fa, Da, De, (1 - fa) / tex]
    if lmax == 2:
        th.append(rng.uniform(0.1, 0.9, nv))
    return np.stack(th)


@pytest.mark.parametrize("kind", ["r2s", "joint", "nexi_trapz", "nexi_gl64", "nexi_l2", "nexi_rice"])
def test_fwd_matches_oracle(kind):
    rng = np.random.default_rng(3)
    nv = 777
    extra = None
    if kind == "r2s":
        _, _, md = r2s_setup({})
        th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    elif kind == "joint":
        _, _, md = _joint_md()
        th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
        extra = rng.uniform(0.8, 1.2, (1, nv))
    else:
        lmax = 2 if kind == "nexi_l2" else 0
        _, _, md = _nexi_md(quad="gl64" if kind == "nexi_gl64" else "trapz14", lmax=lmax,
                            rician=kind == "nexi_rice")
        th = _nexi_theta(nv, rng, lmax)
        if kind == "nexi_rice":
            extra = np.full((1, nv), 0.05)
    ref = orc.fwd_value(md, th, extra)
    got = gpu_abi.fwd(md, th, None if extra is None else extra.astype(np.float32))
    # FP32 kernel vs FP64 oracle: relative 2e-5 (transcendentals in FP32)
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-6)


@pytest.mark.parametrize("kind", ["r2s", "joint", "nexi_trapz", "nexi_l2", "nexi_rice"])
@pytest.mark.parametrize("loss", [_abi.GC_LOSS_L1, _abi.GC_LOSS_L2, _abi.GC_LOSS_HUBER, _abi.GC_LOSS_MSE])
def test_fwd_adj_matches_oracle_gradient(kind, loss):
    """Hand-written adjoint vs the oracle's forward-mode dual derivative."""
    rng = np.random.default_rng(7)
    nv = 500
    extra = None
    if kind == "r2s":
        _, _, md = r2s_setup({})
        th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    elif kind == "joint":
        _, _, md = _joint_md()
        th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
        extra = rng.uniform(0.8, 1.2, (1, nv))
    else:
        lmax = 2 if kind == "nexi_l2" else 0
        _, _, md = _nexi_md(lmax=lmax, rician=kind == "nexi_rice")
        th = _nexi_theta(nv, rng, lmax)
        if kind == "nexi_rice":
            extra = np.full((1, nv), 0.05)
    s, jac = orc.fwd(md, th, extra)
    y = s + 0.02 * rng.standard_normal(s.shape)
    w = rng.uniform(0.2, 1.0, s.shape)
    r = s - y
    if loss == _abi.GC_LOSS_L1:
        dl = w * np.sign(r)
        lsum = np.sum(w * np.abs(r))
    elif loss == _abi.GC_LOSS_L2:
        dl = 2 * w * r
        lsum = np.sum(w * r * r)
    elif loss == _abi.GC_LOSS_HUBER:
        a = np.abs(r)
        dl = np.where(a <= 1, w * r, w * np.sign(r))
        lsum = np.sum(np.where(a <= 1, 0.5 * w * r * r, w * (a - 0.5)))
    else:
        dl = r
        lsum = np.sum(0.5 * r * r)
        w = None
    g_ref = np.einsum("mv,mpv->pv", dl, jac)
    out = gpu_abi.fwd_adj(md, loss, _abi.GC_TRANSFORM_NONE, th, y.astype(np.float32),
                          None if w is None else w.astype(np.float32),
                          None if extra is None else extra.astype(np.float32), scale=1.0)
    np.testing.assert_allclose(out["loss_sum"], lsum, rtol=2e-5)
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-12
    np.testing.assert_allclose(out["grad"] / scale, g_ref / scale, atol=5e-4)


def test_askadam_l2_trajectory_matches_oracle():
    """2000 iterations, L2 loss, stopping off: per-iteration loss and final
    parameters against the FP64 oracle (relative 1e-4 on parameters)."""
    nv = 4096
    y, th_true, _ = r2s_problem(nv, seed=11)
    _, f, md = r2s_setup(no_stop(iteration=2000, lossFunction="L2", initialLearnRate=1e-2))
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0, trace=True)
    assert got["n_iteration"] == ref["n_iteration"] == 2000
    # FP32 device arithmetic vs FP64 oracle over 2000 iterations
    np.testing.assert_allclose(got["trace"], ref["trace"], rtol=5e-5)
    for p in range(2):
        th_g = md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-got["z_final"][p].astype(np.float64)))
        th_r = md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-ref["z_final"][p]))
        rel = np.abs(th_g - th_r) / np.abs(th_r)
        # per-parameter relative error against the FP64 oracle: <= 1e-4
        assert np.max(rel) < 1e-4, (p, np.max(rel))
    # the minimum-loss iteration is the argmin of the run's own loss trace
    # (on a plateau FP32 and FP64 may pick different, equally low iterations)
    assert got["trace"][got["min_iteration"]] == got["trace"].min()
    assert got["min_iteration"] == int(np.argmin(got["trace"]))
    np.testing.assert_allclose(got["min_loss"], ref["min_loss"], rtol=1e-5)
    np.testing.assert_allclose(got["residual"], ref["residual"], rtol=1e-3, atol=1e-6)


@pytest.mark.parametrize("opt", ["adam", "sgdm", "rmsprop"])
@pytest.mark.parametrize("form", ["effective", "kingma"])
def test_askadam_optimisers_short(opt, form):
    if opt != "adam" and form == "kingma":
        pytest.skip("form only applies to Adam")
    nv = 1000
    y, _, _ = r2s_problem(nv, seed=5)
    _, f, md = r2s_setup(no_stop(iteration=50, lossFunction="L2", optimiser=opt, adamForm=form,
                                 initialLearnRate=1e-2, decayRate=0.01))
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.full(nv, 0.8), np.full(nv, 25.0)]), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)])[0]
    ref = orc.askadam(md, cfg, y, z0)
    # RMSProp/Adam steps are normalised: where a gradient component is ~0
    # the FP32 and FP64 sign decisions may differ (step ~ lr either way)
    close = np.isclose(got["z_final"], ref["z_final"], rtol=2e-5, atol=2e-5)
    assert close.mean() > 0.95, close.mean()
    np.testing.assert_allclose(got["loss"], ref["loss"], rtol=1e-4)


@pytest.mark.parametrize("model", ["ema", "linear"])
def test_askadam_loss_stopping(model):
    """Loss-convergence stopping: the device run stops exactly where the
    reference loop would on the device's own loss trace, and within a few
    percent of the FP64 oracle's stop iteration."""
    nv = 2000
    y, _, _ = r2s_problem(nv, seed=2)
    _, f, md = r2s_setup(dict(iteration=3000, verbose=False, initialLearnRate=1e-2, convergenceValue=1e-5,
                              convergenceModel=model, convergenceStepTol=0, convergenceGradTol=0))
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0)
    e, reason = replay_loss_stopping(got["trace"], f)
    assert (got["n_iteration"], got["stop_reason"]) == (e, reason)
    assert got["stop_reason"] == ref["stop_reason"] != _abi.GC_STOP_NONE
    assert abs(got["n_iteration"] - ref["n_iteration"]) <= max(3, 0.03 * ref["n_iteration"])
    assert got["min_iteration"] == int(np.argmin(got["trace"])) or got["trace"][got["min_iteration"]] == \
        got["trace"].min()


@pytest.mark.parametrize("which", ["step", "grad", "tol"])
def test_askadam_other_stopping_signals(which):
    """Step-norm, gradient-norm and loss-tolerance stops land on the same
    iteration as the oracle (thresholds chosen so FP32/FP64 agree)."""
    nv = 1000
    y, _, _ = r2s_problem(nv, seed=3)
    extra = dict(iteration=500, verbose=False, initialLearnRate=1e-2, convergenceValue=-1, tol=0,
                 convergenceStepTol=0, convergenceGradTol=0)
    if which == "step":
        extra["convergenceStepTol"] = 1e3
    elif which == "grad":
        extra["convergenceGradTol"] = 1e3
    else:
        extra["tol"] = 1.0
    _, f, md = r2s_setup(extra)
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)])[0]
    ref = orc.askadam(md, cfg, y, z0)
    assert (got["n_iteration"], got["stop_reason"]) == (ref["n_iteration"], ref["stop_reason"])
    np.testing.assert_allclose(got["z_final"], ref["z_final"], rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("k", [1, 7, 64])
def test_block_replay_is_exact(k):
    """Rollback/replay when stops and minima fall inside a launch: results
    are bit-identical whatever the number of iterations per launch."""
    nv = 1500
    y, _, _ = r2s_problem(nv, seed=9)
    base = dict(iteration=400, verbose=False, initialLearnRate=2e-2, convergenceValue=1e-4)
    _, f, md = r2s_setup(dict(base, iterationsPerLaunch=k))
    _, f1, _ = r2s_setup(dict(base, iterationsPerLaunch=400))
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    a = gpu_abi.askadam(md, askadam().make_cfg(f, None), [dict(y=y, z0=z0)])[0]
    b = gpu_abi.askadam(md, askadam().make_cfg(f1, None), [dict(y=y, z0=z0)])[0]
    assert a["n_iteration"] == b["n_iteration"]
    assert a["min_iteration"] == b["min_iteration"]
    np.testing.assert_array_equal(a["z_final"], b["z_final"])
    np.testing.assert_array_equal(a["z_min"], b["z_min"])


def test_multi_volume_batch_equals_single_runs():
    """Uneven volumes in one batched launch: each volume's result equals its
    own single-volume run (own N, stopping state and outputs)."""
    _, f, md = r2s_setup(dict(iteration=600, verbose=False, initialLearnRate=1e-2, convergenceValue=1e-5))
    cfg = askadam().make_cfg(f, None)
    vols = []
    for i, nv in enumerate([300, 5000, 1234]):
        y, _, _ = r2s_problem(nv, seed=100 + i)
        z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
        vols.append(dict(y=y, z0=z0))
    batch = gpu_abi.askadam(md, cfg, vols)
    for v, b in zip(vols, batch):
        s = gpu_abi.askadam(md, cfg, [v])[0]
        assert s["n_iteration"] == b["n_iteration"]
        assert s["stop_reason"] == b["stop_reason"]
        np.testing.assert_array_equal(s["z_final"], b["z_final"])


def test_joint_askadam_matches_oracle():
    nv = 2000
    rng = np.random.default_rng(4)
    m, f, md = _joint_md(no_stop(iteration=300, lossFunction="L2", initialLearnRate=1e-2))
    th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
    b1 = rng.uniform(0.8, 1.2, (1, nv))
    s = orc.fwd_value(md, th, b1)
    y = (s + 0.01 * rng.standard_normal(s.shape)).astype(np.float32)
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.ones(nv), np.full(nv, 30.0)]), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, extra=b1.astype(np.float32))])[0]
    ref = orc.askadam(md, cfg, y, z0, extra=b1)
    np.testing.assert_allclose(got["loss"], ref["loss"], rtol=1e-5)
    np.testing.assert_allclose(got["z_final"], ref["z_final"], rtol=1e-4, atol=1e-4)


def test_python_mirror_end_to_end():
    """askadam().optimisation through the model class on an N-D volume."""
    rng = np.random.default_rng(21)
    X, Y = 20, 30
    s0 = rng.uniform(0.5, 1.5, (X, Y, 1))
    r2 = rng.uniform(10, 60, (X, Y, 1))
    data = s0[..., None] * np.exp(-r2[..., None] * TE12.reshape(1, 1, 1, -1))
    mask = np.ones((X, Y, 1), bool)
    mask[0, 0, 0] = False
    m = models.gpuR2starMapping(TE12)
    out = m.fit(data.astype(np.float32), mask, dict(start="default", iteration=1500, lossFunction="L2",
                                                    initialLearnRate=5e-2, verbose=False, tol=0,
                                                    convergenceValue=-1))
    est = out["final"]["R2star"][..., 0]
    err = np.abs(est - r2[..., 0])[mask[..., 0]]
    assert np.median(err) < 0.5
    assert est[0, 0] == 0  # masked out


def test_reused_cached_blocks_give_identical_results():
    """The library's block caches hand released device and pinned buffers to
    later calls: a run repeated after calls of other sizes is bit-identical
    to the first."""
    from gacelle_amd import runtime

    _, f, md = r2s_setup(dict(iteration=300, verbose=False, initialLearnRate=1e-2, convergenceValue=1e-5))
    cfg = askadam().make_cfg(f, None)
    nv = 3000
    y, _, _ = r2s_problem(nv, seed=41)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    first = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)])[0]
    for n2, seed in ((12000, 42), (700, 43)):
        y2, _, _ = r2s_problem(n2, seed=seed)
        gpu_abi.askadam(md, cfg, [dict(y=y2, z0=logit_start(np.stack([np.ones(n2), np.full(n2, 30.0)]), md))])
    again = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)])[0]
    for key in ("z_final", "z_min"):
        np.testing.assert_array_equal(first[key], again[key])
    assert (first["n_iteration"], first["min_iteration"]) == (again["n_iteration"], again["min_iteration"])
    assert runtime.context() is not None


@pytest.mark.parametrize("true_fa", [False, True])
def test_joint_askadam_unequal_echoes_matches_oracle(true_fa):
    """The joint model's grouped blocked adjoint on the guarded variant
    (5 unequally spaced echoes x 3 flip angles: per-echo exponentials, runtime
    echo and flip-angle counts), with b1 or with extraData.trueFlipAngle
    (R1R2s/gpuJointR1R2starMapping.m:387-391), against the oracle."""
    te = np.array([2.0, 3.5, 6.0, 9.0, 13.0]) * 1e-3
    fa = np.array([4.0, 12.0, 20.0])
    m = models.gpuJointR1R2starMapping(te, 25e-3, fa)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1], ub=[2, 10, 200],
                                             **no_stop(iteration=300, lossFunction="L2", initialLearnRate=1e-2)))
    nv = 3000
    rng = np.random.default_rng(14)
    b1 = rng.uniform(0.8, 1.2, (1, nv))
    extra_data = {"trueFlipAngle": (b1.T * np.deg2rad(fa)).T} if true_fa else {"b1": b1}
    md = m.model_for_extra(m.gc_model(f, "askadam", te.size * fa.size), extra_data)
    assert not (md.flags & _abi.GC_FLAG_UNIFORM_TE)
    ex = m.extra_rows(extra_data, np.ones(nv, bool), np.arange(nv), nv).astype(np.float32)
    th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
    s = orc.fwd_value(md, th, ex.astype(np.float64))
    y = (s + 0.01 * rng.standard_normal(s.shape)).astype(np.float32)
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.ones(nv), np.full(nv, 30.0)]), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, extra=ex)])[0]
    ref = orc.askadam(md, cfg, y, z0, extra=ex.astype(np.float64))
    np.testing.assert_allclose(got["loss"], ref["loss"], rtol=1e-5)
    np.testing.assert_allclose(got["z_final"], ref["z_final"], rtol=1e-4, atol=1e-4)
