"""GPU parity of the kernels' reduced-transcendental forms against the FP64
oracle: the echo decays from one exponential when the spacing is one or two
first-echo times (and from two otherwise), the batched reciprocals of the
sigmoids and Adam denominators in and out of their range, and the joint
model's per-voxel flip-angle sines and cosines."""
import numpy as np
import pytest

from gacelle_amd import _abi, models
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import FA_JOINT, TE_JOINT, TR_JOINT, logit_start, no_stop

pytestmark = pytest.mark.gpu

# 12 equally spaced echoes: spacing = 2 x first echo, = first echo, and a
# spacing that is no small multiple of the first echo (two exponentials)
TE_SETS = {
    "k2": np.arange(2, 47, 4) * 1e-3,
    "k1": np.arange(2, 25, 2) * 1e-3,
    "two_exp": np.arange(3, 48, 4) * 1e-3,
}


def _r2s(te, fitting, lb=(0, 0.1)):
    m = models.gpuR2starMapping(te)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=list(lb), ub=[2, 200], **fitting))
    return f, m.gc_model(f, "askadam", te.size)


def _problem(te, nv, seed):
    rng = np.random.default_rng(seed)
    s0 = rng.uniform(0.5, 1.5, nv)
    r2 = rng.uniform(10, 60, nv)
    y = s0 * np.exp(-np.outer(te, r2))
    y = (y + s0.mean() / 100 * rng.standard_normal(y.shape)).astype(np.float32)
    return y, np.stack([s0, r2])


@pytest.mark.parametrize("name", list(TE_SETS))
def test_r2s_echo_spacing_forms_match_oracle(name):
    """Signals within 2e-5 of the oracle, and a 400-iteration L2 fit within
    1e-4 per parameter, for each form of the uniform-spacing decays."""
    te = TE_SETS[name]
    nv = 2048
    f, md = _r2s(te, no_stop(iteration=400, lossFunction="L2", initialLearnRate=1e-2))
    assert md.flags & _abi.GC_FLAG_UNIFORM_TE
    y, th = _problem(te, nv, seed=21)
    np.testing.assert_allclose(gpu_abi.fwd(md, th), orc.fwd_value(md, th), rtol=2e-5, atol=2e-6)
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)])[0]
    ref = orc.askadam(md, cfg, y, z0)
    for p in range(2):
        th_g = md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-got["z_final"][p].astype(np.float64)))
        th_r = md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-ref["z_final"][p]))
        assert np.max(np.abs(th_g - th_r) / np.abs(th_r)) < 1e-4, (name, p)


def test_batched_reciprocals_out_of_range_match_oracle():
    """Starting points far in the sigmoid's tail make the product of the two
    denominators 1 + exp(-z) exceed 2^100, so every lane takes the exact
    per-value reciprocals; half the voxels start there, half at ordinary
    points, so wavefronts mix both.  M0's lower bound 0.5 keeps the tail
    voxels' signal (0.5 exp(-0.1 TE)) and loss away from zero, so the
    reciprocals' values show in the loss."""
    te = TE_SETS["k2"]
    nv = 1024
    f, md = _r2s(te, no_stop(iteration=60, lossFunction="L2", initialLearnRate=1e-2), lb=(0.5, 0.1))
    y, _ = _problem(te, nv, seed=5)
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    z0[:, ::2] = -40.0  # (1 + e^40)^2 ~ 5.5e34 > 2^100
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)])[0]
    ref = orc.askadam(md, cfg, y, z0)
    assert np.all(np.isfinite(got["z_final"]))
    np.testing.assert_allclose(got["z_final"], ref["z_final"], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(got["loss"], ref["loss"], rtol=1e-5)


@pytest.mark.parametrize("true_fa", [False, True])
def test_joint_voxel_prep_matches_oracle(true_fa):
    """The 12 x 4 joint fit (equally spaced echoes, spacing = first echo)
    with its per-voxel sines and cosines (B1 x FA, or the supplied true flip
    angles) and batched steady-state reciprocals: 300 L2 iterations against
    the oracle."""
    m = models.gpuJointR1R2starMapping(TE_JOINT, TR_JOINT, FA_JOINT)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1], ub=[2, 10, 200],
                                             **no_stop(iteration=300, lossFunction="L2", initialLearnRate=1e-2)))
    nv = 3000
    rng = np.random.default_rng(8)
    b1 = rng.uniform(0.8, 1.2, (1, nv))
    extra_data = {"trueFlipAngle": (b1.T * np.deg2rad(FA_JOINT)).T} if true_fa else {"b1": b1}
    md = m.model_for_extra(m.gc_model(f, "askadam", TE_JOINT.size * FA_JOINT.size), extra_data)
    assert md.flags & _abi.GC_FLAG_UNIFORM_TE
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
