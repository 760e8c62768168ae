"""GPU parity of the parameter transforms against the FP64 oracle:
gc_rescale_parameters / gc_unscale_parameters (askadam.rescale_parameters /
unscale_parameters, utils/askadam.m:1063-1099), and askadam with
parameterTransform 'linear': set_boundary01 of the starting point
(:183-185), the clamped local copy that forward_pass differentiates while the
update applies to the unclamped parameters (:424-426), and set_boundary01 of
both results (:312-316)."""
import numpy as np
import pytest

from gacelle_amd import _abi
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import TE12, no_stop, r2s_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("transform", [_abi.GC_TRANSFORM_SIGMOID, _abi.GC_TRANSFORM_LINEAR])
def test_rescale_unscale_match_oracle(transform):
    md = r2s_setup({})[2]
    rng = np.random.default_rng(2)
    nv = 5000
    # inside the bounds, on them and beyond them (the sigmoid clamp at 1e-4 of the range)
    th = np.stack([rng.uniform(-0.5, 2.5, nv), rng.uniform(-20, 230, nv)]).astype(np.float32)
    th[:, :4] = [[0, 2, 1e-6, 2 - 1e-6], [0.1, 200, 0.1 + 1e-3, 200 - 1e-3]]
    z = gpu_abi.rescale(md, transform, th)
    zr = orc.rescale_parameters(md, transform, th.astype(np.float64))
    # compared in parameter space: at the clamp (1e-4 of the range from a
    # bound) the FP32 logit carries the cancellation of 1 - t, ~1e-3 relative
    # in z, which is ~1e-8 of the range in theta (the reference computes it in
    # single precision too)
    back, back_r = orc.unscale_parameters(md, transform, z), orc.unscale_parameters(md, transform, zr)
    rng_p = np.array([md.ub[p] - md.lb[p] for p in range(2)])[:, None]
    np.testing.assert_allclose(back / rng_p, back_r / rng_p, rtol=0, atol=1e-6)
    inner = np.abs(zr) < 5
    np.testing.assert_allclose(z[inner], zr[inner], rtol=2e-6, atol=2e-6)
    zz = np.stack([rng.uniform(-12, 12, nv), rng.uniform(-12, 12, nv)]).astype(np.float32)
    if transform == _abi.GC_TRANSFORM_LINEAR:
        zz = zz / 8
    t = gpu_abi.unscale(md, transform, zz)
    tr = orc.unscale_parameters(md, transform, zz.astype(np.float64))
    np.testing.assert_allclose(t, tr, rtol=2e-6, atol=1e-6)


def test_askadam_linear_transform_matches_oracle():
    """Adam on the linear transform, L2 loss: starting points partly outside
    [0, 1] (set_boundary01 brings them onto the boundary) and voxels whose
    data need R2* above ub (their R2* leaves [0, 1], gets no gradient from
    the clamped copy and keeps its momentum, and is clamped at the end).
    Per-parameter relative error against the FP64 oracle <= 1e-4."""
    nv = 3000
    rng = np.random.default_rng(9)
    s0 = rng.uniform(0.5, 1.5, nv)
    r2 = rng.uniform(10, 60, nv)
    r2[:300] = rng.uniform(230, 300, 300)   # beyond ub = 200
    y = (s0 * np.exp(-np.outer(TE12, r2)) + 0.005 * rng.standard_normal((12, nv))).astype(np.float32)
    # 600 iterations: converged, but not yet into the FP32 jitter of Adam at an
    # exact L2 minimum (gradients at rounding level), which is not the point here
    _, f, md = r2s_setup(no_stop(iteration=600, lossFunction="L2", initialLearnRate=3e-3,
                                 parameterTransform="linear"))
    cfg = askadam().make_cfg(f, None)
    assert cfg.transform == _abi.GC_TRANSFORM_LINEAR
    z0 = np.stack([rng.uniform(-0.3, 1.3, nv), rng.uniform(-0.3, 1.3, nv)]).astype(np.float32)
    got = gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0)], trace=True)[0]
    ref = orc.askadam(md, cfg, y, z0, trace=True, nthreads=8)
    # both results set_boundary01'ed; the beyond-ub voxels end on the boundary
    for z in (got["z_final"], got["z_min"]):
        assert z.min() >= 0 and z.max() <= 1
    at_ub = got["z_final"][1] == 1.0
    assert at_ub[:300].sum() > 100
    np.testing.assert_array_equal(at_ub, ref["z_final"][1] == 1.0)
    np.testing.assert_allclose(got["trace"], ref["trace"], rtol=1e-4)
    for p in range(2):
        th_g = got["z_final"][p].astype(np.float64) * (md.ub[p] - md.lb[p]) + md.lb[p]
        th_r = ref["z_final"][p] * (md.ub[p] - md.lb[p]) + md.lb[p]
        assert np.max(np.abs(th_g - th_r) / np.abs(th_r)) <= 1e-4, p
        thm_g = got["z_min"][p].astype(np.float64) * (md.ub[p] - md.lb[p]) + md.lb[p]
        thm_r = ref["z_min"][p] * (md.ub[p] - md.lb[p]) + md.lb[p]
        if got["min_iteration"] == ref["min_iteration"]:
            assert np.max(np.abs(thm_g - thm_r) / np.abs(thm_r)) <= 1e-4, p
