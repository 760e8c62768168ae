"""Parity of the joint R1-R2* workload against the FP64 oracle.

VFA-SPGR x multi-echo (R1R2s/gpuJointR1R2starMapping.m:375-422,
:513-525), 4 flip angles {3, 8, 15, 25} deg x 12 echoes (TE 2:2:24 ms),
TR 30 ms, M0 ~ U(0.5, 2), R1 ~ U(0.5, 2) /s, R2* ~ U(10, 60) /s,
B1 ~ U(0.8, 1.2), Rician noise at SNR 80; askadam with the reference's
defaults (L1, Adam lr 1e-3), 2000 iterations, stopping off.  Three volumes
of uneven size go through one batched launch, each with its own N; each is
compared with the matched-sign FP64 oracle run of that volume alone.

R1 is the weakly conditioned parameter of this protocol (four flip angles,
TR 30 ms): along the same sign path the oracle's own FP32 restatement
(std::exp, IEEE division) departs from FP64 by up to ~2e-4 in R1 over a few
thousand voxels, while the median departure is ~2e-6.  The bar is therefore
1e-4 on the 99.9th percentile of every parameter, and a maximum within 3x
of what the FP32 restatement shows on the same voxels.
"""
import os

import numpy as np
import pytest

from gacelle_amd import models
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import no_stop

pytestmark = pytest.mark.gpu

NTHREADS = max(1, min(16, len(os.sched_getaffinity(0))))
TE = np.arange(2, 25, 2) * 1e-3
FA = np.array([3.0, 8.0, 15.0, 25.0])


def _setup(iteration):
    m = models.gpuJointR1R2starMapping(TE, 30e-3, FA)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1],
                                             ub=[2, 10, 200], **no_stop(iteration=iteration)))
    return m.gc_model(f, "askadam", None), askadam().make_cfg(f, m)


def _volume(nv, rng, md):
    th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
    b1 = rng.uniform(0.8, 1.2, (1, nv))
    s = orc.fwd_value(md, th, b1)
    sig = th[0].mean() / 80
    y = np.sqrt((s + sig * rng.standard_normal(s.shape)) ** 2 + (sig * rng.standard_normal(s.shape)) ** 2)
    zc = [np.log(t / (1 - t)) for t in ((1.0 - 0) / 2, (1.0 - 0.1) / 9.9, (30.0 - 0.1) / 199.9)]
    z0 = np.repeat(np.array(zc)[:, None], nv, axis=1)
    return dict(y=y.astype(np.float32), extra=b1.astype(np.float32), z0=z0.astype(np.float32))


def _theta(z, md):
    z = np.asarray(z, np.float64)
    return np.stack([md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-z[p])) for p in range(3)])


def test_joint_three_uneven_volumes_l1_matched_sign_parity():
    md, cfg = _setup(2000)
    rng = np.random.default_rng(31)
    vols = [_volume(n, rng, md) for n in (1200, 2500, 4100)]
    got = gpu_abi.askadam(md, cfg, vols, signs=True)
    # one batched launch gives each volume exactly its own single-volume run
    alone = gpu_abi.askadam(md, cfg, [vols[1]])[0]
    np.testing.assert_array_equal(got[1]["z_final"], alone["z_final"])
    for i, (v, g) in enumerate(zip(vols, got)):
        ref = orc.askadam(md, cfg, v["y"], v["z0"], extra=v["extra"], sign_trace=g["signs"], nthreads=NTHREADS)
        r32 = orc.askadam(md, cfg, v["y"], v["z0"], extra=v["extra"], sign_trace=g["signs"], nthreads=NTHREADS,
                          precision=1)
        th_g, th_r, th_32 = _theta(g["z_final"], md), _theta(ref["z_final"], md), _theta(r32["z_final"], md)
        for p, name in enumerate(["M0", "R1", "R2*"]):
            e = np.abs(th_g[p] - th_r[p]) / np.abs(th_r[p])
            e32 = np.abs(th_32[p] - th_r[p]) / np.abs(th_r[p])
            print(f"\njoint volume {i} ({v['y'].shape[1]} voxels) {name}, matched signs: median {np.median(e):.1e} "
                  f"p99.9 {np.quantile(e, 0.999):.1e} max {e.max():.1e} (oracle FP32: max {e32.max():.1e})")
            assert np.quantile(e, 0.999) <= 1e-4, (i, name)
            assert e.max() <= max(1e-4, 3 * e32.max()), (i, name, e.max(), e32.max())
        np.testing.assert_allclose(g["loss"], ref["loss"], rtol=5e-5)
