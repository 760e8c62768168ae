"""GPU checks at the benchmark workloads' full sizes.

The FP64 oracle cannot replay millions of voxels for 2000 iterations in a
test, so full-size parity goes through properties that hold at any size:

- askadam without a regulariser fits every voxel independently (its
  gradient is scaled by the loss mean's N, gc_volume.n_mean).  A random
  subset re-fitted on its own with the full run's N must be bit-identical
  to the full-size run at those voxels: batching, block layout, launch
  count and volume size change nothing.
- That subset, small enough for the oracle, must then match the FP64
  oracle along the device's residual signs (matched-sign mode), to the
  same bars as the small-size parity tests.

Workloads: the headline (mono-exponential R2*, 12 echoes TE 2:4:46 ms,
5,873,183 voxels, L1, Adam lr 1e-3, 2000 iterations, stopping off;
R2star/gpuR2starMapping.m:321-363, utils/askadam.m:86-88) and the joint
R1-R2* workload (4 flip angles x 12 echoes, three volumes of 0.4M / 0.9M /
1.6M voxels in one batched launch; R1R2s/gpuJointR1R2starMapping.m:375-422).
"""
import ctypes
import os

import numpy as np
import pytest

from gacelle_amd import _abi, models, runtime
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import logit_start, no_stop, r2s_problem, r2s_setup

pytestmark = pytest.mark.gpu

NTHREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _run(md, cfg, vols, signs=False):
    """The batched askadam call with the final variables only (no residual
    outputs).  vols: dicts of y, z0, extra (optional), n_mean (optional).
    Returns the final z of every volume, and with signs=True the residual
    sign trace of every volume."""
    c = gpu_abi.ctx()
    arr = (_abi.gc_volume * len(vols))()
    keep, zs, sgs = [], [], []
    for i, v in enumerate(vols):
        y = np.ascontiguousarray(v["y"], np.float32)
        M, nv = y.shape
        z0 = np.ascontiguousarray(v["z0"], np.float32)
        yd, zd = c.to_device(y), c.to_device(z0)
        zm = c.empty(z0.shape)
        ex = c.to_device(v["extra"]) if v.get("extra") is not None else None
        sg = c.zeros((int(cfg.iteration), (M + 15) // 16, nv), np.int32) if signs else None
        a = arr[i]
        a.nv = nv
        a.n_mean = int(v.get("n_mean", 0))
        a.y_dev, a.z_dev, a.zmin_dev = runtime.ptr(yd), runtime.ptr(zd), runtime.ptr(zm)
        a.extra_dev = runtime.ptr(ex)
        a.sign_trace_dev = runtime.ptr(sg)
        keep += [yd, zm, ex]
        zs.append(zd)
        sgs.append(sg)
    c.check(c.lib.gc_askadam_run(c.handle, ctypes.byref(md), ctypes.byref(cfg), arr, len(vols)))
    z = [runtime.to_host(zd) for zd in zs]
    if not signs:
        return z, [a.n_iteration for a in arr]
    return z, [runtime.to_host(s).view(np.uint32) for s in sgs]


def _theta(z, md):
    z = np.asarray(z, np.float64)
    return np.stack([md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-z[p])) for p in range(z.shape[0])])


def _rel(a, b):
    return np.abs(a - b) / np.abs(b)


def test_headline_full_size_subset_refit_and_oracle():
    nv = 5_873_183
    y, th_true, _ = r2s_problem(nv, seed=5438973)
    _, f, md = r2s_setup(no_stop(iteration=2000, lossFunction="L1", initialLearnRate=1e-3))
    cfg = askadam().make_cfg(f, None)
    z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    print(f"\nheadline full size: {nv} voxels", flush=True)
    (zf,), (nit,) = _run(md, cfg, [dict(y=y, z0=z0)])
    assert nit == 2000 and np.all(np.isfinite(zf))
    th = _theta(zf, md)
    med = [float(np.median(_rel(th[p], th_true[p]))) for p in range(2)]
    print(f"median rel err against the ground truth: M0 {med[0]:.2e} R2* {med[1]:.2e}", flush=True)
    assert max(med) < 0.05   # SNR 100: the noise floor is about 1-2%

    idx = np.sort(np.random.default_rng(1).choice(nv, 4096, replace=False))
    ys, z0s = np.ascontiguousarray(y[:, idx]), np.ascontiguousarray(z0[:, idx])
    (zs,), (sg,) = _run(md, cfg, [dict(y=ys, z0=z0s, n_mean=nv)], signs=True)
    np.testing.assert_array_equal(zs, zf[:, idx])

    ref = orc.askadam(md, cfg, ys, z0s, sign_trace=sg, n_mean=nv, nthreads=NTHREADS)
    th_g, th_r = _theta(zs, md), _theta(ref["z_final"], md)
    err = [float(np.max(_rel(th_g[p], th_r[p]))) for p in range(2)]
    print(f"subset of 4096 against the matched-sign oracle: max rel err M0 {err[0]:.2e} R2* {err[1]:.2e}")
    assert max(err) <= 1e-4, err


def test_joint_full_size_batched_subset_refit_and_oracle():
    te = np.arange(2, 25, 2) * 1e-3
    fa = np.array([3.0, 8.0, 15.0, 25.0])
    m = models.gpuJointR1R2starMapping(te, 30e-3, fa)
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1],
                                             ub=[2, 10, 200], **no_stop(iteration=2000)))
    md, cfg = m.gc_model(f, "askadam", None), askadam().make_cfg(f, m)
    rng = np.random.default_rng(3)
    zc = [np.log(t / (1 - t)) for t in ((1.0 - 0) / 2, (1.0 - 0.1) / 9.9, (30.0 - 0.1) / 199.9)]
    vols = []
    for nv in (100 * 100 * 40, 150 * 150 * 40, 200 * 200 * 40):
        th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
        b1 = rng.uniform(0.8, 1.2, (1, nv)).astype(np.float32)
        s = gpu_abi.fwd(md, th, b1).astype(np.float64)
        sig = th[0].mean() / 80
        y = np.sqrt((s + sig * rng.standard_normal(s.shape)) ** 2 + (sig * rng.standard_normal(s.shape)) ** 2)
        vols.append(dict(y=y.astype(np.float32), extra=b1,
                         z0=np.repeat(np.array(zc, np.float32)[:, None], nv, axis=1)))
    print(f"\njoint full size: volumes of {[v['y'].shape[1] for v in vols]} voxels", flush=True)
    zb, nits = _run(md, cfg, vols)
    assert nits == [2000, 2000, 2000] and all(np.all(np.isfinite(z)) for z in zb)

    vi = 1
    nv = vols[vi]["y"].shape[1]
    idx = np.sort(np.random.default_rng(2).choice(nv, 2048, replace=False))
    sub = dict(y=np.ascontiguousarray(vols[vi]["y"][:, idx]), extra=np.ascontiguousarray(vols[vi]["extra"][:, idx]),
               z0=np.ascontiguousarray(vols[vi]["z0"][:, idx]), n_mean=nv)
    (zs,), (sg,) = _run(md, cfg, [sub], signs=True)
    np.testing.assert_array_equal(zs, zb[vi][:, idx])

    ref = orc.askadam(md, cfg, sub["y"], sub["z0"], extra=sub["extra"], sign_trace=sg, n_mean=nv,
                      nthreads=NTHREADS)
    th_g, th_r = _theta(zs, md), _theta(ref["z_final"], md)
    p999 = [float(np.quantile(_rel(th_g[p], th_r[p]), 0.999)) for p in range(3)]
    mx = [float(np.max(_rel(th_g[p], th_r[p]))) for p in range(3)]
    print(f"subset of 2048 against the matched-sign oracle: p99.9 rel err {p999}, max {mx}")
    # the small-size joint test's bars: 1e-4 on the 99.9th percentile; R1
    # is the weakly conditioned parameter (up to ~2e-4 between FP32 and FP64
    # restatements along the same sign path)
    assert max(p999) <= 1e-4, p999
    assert max(mx) <= 1e-3, mx
