"""gc_fwd_adj_noalloc against gc_fwd_adj through the C ABI: the same
gradients, loss sums and residuals (bit for bit: the same kernels and the
same fixed-order reduction), on two calls in a row whose second input is
smaller, so stale partials or a stale model copy in the shared scratch
array would show."""
import ctypes

import numpy as np
import pytest

from gacelle_amd import _abi, models, runtime
from gacelle_amd.askadam import askadam

from .helpers import TE12, TE_JOINT, TR_JOINT, FA_JOINT, r2s_setup

pytestmark = pytest.mark.gpu


def _call(fn, md, loss, transform, x, y, w, extra, scale):
    c = runtime.context()
    P, nv = x.shape
    xd, yd = c.to_device(x), c.to_device(y)
    wd = c.to_device(w) if w is not None else None
    ex = c.to_device(extra) if extra is not None else None
    g = c.empty((P, nv))
    ls = c.zeros((1,), np.float64)
    rs = c.empty((md.n_meas, nv))
    c.check(fn(c.handle, ctypes.byref(md), loss, transform, runtime.ptr(xd), runtime.ptr(yd), runtime.ptr(wd),
               runtime.ptr(ex), nv, scale, runtime.ptr(g), runtime.ptr(ls), runtime.ptr(rs)))
    c.synchronize()
    return runtime.to_host(g), float(runtime.to_host(ls)[0]), runtime.to_host(rs)


def _problem(kind, nv, rng):
    extra = None
    if kind == "r2s":
        md = r2s_setup({})[2]
        th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    elif kind == "joint":
        m = models.gpuJointR1R2starMapping(TE_JOINT, TR_JOINT, FA_JOINT)
        f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1],
                                                 ub=[2, 10, 200]))
        md = m.gc_model(f, "askadam", TE_JOINT.size * FA_JOINT.size)
        th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
        extra = rng.uniform(0.8, 1.2, (1, nv)).astype(np.float32)
    else:
        b = np.repeat(np.arange(1, 7, dtype=float), 2)
        m = models.gpuNEXI(b, np.tile([20.0, 30.0], 6), quad="trapz14", rician=True)
        names = ["fa", "Da", "De", "ra"]
        f = askadam.check_set_default_basic(dict(modelParams=names, lb=m.lb[:4], ub=m.ub[:4]))
        md = m.gc_model(f, "askadam", None)
        fa = rng.uniform(0.2, 0.8, nv)
        th = np.stack([fa, rng.uniform(1.5, 3, nv), rng.uniform(0.5, 1.5, nv), (1 - fa) / rng.uniform(2, 40, nv)])
        extra = np.full((1, nv), 0.05, np.float32)
    z = np.stack([np.log((th[p] - md.lb[p]) / (md.ub[p] - th[p])) for p in range(th.shape[0])]).astype(np.float32)
    y = (rng.uniform(0.1, 1.0, (md.n_meas, nv))).astype(np.float32)
    w = rng.uniform(0.2, 1.0, (md.n_meas, nv)).astype(np.float32)
    return md, z, y, w, extra


@pytest.mark.parametrize("kind", ["r2s", "joint", "nexi_rice"])
@pytest.mark.parametrize("loss", [_abi.GC_LOSS_L1, _abi.GC_LOSS_L2])
def test_noalloc_matches_fwd_adj_on_two_calls(kind, loss):
    lib = runtime.context().lib
    rng = np.random.default_rng(17)
    for nv in (70001, 3333):  # the second call smaller than the first
        md, z, y, w, extra = _problem(kind, nv, rng)
        scale = 1.0 / (md.n_meas * nv)
        for ww in (w, None):
            ref = _call(lib.gc_fwd_adj, md, loss, _abi.GC_TRANSFORM_SIGMOID, z, y, ww, extra, scale)
            got = _call(lib.gc_fwd_adj_noalloc, md, loss, _abi.GC_TRANSFORM_SIGMOID, z, y, ww, extra, scale)
            np.testing.assert_array_equal(got[0], ref[0])
            assert got[1] == ref[1]
            np.testing.assert_array_equal(got[2], ref[2])


def test_noalloc_rejects_sizes_beyond_its_array():
    c = runtime.context()
    md = r2s_setup({})[2]
    # null buffers: only the size check runs, nothing is launched
    rc = c.lib.gc_fwd_adj_noalloc(c.handle, ctypes.byref(md), _abi.GC_LOSS_L1, _abi.GC_TRANSFORM_SIGMOID, None, None,
                                  None, None, _abi.GC_NOALLOC_MAX_NV + 1, 1.0, None, None, None)
    assert rc == _abi.GC_ERR_ARG
    assert b"at most" in c.lib.gc_last_error(c.handle)
