"""GPU parity of the dictionary search (gc_dict_search; gpuNEXI.likelihood,
NEXI/gpuNEXI.m:822-849) against the FP64 oracle through the C ABI, and
the likelihood start and estimate() glue of gpuNEXI and
gpuJointR1R2starMapping end to end."""
import numpy as np
import pytest

from gacelle_amd import _abi, models, utils
from gacelle_amd._abi import GacelleError
from oracle import oracle as orc

from . import gpu_abi

pytestmark = pytest.mark.gpu


def _problem(m, nv, nd, seed):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.05, 1.0, (m, nd)).astype(np.float32)
    y = rng.uniform(0.05, 1.0, (m, nv)).astype(np.float32)
    return y, d


@pytest.mark.parametrize("m", [2, 3, 4, 8, 13, 16, 32])
@pytest.mark.parametrize("weighted", [False, True])
def test_dict_search_parity(m, weighted):
    """Ragged sizes (voxels not a multiple of 64, entries not a multiple of
    16); the index is bit-exact, the FP64 sse within 1e-9 of the oracle's."""
    print(f"dict parity m={m} weighted={weighted}", flush=True)
    y, d = _problem(m, 1037, 517, seed=m + 100 * weighted)
    w = np.random.default_rng(m).uniform(0.1, 1.0, m).astype(np.float32) if weighted else None
    b, s = gpu_abi.dict_search(y, d, w)
    bo, so = orc.dict_search(y, d, None if w is None else w.astype(np.float64))
    np.testing.assert_array_equal(b, bo)
    y2 = ((w if w is not None else 1.0) * y.astype(np.float64).T ** 2).sum(axis=1)
    np.testing.assert_allclose(s, so, rtol=1e-9, atol=1e-13 * float(np.max(y2)))


def test_dict_search_edges():
    """Planted exact matches, first-index ties, a zero entry (NaN score),
    an all-NaN voxel (-1), a zero voxel (first finite entry), one voxel,
    one entry."""
    print("dict edges", flush=True)
    y, d = _problem(5, 70, 200, seed=1)
    k = np.arange(70) * 2 % 200
    y = (d[:, k] * np.linspace(0.5, 2.0, 70)).astype(np.float32)
    d[:, 0] = 0
    d[:, 17] = d[:, 102]                    # entry 102 (planted for voxel 51) has an earlier twin
    y[:, 0] = np.nan
    y[:, 1] = 0
    b, s = gpu_abi.dict_search(y, d)
    bo, so = orc.dict_search(y, d)
    np.testing.assert_array_equal(b, bo)
    assert b[0] == -1 and np.isnan(s[0]) and b[1] == 1 and s[1] == 0
    assert b[51] == 17 and s[51] < 1e-12
    np.testing.assert_array_equal(b[2:51], k[2:51])
    b1, _ = gpu_abi.dict_search(y[:, 5:6], d)
    assert b1[0] == k[5]
    bd, _ = gpu_abi.dict_search(y, d[:, 7:8])
    np.testing.assert_array_equal(bd[1:], 0)
    # one measurement: every entry fits exactly (sse 0 up to rounding), so
    # the index is arbitrary in any finite precision
    b1m, s1m = gpu_abi.dict_search(y[:1, 2:], d[:1])
    assert np.all(b1m >= 1) and np.all(np.abs(s1m) <= 1e-12 * y[0, 2:].astype(np.float64) ** 2)
    with pytest.raises(GacelleError):
        gpu_abi.dict_search(np.zeros((33, 4), np.float32), np.ones((33, 4), np.float32))


def test_dict_search_full_size_subset():
    """10,000 entries (the reference's dictionary) against 100,000 voxels;
    a random subset checked against the oracle."""
    print("dict full size", flush=True)
    y, d = _problem(8, 100_000, 10_000, seed=7)
    b, s = gpu_abi.dict_search(y, d)
    sub = np.random.default_rng(0).choice(100_000, 400, replace=False)
    bo, so = orc.dict_search(y[:, sub], d, threads=8)
    np.testing.assert_array_equal(b[sub], bo)
    np.testing.assert_allclose(s[sub], so, rtol=1e-9)


def _nexi(lmax=0):
    b = np.array([0, 1, 2.5, 4, 6] * 2, float)
    D = np.array([20] * 5 + [30] * 5, float)
    return models.gpuNEXI(b, D)


def test_nexi_dictionary_signals_match_oracle():
    """traindata's signals (the library's trapezoid-14 forward model) equal
    the oracle's forward model of the drawn parameters."""
    print("nexi dictionary", flush=True)
    for lmax in (0, 2):
        m = _nexi()
        x, S = m.traindata(2000, lmax)
        assert x.shape == (4 + lmax // 2, 2000) and S.shape == (8 * (1 + lmax // 2), 2000)
        tex = (1 - x[0]) / x[3]
        assert np.all((x[0] >= 0.01) & (x[0] <= 0.99) & (tex >= 1) & (tex <= 100))
        names = ["fa", "Da", "De", "ra", "p2"][: x.shape[0]]
        f = dict(modelParams=names, lb=[0] * len(names), ub=[1e3] * len(names))
        md = models.gpuNEXI(m.b, m.Delta, quad="trapz14").gc_model(f, "askadam", None)
        ref = orc.fwd_value(md, x)
        np.testing.assert_allclose(S, ref, rtol=2e-5, atol=2e-6)


def test_nexi_likelihood_start_recovers_planted_entries():
    """Voxels that are scaled dictionary signals get those entries back;
    low-b CSF-like voxels (D0 > 1.5) have fa, ra (and p2) scaled by 0.01."""
    print("nexi likelihood start", flush=True)
    for lmax in (0, 2):
        m = _nexi()
        f = m.check_set_default(dict(lmax=lmax, start="likelihood"))
        m.updateProperty(f)
        x, S = m.traindata(10000, lmax)
        dims = (6, 5, 2)
        nv = int(np.prod(dims))
        rng = np.random.default_rng(3 + lmax)
        k = rng.choice(10000, nv, replace=False)
        scale = rng.uniform(0.8, 1.0, nv)
        y = (S[:, k] * scale).T.reshape(dims + (S.shape[0],), order="F")
        mask = np.ones(dims, bool)
        mask[0, 0, 0] = False
        p = m.estimate_prior(y, mask, None, lmax)
        bb = m.b[m.b <= max(m.b.min(), 1.1)].astype(np.float64)
        flat = y.reshape(nv, -1, order="F")[:, : m.b.size][:, m.b <= max(m.b.min(), 1.1)]
        csf = (-np.log(flat) @ bb / (bb @ bb)) > 1.5
        mult = np.array([0.01, 1, 1, 0.01, 0.01])
        for j, name in enumerate(m.modelParams[: x.shape[0]]):
            got = p[name].reshape(nv, order="F")
            exp = x[j, k] * np.where(csf, mult[j], 1.0)
            exp[0] = 0.0                                   # outside the mask
            np.testing.assert_allclose(got, exp.astype(np.float32), rtol=1e-6)


def test_nexi_estimate_end_to_end():
    """Raw diffusion volumes (isotropic per shell) -> invariants -> masks ->
    likelihood start -> askadam: the fit lowers the loss of the start."""
    print("nexi estimate", flush=True)
    m0 = _nexi()
    f = m0.check_set_default(dict(lmax=0))
    m0.updateProperty(f)
    names = m0.modelParams
    dims = (5, 4, 2)
    nv = int(np.prod(dims))
    rng = np.random.default_rng(11)
    fa = rng.uniform(0.3, 0.7, nv)
    th = np.stack([fa, rng.uniform(1.8, 2.8, nv), rng.uniform(0.6, 1.2, nv), (1 - fa) / rng.uniform(5, 40, nv)])
    md = models.gpuNEXI(m0.b, m0.Delta).gc_model(dict(modelParams=names, lb=m0.lb, ub=m0.ub), "askadam", None)
    sl0 = orc.fwd_value(md, th)                                   # [8, nv]
    bvals = np.concatenate([[0, 0], np.repeat(m0.b, 6)])
    bdel = np.concatenate([[20, 20], np.repeat(m0.Delta, 6)])
    g = rng.normal(size=(bvals.size, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    s0 = rng.uniform(800, 1200, nv)
    raw = np.concatenate([np.stack([s0, s0], axis=1), np.repeat(sl0.T, 6, axis=1) * s0[:, None]], axis=1)
    raw = raw + rng.normal(0, 2.0, raw.shape)
    raw = raw.reshape(dims + (-1,), order="F")
    mask = np.ones(dims, bool)
    common = dict(lmax=0, verbose=False, tol=0, convergenceValue=-1, lossFunction="L2", initialLearnRate=0.005,
                  start="likelihood")
    ex = dict(bval=bvals, bvec=g.T, BDELTA=bdel)
    o1 = _nexi().estimate(raw, mask, ex, dict(common, iteration=1))
    o2 = _nexi().estimate(raw, mask, ex, dict(common, iteration=400))
    assert o2["mask"].sum() >= 0.9 * nv
    assert o2["final"]["loss"] < o1["final"]["loss"]
    assert set(o2["final"]).issuperset({"fa", "Da", "De", "ra"})


TE = np.arange(2, 25, 3) * 1e-3
FA = np.array([4.0, 10.0, 20.0, 35.0])
TR = 30e-3


def test_joint_estimate_end_to_end():
    """gpuJointR1R2starMapping.estimate on noise-free VFA x multi-echo data:
    DESPOT1 scaling, the closed-form prior, askadam, M0 back to the data
    scale; R1, R2* and M0 within a few percent of the truth."""
    print("joint estimate", flush=True)
    rng = np.random.default_rng(5)
    dims = (6, 5, 2)
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
    m = models.gpuJointR1R2starMapping(TE, TR, FA)
    out = m.estimate(data, np.ones(dims, bool), None, dict(iteration=1500, verbose=False, tol=0,
                                                          convergenceValue=-1, lossFunction="L2",
                                                          initialLearnRate=0.003))
    for name, truth in (("R1", r1), ("R2star", r2), ("M0", m0)):
        rel = np.abs(out["final"][name] - truth) / truth
        assert np.median(rel) < 0.02, (name, np.median(rel))
