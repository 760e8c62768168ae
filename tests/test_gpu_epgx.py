"""GPU parity of the MCR-MWI EPG-X MLP path (isEPG): the k_epgx pre-pass
(features, both networks, their Jacobians) and the EPG model's forward,
adjoint and askadam trajectories against the FP64 oracle through the C ABI;
gpuMCRMWI.fit with caller-supplied networks end to end."""
import numpy as np
import pytest

from gacelle_amd import _abi, models, utils
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import logit_start
from .test_oracle_epgx import FA4, TE24, random_net

pytestmark = pytest.mark.gpu


def _epg(cplx=True, widths=((11, 24, 24, 1), (11, 24, 24, 2)), seed=7):
    """widths None: the reference's trained networks, the model's default."""
    m = models.gpuMCRMWI(TE24, 38e-3, FA4)
    if widths is None:
        f = m.check_set_default(dict(isComplex=cplx, isFitExchange=True))
    else:
        rng = np.random.default_rng(seed)
        nets = (random_net(rng, list(widths[0])), random_net(rng, list(widths[1]), scale=0.1))
        f = m.check_set_default(dict(isComplex=cplx, isFitExchange=True, epgx_phase_ann=nets[0],
                                     epgx_mag_ann=nets[1]))
    assert f["isEPG"]
    m.updateProperty(f)
    f["modelParams"], f["lb"], f["ub"] = list(m.modelParams), list(m.lb), list(m.ub)
    md = m.gc_model(f, "askadam", None)          # host network pointers: the oracle's form
    return m, f, md


def _device_md(m, md, keep):
    mdd = _abi.gc_model.from_buffer_copy(md)
    return m.bind_device(mdd, gpu_abi.ctx(), keep)


def _theta(m, f, nv, rng):
    th = []
    for name, sl in m.param_slots(f):
        k = m.modelParams.index(name)
        lo, hi = m.lb[k], m.ub[k]
        th.append(lo + (hi - lo) * rng.uniform(0.3, 0.7, (sl.stop - sl.start, nv)))
    return np.concatenate(th, axis=0).astype(np.float32)


def _extra(m, nv, rng):
    ex = np.zeros((3 + m.fa.size, nv), np.float32)
    ex[0] = rng.uniform(0.9, 1.1, nv)
    ex[1] = rng.uniform(-0.5, 0.5, nv)
    ex[2] = 1.0
    ex[3:] = rng.uniform(-0.05, 0.05, (m.fa.size, nv))
    return ex


@pytest.mark.parametrize("widths", [((11, 24, 24, 1), (11, 24, 24, 2)), ((11, 40, 40, 40, 1), (11, 64, 2)),
                                    ((11, 8, 1), (11, 12, 2)), ((11, 80, 80, 80, 80, 1), (11, 80, 2)),
                                    None])
def test_epgx_forward_matches_oracle(widths):
    """gc_fwd runs the pre-pass then the EPG model: networks of one to five
    16-unit tiles per layer (the widest the ABI takes, 80 units), and the
    reference's trained 8-layer networks (None)."""
    print("epgx forward", widths, flush=True)
    m, f, md = _epg(True, widths)
    rng = np.random.default_rng(4)
    nv = 300
    th = _theta(m, f, nv, rng)
    ex = _extra(m, nv, rng)
    keep = []
    g = gpu_abi.fwd(_device_md(m, md, keep), th, ex)
    o = orc.fwd_value(md, th, ex)
    scale = np.abs(o).max()
    assert np.abs(g - o).max() / scale < 3e-5
    # the networks matter: the BM path differs
    f_bm = dict(f, isEPG=False)
    bm = models.gpuMCRMWI(TE24, 38e-3, FA4).gc_model(f_bm, "askadam", None)
    assert np.abs(orc.fwd_value(bm, th, ex) - o).max() / scale > 1e-3


@pytest.mark.parametrize("nets", ["random", "reference"])
@pytest.mark.parametrize("cplx", [True, False])
@pytest.mark.parametrize("loss", [_abi.GC_LOSS_L1, _abi.GC_LOSS_L2])
def test_epgx_adjoint_matches_oracle_gradient(cplx, loss, nets):
    print("epgx adjoint", cplx, loss, nets, flush=True)
    m, f, md = _epg(cplx, None if nets == "reference" else ((11, 24, 24, 1), (11, 24, 24, 2)))
    rng = np.random.default_rng(5)
    nv = 200
    th = _theta(m, f, nv, rng)
    ex = _extra(m, nv, rng)
    y = (orc.fwd_value(md, _theta(m, f, nv, rng), ex) + 0.01 * rng.standard_normal((md.n_meas, nv)))
    y = y.astype(np.float32)
    z = logit_start(th.astype(np.float64), md)
    keep = []
    g = gpu_abi.fwd_adj(_device_md(m, md, keep), loss, _abi.GC_TRANSFORM_SIGMOID, z, y, extra=ex, scale=1.0)
    P = md.n_par
    lb = np.array([md.lb[p] for p in range(P)])[:, None]
    ub = np.array([md.ub[p] for p in range(P)])[:, None]
    sg = 1 / (1 + np.exp(-z.astype(np.float64)))
    s, jac = orc.fwd(md, lb + (ub - lb) * sg, ex)
    r = s - y
    dl = np.sign(r) if loss == _abi.GC_LOSS_L1 else 2 * r
    ref = np.einsum("mv,mpv->pv", dl, jac) * (ub - lb) * sg * (1 - sg)
    for p in range(P):
        err = np.abs(g["grad"][p] - ref[p]) / max(np.abs(ref[p]).max(), 1e-12)
        assert np.quantile(err, 0.99) < 1e-4 and err.max() < 5e-3, (p, err.max())


@pytest.mark.parametrize("graphs,nets", [(1, "random"), (0, "random"), (1, "reference")])
def test_askadam_epgx_matches_oracle(graphs, nets):
    """60 Adam iterations of the step form with the pre-pass before every
    step kernel (graph replay and direct launches; the reference's networks)."""
    print("epgx askadam", graphs, nets, flush=True)
    m, f, md = _epg(True, None if nets == "reference" else ((11, 24, 24, 1), (11, 24, 24, 2)))
    dims = (8, 7, 3)
    rng = np.random.default_rng(21)
    mask = rng.random(dims) < 0.8
    vox = utils.find_mask(mask)
    nv = vox.size
    th = _theta(m, f, nv, rng)
    ex = _extra(m, nv, rng)
    y = (orc.fwd_value(md, th, ex) + 0.005 * rng.standard_normal((md.n_meas, nv))).astype(np.float32)
    f2 = askadam.check_set_default_basic(dict(f, iteration=60, lossFunction="L2", initialLearnRate=0.01, tol=0,
                                              convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0))
    cfg = askadam().make_cfg(f2, m)
    z0 = logit_start(_theta(m, f, nv, np.random.default_rng(22)).astype(np.float64), md)
    keep = []
    c = gpu_abi.ctx()
    c.set_option(_abi.GC_OPT_GRAPHS, graphs)
    try:
        got = gpu_abi.askadam(_device_md(m, md, keep), cfg, [dict(y=y, z0=z0, extra=ex, vox=vox, dims=dims)],
                              trace=True)[0]
    finally:
        c.set_option(_abi.GC_OPT_GRAPHS, 1)
    ref = orc.askadam(md, cfg, y, z0, extra=ex, vox=vox, dims=dims, trace=True)
    assert got["n_iteration"] == ref["n_iteration"] == 60
    np.testing.assert_allclose(got["trace"], ref["trace"], rtol=1e-4)
    P = md.n_par
    lb = np.array([md.lb[p] for p in range(P)])[:, None]
    ub = np.array([md.ub[p] for p in range(P)])[:, None]
    tg = lb + (ub - lb) / (1 + np.exp(-got["z_final"].astype(np.float64)))
    to = lb + (ub - lb) / (1 + np.exp(-ref["z_final"]))
    assert (np.abs(tg - to) / (ub - lb)).max() < 1e-4
    # the residual outputs come from the pre-pass at the final parameters too
    rf = (orc.fwd_value(md, tg, ex) - y) ** 2
    np.testing.assert_allclose(got["residual"], rf, rtol=2e-3, atol=1e-6 * rf.max())


def test_estimate_end_to_end_with_networks(tmp_path):
    """gpuMCRMWI.estimate with isEPG and the networks read from .npz files
    (fc<l>_Weights / fc<l>_Bias): scaling, validation, prior, weights and the
    askadam fit with the pre-pass run end to end, and the fit lowers the loss
    well below the starting point's."""
    print("epgx estimate", flush=True)
    mt, f, md = _epg(True)
    nets = {}
    for key in ("epgx_phase_ann", "epgx_mag_ann"):
        path = str(tmp_path / f"{key}.npz")
        p = f[key]["parameters"]
        np.savez(path, **{f"{k}_{n}": np.asarray(v[n]) for k, v in p.items() for n in ("Weights", "Bias")})
        nets[key] = path
    dims = (6, 5, 2)
    nv = int(np.prod(dims))
    rng = np.random.default_rng(31)
    th = _theta(mt, f, nv, rng).astype(np.float64)
    k = mt.modelParams.index("S0")
    th[k] = rng.uniform(0.8, 1.2, nv) * 1000.0
    md.lb[k], md.ub[k] = 0.0, 4000.0
    fb = rng.uniform(-0.05, 0.05, dims + (4,))
    ex = np.zeros((3 + 4, nv))
    ex[0], ex[2] = 1.0, 1.0
    ex[3:7] = fb.reshape(nv, 4, order="F").T
    s = orc.fwd_value(md, th, ex)
    nte, nfa = 24, 4
    zc = (s[: nte * nfa] + 1j * s[nte * nfa:]).reshape(nte, nfa, nv)
    img = np.transpose(zc, (2, 0, 1)).reshape(dims + (nte, nfa), order="F")
    img = img + 2.0 * (rng.standard_normal(img.shape) + 1j * rng.standard_normal(img.shape))
    mask = np.ones(dims, bool)
    common = dict(isComplex=True, isFitExchange=True, isEPG=True, initialLearnRate=0.01, lossFunction="L2",
                  verbose=False, tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0, **nets)
    o1 = models.gpuMCRMWI(TE24, 38e-3, FA4).estimate(img, mask, dict(freqBKG=fb), dict(common, iteration=1))
    o2 = models.gpuMCRMWI(TE24, 38e-3, FA4).estimate(img, mask, dict(freqBKG=fb), dict(common, iteration=400))
    assert o2["final"]["S0"].shape == dims
    assert o2["final"]["loss"] < 0.8 * o1["final"]["loss"], (o1["final"]["loss"], o2["final"]["loss"])
