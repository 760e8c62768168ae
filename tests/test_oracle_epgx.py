"""CPU tests of the MCR-MWI EPG-X MLP path in the oracle: its features and
networks against a numpy restatement of feature_preprocess_MCRMWI_MLP_EPGX_
leakyrelu.m and mlp_model_leakyRelu.m, its derivatives against central
differences, the composed forward model against the BM path (zero
networks, and constant corrections added in closed form), and the mirror's
network packing."""
import numpy as np
import pytest

from gacelle_amd import _abi, models
from oracle import oracle as orc

TE24 = np.arange(2, 49, 2) * 1e-3
FA4 = np.array([5.0, 10.0, 20.0, 50.0])


def random_net(rng, widths, scale=0.3, bias=0.0):
    """dlnet-style parameters: fc<l>.Weights [out, in], fc<l>.Bias [out]."""
    p = {}
    for l in range(1, len(widths)):
        # FP32 values: the C ABI carries the parameters in FP32
        W = rng.normal(0, scale / np.sqrt(widths[l - 1]), (widths[l], widths[l - 1]))
        b = rng.normal(0, 0.1, widths[l]) + bias
        p[f"fc{l}"] = {"Weights": W.astype(np.float32).astype(float), "Bias": b.astype(np.float32).astype(float)}
    return {"parameters": p, "alpha": 0.01}


def const_net(widths, out):
    p = {}
    for l in range(1, len(widths)):
        p[f"fc{l}"] = {"Weights": np.zeros((widths[l], widths[l - 1])), "Bias": np.zeros(widths[l])}
    p[f"fc{len(widths) - 1}"]["Bias"] = np.asarray(out, float)
    return {"parameters": p}


def np_features(fm, t1iew, kiewmw, fa, tr, t2iew, t1mw=234e-3):
    """feature_preprocess_MCRMWI_MLP_EPGX_leakyrelu.m:24-95, literally."""
    kmwiew = kiewmw * (1 - fm) / fm
    f = np.zeros(11)
    f[0] = fm
    f[1] = np.exp(-tr / t1iew)
    f[2] = np.exp(-tr * kiewmw)
    f[3] = np.sin(fa)
    f[4] = 1 - f[1] * np.cos(fa)
    f[5] = np.exp(-tr * kmwiew)
    f[6] = np.exp(-tr * (1 / t1iew + kiewmw))
    f[7] = np.exp(-tr * (1 / t1mw + kmwiew))
    bloch = lambda m0, t1: m0 * (1 - np.exp(-tr / t1)) * np.sin(fa) / (1 - np.exp(-tr / t1) * np.cos(fa))
    f[8] = bloch(1 - fm, t1iew)
    f[9] = bloch(fm, t1mw)
    f[10] = np.exp(-tr / t2iew)
    return f


def np_mlp(net, x):
    """mlp_model_leakyRelu.m:23-47: fullyconnect, then leakyrelu + fullyconnect."""
    p = net["parameters"]
    a = np.asarray(p["fc1"]["Weights"]) @ x + p["fc1"]["Bias"]
    l = 2
    while f"fc{l}" in p:
        a = np.where(a > 0, a, float(np.float32(0.01)) * a)   # the ABI's FP32 alpha
        a = np.asarray(p[f"fc{l}"]["Weights"]) @ a + p[f"fc{l}"]["Bias"]
        l += 1
    return a


def _model(nets, cplx=True):
    """nets: (phase, magnitude) networks, None for the BM path, or "default"
    for the model's defaults (the reference's trained networks)."""
    m = models.gpuMCRMWI(TE24, 38e-3, FA4)
    fit = dict(isComplex=cplx, isFitExchange=True)
    if nets is None:
        fit["isEPG"] = False
    elif nets != "default":
        fit.update(epgx_phase_ann=nets[0], epgx_mag_ann=nets[1])
    f = m.check_set_default(fit)
    m.updateProperty(f)
    f["modelParams"], f["lb"], f["ub"] = list(m.modelParams), list(m.lb), list(m.ub)
    return m, f, m.gc_model(f, "askadam", None)


def test_epgx_sample_matches_numpy_restatement_and_differences():
    rng = np.random.default_rng(1)
    nets = (random_net(rng, [11, 20, 20, 1]), random_net(rng, [11, 24, 2]))
    m, f, md = _model(nets)
    assert md.flags & _abi.GC_FLAG_MCRMWI_EPG
    args = dict(mvf=0.12, r1iew=1.1, kfr=2.5, r2siw=22.0, fa=np.deg2rad(20.0) * 1.05)
    out = orc.epgx_sample(md, **args)

    def np_out(a):
        x = np_features(a["mvf"], 1 / a["r1iew"], a["kfr"], a["fa"], float(np.float32(38e-3)), 1 / a["r2siw"],
                        float(np.float32(234e-3)))
        return np.concatenate([np_mlp(nets[0], x), np_mlp(nets[1], x)])

    ref = np_out(args)
    np.testing.assert_allclose(out[[0, 5, 10]], ref, rtol=1e-12, atol=1e-14)
    for q, key in enumerate(("mvf", "r1iew", "kfr", "r2siw")):
        h = 1e-6 * max(abs(args[key]), 1)
        a1, a2 = dict(args), dict(args)
        a1[key] += h
        a2[key] -= h
        fd = (np_out(a1) - np_out(a2)) / (2 * h)
        np.testing.assert_allclose(out[[1 + q, 6 + q, 11 + q]], fd, rtol=1e-5, atol=1e-9, err_msg=key)


def _npz_net(path):
    with np.load(path, allow_pickle=False) as d:
        p, l = {}, 1
        while f"fc{l}_Weights" in d:
            p[f"fc{l}"] = {"Weights": d[f"fc{l}_Weights"].astype(float), "Bias": d[f"fc{l}_Bias"].astype(float)}
            l += 1
    return {"parameters": p, "alpha": 0.01}


def test_default_model_runs_the_reference_networks():
    """With no networks given, gpuMCRMWI takes the EPG-X path (isEPG = true,
    MCRMWI/gpuMCRMWI.m:1094) with the reference's trained networks
    (:1104-1105, 8 layers 11-20-30-40-45-60-65-75 then 1 / 2 outputs, as
    extracted into gacelle_amd/data).  The oracle's features, both networks
    and the derivatives of the three outputs match the numpy restatement of
    the feature script and of mlp_model_leakyRelu on those weights, and
    central differences of it, at samples across the parameter ranges."""
    import os

    m, f, md = _model("default")
    assert f["isEPG"] and md.flags & _abi.GC_FLAG_MCRMWI_EPG
    widths = [11, 20, 30, 40, 45, 60, 65, 75]
    for n, nout in ((0, 1), (1, 2)):
        assert md.mlp[n].n_layer == 8
        assert [md.mlp[n].width[l] for l in range(9)] == widths + [nout]
    nets = [_npz_net(os.path.join(models._DATA_DIR, f"mcrmwi_epgx_{k}.npz")) for k in ("phase", "magn")]
    for args in (dict(mvf=0.12, r1iew=1.1, kfr=2.5, r2siw=22.0, fa=np.deg2rad(20.0) * 1.05),
                 dict(mvf=0.05, r1iew=0.8, kfr=1.0, r2siw=15.0, fa=np.deg2rad(5.0) * 0.9),
                 dict(mvf=0.3, r1iew=1.6, kfr=5.0, r2siw=30.0, fa=np.deg2rad(50.0) * 1.1)):
        out = orc.epgx_sample(md, **args)

        def np_out(a):
            x = np_features(a["mvf"], 1 / a["r1iew"], a["kfr"], a["fa"], float(np.float32(38e-3)), 1 / a["r2siw"],
                            float(np.float32(234e-3)))
            return np.concatenate([np_mlp(nets[0], x), np_mlp(nets[1], x)])

        ref = np_out(args)
        np.testing.assert_allclose(out[[0, 5, 10]], ref, rtol=1e-10, atol=1e-12)
        for q, key in enumerate(("mvf", "r1iew", "kfr", "r2siw")):
            h = 1e-6 * max(abs(args[key]), 1)
            a1, a2 = dict(args), dict(args)
            a1[key] += h
            a2[key] -= h
            fd = (np_out(a1) - np_out(a2)) / (2 * h)
            np.testing.assert_allclose(out[[1 + q, 6 + q, 11 + q]], fd, rtol=2e-5, atol=1e-8, err_msg=key)


@pytest.mark.parametrize("cplx", [True, False])
def test_zero_networks_reduce_to_the_bm_path(cplx):
    nets = (const_net([11, 8, 1], [0.0]), const_net([11, 8, 2], [0.0, 0.0]))
    _, f, md_e = _model(nets, cplx)
    _, _, md_b = _model(None, cplx)
    assert not md_b.flags & _abi.GC_FLAG_MCRMWI_EPG
    rng = np.random.default_rng(2)
    nv = 20
    th = np.stack([rng.uniform(0.8, 1.2, nv), rng.uniform(0.05, 0.2, nv), rng.uniform(0.4, 0.8, nv),
                   rng.uniform(0.8, 1.4, nv), rng.uniform(1, 3, nv), rng.uniform(80, 150, nv),
                   rng.uniform(15, 30, nv), rng.uniform(15, 30, nv), rng.uniform(0.02, 0.1, nv),
                   rng.uniform(-0.05, 0, nv)] + ([np.zeros(nv)] * 5 if cplx else []))
    ex = np.zeros((7, nv))
    ex[0], ex[1], ex[2] = rng.uniform(0.9, 1.1, nv), rng.uniform(-0.3, 0.3, nv), 1.0
    np.testing.assert_allclose(orc.fwd_value(md_e, th, ex), orc.fwd_value(md_b, th, ex), rtol=1e-13, atol=1e-15)


def test_constant_corrections_add_in_closed_form():
    """Constant network outputs (phase c, IEW d_i, MW d_m): the complex
    signal is the BM one plus d_m tv rho e^{-te R2sMW} e^{i pMW} and
    (d_i tv) times the IW/EW decays, with the IW and EW phases rotated by c."""
    c, di, dm = (float(np.float32(v)) for v in (0.3, 0.02, -0.01))   # FP32 in the ABI
    nets = (const_net([11, 4, 1], [c]), const_net([11, 4, 2], [di, dm]))
    _, f, md_e = _model(nets)
    _, _, md_b = _model(None)
    nv = 3
    rng = np.random.default_rng(3)
    S0, mwf, iwf = 1.0, 0.1, 0.6
    th = np.array([S0, mwf, iwf, 1.0, 2.0, 100.0, 20.0, 25.0, 0.05, -0.02, 0, 0, 0, 0, 0.1], float)[:, None] * \
        np.ones((1, nv))
    ex = np.zeros((7, nv))
    ex[0], ex[1], ex[2] = 1.0, rng.uniform(-0.3, 0.3, nv), 1.0
    se, sb = orc.fwd_value(md_e, th, ex), orc.fwd_value(md_b, th, ex)
    rho = md_b.f[3]
    tv = S0 * (1 - mwf) + S0 * mwf / rho
    B0g = md_b.f[1] * md_b.f[2]
    nte, nfa = TE24.size, FA4.size
    te = np.asarray([md_b.table[t] for t in range(nte)])[:, None]
    ze, zb = se[: nte * nfa] + 1j * se[nte * nfa:], sb[: nte * nfa] + 1j * sb[nte * nfa:]
    pini = ex[1][None] + 0.1          # extraData.pini + dpini
    for fi in range(nfa):
        sl = slice(fi * nte, (fi + 1) * nte)
        w = 2 * np.pi * te * B0g
        eMW = np.exp(-te * 100.0) * np.exp(1j * (w * 0.05 + pini))
        eIW = np.exp(-te * 20.0) * np.exp(1j * (w * -0.02 + pini))
        eEW = np.exp(-te * 25.0) * np.exp(1j * (w * md_b.f[5] + pini))
        # BM amplitudes from the BM signal: solve for aIW + aEW split by iwf
        # through the three decays (least squares on the echoes)
        A = np.stack([eMW[:, 0], iwf * eIW[:, 0] + (1 - iwf) * eEW[:, 0]], axis=1)
        amp = np.linalg.lstsq(A, zb[sl, 0], rcond=None)[0]
        exp_e = (amp[0] + dm * tv * rho) * eMW[:, 0] + (amp[1] + di * tv) * np.exp(1j * c) * \
            (iwf * eIW[:, 0] + (1 - iwf) * eEW[:, 0])
        np.testing.assert_allclose(ze[sl, 0], exp_e, rtol=1e-9, atol=1e-12)


def test_epgx_model_gradient_matches_central_differences():
    rng = np.random.default_rng(4)
    nets = (random_net(rng, [11, 16, 16, 1]), random_net(rng, [11, 16, 16, 2], scale=0.1))
    m, f, md = _model(nets)
    nv = 4
    th = np.stack([rng.uniform(0.8, 1.2, nv), rng.uniform(0.05, 0.2, nv), rng.uniform(0.4, 0.8, nv),
                   rng.uniform(0.8, 1.4, nv), rng.uniform(1, 3, nv), rng.uniform(80, 150, nv),
                   rng.uniform(15, 30, nv), rng.uniform(15, 30, nv), rng.uniform(0.02, 0.1, nv),
                   rng.uniform(-0.05, 0, nv)] + [rng.uniform(-0.02, 0.02, nv) for _ in range(5)])
    ex = np.zeros((7, nv))
    ex[0], ex[1], ex[2] = rng.uniform(0.9, 1.1, nv), rng.uniform(-0.3, 0.3, nv), 1.0
    s, jac = orc.fwd(md, th, ex)
    for p in range(th.shape[0]):
        h = 1e-6 * max(abs(th[p]).max(), 1.0)
        a, b = th.copy(), th.copy()
        a[p] += h
        b[p] -= h
        fd = (orc.fwd_value(md, a, ex) - orc.fwd_value(md, b, ex)) / (2 * h)
        np.testing.assert_allclose(jac[:, p], fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max() + 1e-12, err_msg=str(p))


def test_mlp_pack_layout_and_checks(tmp_path):
    rng = np.random.default_rng(5)
    net = random_net(rng, [11, 6, 2])
    widths, alpha, packed = models.mlp_pack(net, 2)
    assert widths == [11, 6, 2] and alpha == 0.01
    W1 = np.asarray(net["parameters"]["fc1"]["Weights"], np.float32)
    np.testing.assert_array_equal(packed[:66], W1.ravel(order="F"))
    np.testing.assert_array_equal(packed[66:72], np.float32(net["parameters"]["fc1"]["Bias"]))
    path = str(tmp_path / "net.npz")
    np.savez(path, fc1_Weights=W1, fc1_Bias=net["parameters"]["fc1"]["Bias"],
             fc2_Weights=net["parameters"]["fc2"]["Weights"], fc2_Bias=net["parameters"]["fc2"]["Bias"])
    w2, a2, p2 = models.mlp_pack(path, 2)
    assert w2 == widths and a2 == 0.01
    np.testing.assert_array_equal(p2, packed)
    with pytest.raises(_abi.GacelleError, match="output"):
        models.mlp_pack(net, 1)
    with pytest.raises(_abi.GacelleError, match="isEPG needs fitting"):
        m = models.gpuMCRMWI(TE24, 38e-3, FA4)
        m.gc_model(dict(isEPG=True, isFitExchange=True, modelParams=m.modelParams, lb=m.lb, ub=m.ub), "askadam", None)
