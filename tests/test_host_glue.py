"""estimate() glue of gpuJointR1R2starMapping and gpuNEXI on the host
(R1R2s/gpuJointR1R2starMapping.m:105-491, NEXI/gpuNEXI.m:148-849,
NEXI/gpuNEXIrice.m:104-447, utils/DWIutility.m): rotational invariants
against the closed form of a + c P2, the shell table, MATLAB's twister
stream, the NEXI weights and masks, and the joint model's prior."""
import numpy as np
import pytest

from gacelle_amd import dwi, models
from gacelle_amd._abi import GacelleError


def _dirs(n, seed):
    g = np.random.default_rng(seed).normal(size=(n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def _p2(x):
    return 0.5 * (3 * x * x - 1)


def test_rotinv_closed_form_and_basis():
    """S(g) = a + c P2(g . n): F0 = a, F2 = |c| / 5 (P2(g.n) = 4 pi / 5
    sum_m Y2m(g) Y2m(n) and sum_m Y2m(n)^2 = 5 / (4 pi)), for any axis n;
    and the basis is orthonormal on the sphere."""
    g = _dirs(45, 1)
    rng = np.random.default_rng(2)
    for _ in range(4):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        a, c = rng.uniform(0.2, 1.0), rng.uniform(-0.5, 0.5)
        F = dwi.sh_rotinv((a + c * _p2(g @ n))[:, None], g, 2)[:, 0]
        np.testing.assert_allclose(F, [a, abs(c) / 5], rtol=1e-12, atol=1e-13)
    # a degree-4 term leaves F0 and F2 alone at lmax = 4 and shows in F4
    n = np.array([0.0, 0.6, 0.8])
    x = g @ n
    p4 = (35 * x ** 4 - 30 * x ** 2 + 3) / 8
    F = dwi.sh_rotinv((0.5 + 0.2 * _p2(x) + 0.1 * p4)[:, None], g, 4)[:, 0]
    np.testing.assert_allclose(F, [0.5, 0.2 / 5, 0.1 / 9], rtol=1e-11)
    N = 20000
    i = np.arange(N) + 0.5
    Y = dwi.real_sh_even(4, np.pi * (1 + 5 ** 0.5) * i, np.arccos(1 - 2 * i / N))
    np.testing.assert_allclose(Y.T @ Y * 4 * np.pi / N, np.eye(Y.shape[1]), atol=2e-5)


def _dwi_volume(dims=(3, 2, 2), shells=(1.0, 2.5), ndir=30, seed=3):
    rng = np.random.default_rng(seed)
    nvox = int(np.prod(dims))
    bval, bvec = [0.0, 0.0], [np.array([1.0, 0, 0])] * 2
    for k, b in enumerate(shells):
        g = _dirs(ndir, 10 + k)
        bval += [b] * ndir
        bvec += list(g)
    bval = np.array(bval)
    bvec = np.array(bvec).T                                   # [3, N], the reference's layout
    s0 = rng.uniform(500, 1500, nvox)
    n = rng.normal(size=(nvox, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    a = rng.uniform(0.2, 0.9, (nvox, len(shells)))
    c = rng.uniform(0.0, 0.3, (nvox, len(shells)))
    sig = np.zeros((nvox, bval.size))
    sig[:, bval == 0] = s0[:, None]
    for k, b in enumerate(shells):
        sel = bval == b
        sig[:, sel] = s0[:, None] * (a[:, k:k + 1] + c[:, k:k + 1] * _p2(n @ bvec[:, sel]))
    return sig.reshape(dims + (-1,), order="F"), bval, bvec, a, c


def test_rotationally_invariant_signal_layout():
    """One echo time: [Sl0 of each shell, Sl2 of each shell] / b0 (shell
    fastest, l-major), as the NEXI model reads them; the b0-normalised
    variant of the Rician model gives the same values here."""
    sig, bval, bvec, a, c = _dwi_volume()
    out, bo, _, bdo, _ = dwi.compute_rotationally_invariant_signal(sig, bval, bvec, None, 20.0, None, 2)
    flat = out.reshape(-1, out.shape[3], order="F")
    np.testing.assert_allclose(flat, np.concatenate([a, c / 5], axis=1), rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(bo, [1.0, 2.5, 1.0, 2.5])
    np.testing.assert_array_equal(bdo, [20.0] * 4)
    out2, b2 = dwi.get_Sl_all(sig, bval, bvec, 1.0, 20.0, 2)
    np.testing.assert_allclose(out2, out, rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(b2, [1.0, 2.5])


def test_unique_shell_keepb0():
    b = [0, 2, 1, 2, 0, 1, 3]
    D = [20, 30, 20, 20, 20, 30, 30]
    bs, _, ds, _ = dwi.unique_shell_keepb0(b, None, D, None, True)
    np.testing.assert_array_equal(bs, [1, 2, 1, 2, 3])
    np.testing.assert_array_equal(ds, [20, 20, 30, 30, 30])
    bs, _, ds, _ = dwi.unique_shell_keepb0(b, None, D, None, False)
    np.testing.assert_array_equal(bs, [0, 1, 2, 1, 2, 3])
    m = models.gpuNEXI(b, D)
    np.testing.assert_array_equal(m.b, [1, 2, 1, 2, 3])
    np.testing.assert_array_equal(m.Nav, np.ones(5))


def test_twister_stream_is_matlabs():
    """MATLAB's rand after rng('default') (twister, seed 5489) begins
    0.8147, 0.9058, 0.1270 (MATLAB documentation of rand); numpy's
    RandomState gives the same stream, which traindata draws from with
    rng(48463, 'twister')."""
    np.testing.assert_allclose(np.random.RandomState(5489).random_sample(3), [0.8147, 0.9058, 0.1270], atol=5e-5)


def test_nexi_weights():
    m = models.gpuNEXI([1, 2, 3], [20, 20, 20], Nav=[1, 2, 4])
    mask = np.ones((2, 1, 1), bool)
    w2 = m.compute_optimisation_weights(mask, "L2", 2)
    np.testing.assert_allclose(w2[0, 0, 0], np.array([1, 2, 4, 1 / 5, 2 / 5, 4 / 5]) / 4, rtol=1e-6)
    w1 = m.compute_optimisation_weights(mask, "L1", 0)
    np.testing.assert_allclose(w1[1, 0, 0], np.sqrt([1, 2, 4]) / 2, rtol=1e-6)


def test_nexi_prepare_masks():
    """Impossible, background, incoherent and NaN voxels leave the mask."""
    m = models.gpuNEXI([1, 2, 3, 4], [20] * 4)
    rng = np.random.default_rng(4)
    dims = (4, 3, 1)
    base = np.exp(-0.6 * np.array([1, 2, 3, 4]))
    data = base * rng.uniform(0.9, 1.1, dims + (1,)) + rng.normal(0, 1e-3, dims + (4,))
    data[0, 0, 0] = [1.3, 0.5, 0.3, 0.2]          # impossible
    data[1, 0, 0] = [0.005, 0.004, 0.003, 0.002]  # background
    data[2, 0, 0] = base[::-1]                    # rising with b: anti-correlated
    data[3, 0, 0, 2] = np.nan
    mask = np.ones(dims, bool)
    mask[3, 2, 0] = False
    f = m.check_set_default({})
    d2, mk, _ = m.prepare_dwi_data(data, mask, None, 0, f)
    bad = {(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0), (3, 2, 0)}
    for idx in np.ndindex(*dims):
        assert mk[idx] == (idx not in bad), idx
    assert np.all(np.isfinite(d2))
    with pytest.raises(GacelleError):
        m.prepare_dwi_data(data[..., :3], mask, None, 0, f)


TE = np.arange(2, 25, 3) * 1e-3
FA = np.array([4.0, 10.0, 20.0, 35.0])
TR = 30e-3


def _joint(dims=(4, 3, 2), seed=5):
    rng = np.random.default_rng(seed)
    m0 = rng.uniform(800, 1200, dims)
    r1 = rng.uniform(0.6, 1.5, dims)
    r2 = rng.uniform(15, 40, dims)
    e1 = np.exp(-TR * r1)
    sig = []
    for fa in np.float32(FA):
        a = np.deg2rad(fa)
        amp = m0 * np.sin(a) * (1 - e1) / (1 - np.cos(a) * e1)
        sig.append(amp[..., None] * np.exp(-r2[..., None] * TE))
    return np.stack(sig, axis=4), m0, r1, r2


def test_joint_prior_and_prepare():
    """DESPOT1 of the first echo gives R1 exactly and M0 exp(-R2* TE1);
    the data scale is the 98th percentile of that M0; R2* is the closed
    form of the flip-angle mean, clamped; R2* = R1 / 2 where R1 > R2*."""
    data, m0, r1, r2 = _joint()
    m = models.gpuJointR1R2starMapping(TE, TR, FA)
    mask = np.ones(m0.shape, bool)
    p = m.estimate_prior(data, mask, None)
    np.testing.assert_allclose(p["R1"], r1, rtol=1e-5)
    np.testing.assert_allclose(p["M0"], np.minimum(m0 * np.exp(-r2 * TE[0]), 2), rtol=1e-5)
    r2c, _ = models.gpuR2starMapping.R2star_trapezoidal(data.mean(axis=4), TE)
    np.testing.assert_allclose(p["R2star"], np.clip(r2c, 0.1, 200), rtol=1e-5)
    assert np.all(np.abs(p["R2star"] - r2) < 0.1 * r2)
    d2, mk, scale = m.prepare_data(data, mask, None)
    from gacelle_amd import utils
    np.testing.assert_allclose(scale, utils.prctile(m0 * np.exp(-r2 * TE[0]), 98), rtol=1e-6)
    np.testing.assert_allclose(d2, data / scale)
    # R1 above R2*: R2* becomes R1 / 2
    d3 = data.copy()
    d3[0, 0, 0] = d3[0, 0, 0, :1, :]             # flat echoes: R2* = 0 -> lower bound
    p3 = m.estimate_prior(d3, mask, None)
    assert p3["R2star"][0, 0, 0] == pytest.approx(p3["R1"][0, 0, 0] / 2, rel=1e-6)


def test_joint_validate_input():
    data, m0, _, _ = _joint()
    m = models.gpuJointR1R2starMapping(TE, TR, FA)
    mask, ex = m.validate_input(data, None, None)
    a = np.abs(data).max(axis=(3, 4))
    np.testing.assert_array_equal(mask, a / a.max() > 0.05)
    np.testing.assert_array_equal(ex["b1"], np.ones(m0.shape))
    with pytest.raises(GacelleError):
        m.validate_input(data[..., :3], None, None)
    with pytest.raises(GacelleError):
        m.validate_input(data, None, {"other": 1})
    with pytest.raises(GacelleError):
        m.validate_input(data, np.ones((2, 2, 2)), None)
    # a numeric start is ignored: the fixed start points are used
    f = m.check_set_default({"start": [1, 2, 3, 4]})
    f["modelParams"], f["ub"], f["lb"] = m.modelParams, m.ub, m.lb
    x0 = m.determine_x0(data, np.ones(m0.shape, bool), None, f)
    assert np.all(x0["R2star"] == 30) and np.all(x0["R1"] == 1)


def test_nexi_rician_prepare_masks():
    """gpuNEXIrice.prepare_dwi_data (NEXI/gpuNEXIrice.m:338-412): sigma / b0,
    and voxels leave the mask above 1.01, at NaN, at SNR below 1 (sigma > 1
    after normalisation) or when their curve runs against the mean curve."""
    m = models.gpuNEXI([1, 2, 3, 4], [20] * 4, rician=True)
    rng = np.random.default_rng(8)
    dims = (4, 3, 1)
    base = np.exp(-0.6 * np.array([1, 2, 3, 4]))
    data = base * rng.uniform(0.9, 1.0, dims + (1,))
    data[0, 0, 0] = [1.2, 0.5, 0.3, 0.2]          # above 1.01
    data[1, 0, 0, 1] = np.nan
    data[2, 0, 0] = base[::-1] * 0.9              # rising with b
    sigma = np.full(dims, 0.02)
    sigma[3, 0, 0] = 2.0                          # SNR below 1
    ex = dict(bval=np.array([1, 2, 3, 4.0]), sigma=sigma)
    mask = np.ones(dims, bool)
    d2, mk, exn = m.prepare_dwi_data_rice(data, mask, ex, 0)
    bad = {(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0)}
    for idx in np.ndindex(*dims):
        assert mk[idx] == (idx not in bad), idx
    np.testing.assert_allclose(exn["sigma"], sigma)    # no b=0 volume: b0 = 1
    assert np.all(np.isfinite(d2))


def test_gpunexirice_is_the_rician_model_with_the_reference_argument_order(monkeypatch):
    """gpuNEXIrice(b, Delta).fit(dwi, mask, fitting, extraData) (NEXI/
    gpuNEXIrice.m:31, :203): the model carries the Rician mean correction,
    extraData reaches the solver as the FWD's extra data (sigma) and is not
    read as pars0; estimate keeps the (dwi, mask, extraData, fitting) order."""
    import gacelle_amd as ga
    from gacelle_amd import _abi

    b = np.repeat(np.arange(1, 7, dtype=float), 2)
    D = np.tile([20.0, 30.0], 6)
    m = ga.gpuNEXIrice(b, D)
    assert isinstance(m, models.gpuNEXI) and m.rician
    f = m.check_set_default({})
    m.updateProperty(f)
    f["modelParams"], f["lb"], f["ub"] = list(m.modelParams), list(m.lb), list(m.ub)
    assert m.gc_model(f, "askadam", None).flags & _abi.GC_FLAG_NEXI_RICIAN
    assert not (models.gpuNEXI(b, D).gc_model(f, "askadam", None).flags & _abi.GC_FLAG_NEXI_RICIAN)
    seen = {}

    def fake_solve(self, data, mask, w, pars0, fitting, *fwd_args):
        seen["pars0"], seen["fwd_args"] = pars0, fwd_args
        return {}

    monkeypatch.setattr(models.gpuNEXI, "_solve", fake_solve)
    dwi = np.full((2, 2, 1, 12), 0.5, np.float32)
    mask = np.ones((2, 2, 1), bool)
    extra = {"sigma": np.full((2, 2, 1), 0.02, np.float32)}
    m.fit(dwi, mask, {"start": "default"}, extra)
    assert seen["fwd_args"] == (extra,)
    assert set(seen["pars0"]) >= {"fa", "Da", "De", "ra"}
    np.testing.assert_allclose(np.asarray(seen["pars0"]["fa"]), 0.4)   # the fixed start, not sigma
    # gpuNEXI keeps its own order fit(dwi, mask, fitting, pars0, extradata)
    p0 = {k: np.full((2, 2, 1), v, np.float32) for k, v in zip(models.gpuNEXI.modelParams, [0.3, 2, 1, 0.05, 0.2, 0.01])}
    models.gpuNEXI(b, D).fit(dwi, mask, {"start": "default"}, p0, None)
    np.testing.assert_allclose(np.asarray(seen["pars0"]["fa"]), 0.3)
    assert "gpuMCRMWI" in ga.__all__ and ga.gpuMCRMWI is models.gpuMCRMWI


def test_joint_true_flip_angle_rows_and_flag():
    """extraData.trueFlipAngle replaces b1 .* FA in FWD
    (R1R2s/gpuJointR1R2starMapping.m:387-391): one extra row per flip angle
    in G-D order and GC_FLAG_R1R2S_TRUE_FA; given as b1 .* FA the oracle's
    signal equals the b1 form."""
    from gacelle_amd import _abi, utils
    from gacelle_amd.askadam import askadam as A
    from oracle import oracle as orc

    te = np.array([2.0, 3.5, 6.0, 9.0, 13.0]) * 1e-3
    fa = np.array([4.0, 12.0, 20.0])
    m = models.gpuJointR1R2starMapping(te, 25e-3, fa)
    dims = (4, 3, 2)
    rng = np.random.default_rng(3)
    mask = rng.uniform(size=dims) > 0.3
    vox = utils.find_mask(mask)
    nv = vox.size
    b1 = rng.uniform(0.8, 1.2, dims)
    tfa = b1[..., None] * np.deg2rad(fa)  # [X, Y, Z, Nfa]
    f = A.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1], ub=[2, 10, 200]))
    md = m.model_for_extra(m.gc_model(f, "askadam", None), {"trueFlipAngle": tfa})
    assert md.flags & _abi.GC_FLAG_R1R2S_TRUE_FA and md.n_extra == fa.size
    rows = m.extra_rows({"trueFlipAngle": tfa}, mask, vox, nv)
    assert rows.shape == (fa.size, nv)
    np.testing.assert_array_equal(rows, tfa.reshape(-1, fa.size, order="F")[vox].T.astype(np.float32))
    # G-D input ([1, Nv, 1, 1, Nfa], the reference's gpu_reshape_ND2GD_struct form)
    gd = rows.T.reshape(1, nv, 1, 1, fa.size)
    np.testing.assert_array_equal(m.extra_rows({"trueFlipAngle": gd}, mask, vox, nv), rows)
    # without the field: the b1 row and no flag
    md_b1 = m.model_for_extra(m.gc_model(f, "askadam", None), {"b1": b1})
    assert not (md_b1.flags & _abi.GC_FLAG_R1R2S_TRUE_FA) and md_b1.n_extra == 1
    b1_row = m.extra_rows({"b1": b1}, mask, vox, nv)
    th = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
    np.testing.assert_allclose(orc.fwd_value(md, th, rows.astype(np.float64)),
                               orc.fwd_value(md_b1, th, b1_row.astype(np.float64)), rtol=1e-6)
    with pytest.raises(GacelleError):
        m.extra_rows({"trueFlipAngle": tfa[..., :2]}, mask, vox, nv)
