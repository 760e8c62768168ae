"""gpuMCRMWI fit / estimate glue on the host (MCRMWI/gpuMCRMWI.m:217-555,
:873-967, MCRMWI/private/despot1.m): DESPOT1 against the spoiled-GRE
closed form, the default initial phase, the fittable-voxel mask, the
'1stecho' weights and the prior starting points."""
import numpy as np
import pytest

from gacelle_amd import models

TE = np.arange(2, 49, 2) * 1e-3
FA = np.array([5.0, 10.0, 20.0, 50.0])
TR = 38e-3


def _spgr(m0, t1, fa_deg, b1=1.0):
    a = np.deg2rad(fa_deg) * b1
    e1 = np.exp(-TR / t1)
    return m0 * np.sin(a) * (1 - e1) / (1 - np.cos(a) * e1)


def test_despot1_recovers_closed_form():
    m = models.gpuMCRMWI(TE, TR, FA)
    rng = np.random.default_rng(0)
    dims = (4, 3, 2)
    t1 = rng.uniform(0.5, 2.0, dims)
    m0 = rng.uniform(0.5, 1.5, dims)
    b1 = rng.uniform(0.9, 1.1, dims)
    img = np.stack([_spgr(m0, t1, fa, b1) for fa in np.float32(FA)], axis=3)
    mask = np.ones(dims, bool)
    mask[0, 0, 0] = False
    t1e, m0e, fitted = m.despot1(img, mask, b1)
    np.testing.assert_allclose(t1e[mask], t1[mask], rtol=1e-6)
    np.testing.assert_allclose(m0e[mask], m0[mask], rtol=1e-6)
    assert not fitted[0, 0, 0] and t1e[0, 0, 0] == 0 and fitted[mask].all()
    # a zero signal gives T1 = M0 = 0 and is not fitted
    img[1, 1, 1] = 0
    t1z, _, fz = m.despot1(img, mask, b1)
    assert t1z[1, 1, 1] == 0 and not fz[1, 1, 1]


def _volume(dims=(5, 4, 2), freq_per_fa=True, seed=1):
    rng = np.random.default_rng(seed)
    nte, nfa = TE.size, FA.size
    s0 = rng.uniform(800, 1200, dims)
    t1 = rng.uniform(0.8, 1.4, dims)
    r2 = rng.uniform(15, 30, dims)
    fb = rng.uniform(-0.05, 0.05, dims + ((nfa,) if freq_per_fa else ()))
    phase0 = rng.uniform(-1, 1, dims)
    B0g = 3.0 * 42.57747892
    fbk = fb if freq_per_fa else fb[..., None] * np.ones(nfa)
    amp = np.stack([_spgr(s0, t1, fa) for fa in np.float32(FA)], axis=3)     # [x,y,z,nfa]
    sig = amp[:, :, :, None, :] * np.exp(-TE[None, None, None, :, None] * r2[..., None, None]) * \
        np.exp(1j * (2 * np.pi * TE[None, None, None, :, None] * fbk[:, :, :, None, :] * B0g + phase0[..., None, None]))
    return sig, dict(freqBKG=fb), s0, t1, r2, phase0


def test_default_pini_and_mask():
    m = models.gpuMCRMWI(TE, TR, FA)
    f = m.check_set_default({}, None)
    data, ex, s0, t1, r2, phase0 = _volume()
    mask = np.ones(data.shape[:3], bool)
    data[0, 0, 0, 0, 2] = 0          # zero first echo at one flip angle: masked out
    ex2, m2 = m.validate_data(data, ex, mask, f)
    # pini = angle(d(TE1, FA1) / exp(i 2 pi freqBKG B0 gamma TE1)), one map per flip angle
    fb = ex["freqBKG"]
    d11 = data[:, :, :, 0, 0][..., None]
    ref = np.angle(d11 / np.exp(1j * 2 * np.pi * fb * 3.0 * 42.57747892 * TE[0]))
    np.testing.assert_allclose(ex2["pini"], ref.astype(np.float32), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(ex2["pini"][..., 0], phase0.astype(np.float32), atol=1e-5)
    assert not m2[0, 0, 0] and m2.sum() == mask.sum() - 1
    assert np.all(ex2["b1"] == 1)


def test_mask_thresholds():
    m = models.gpuMCRMWI(TE, TR, FA)
    f = m.check_set_default({}, None)
    data, ex, *_ = _volume(freq_per_fa=False)
    dims = data.shape[:3]
    slow = np.stack([_spgr(1000.0, 4.0, fa) for fa in np.float32(FA)])     # T1 4 s > 3.1 s
    data[2, 1, 0] = slow[None, :] * np.exp(-TE * 20.0)[:, None]
    data[3, 1, 0] = data[3, 1, 0, :1] * np.ones((TE.size, 1))            # no decay: R2* = 0 < 2
    _, m2 = m.validate_data(data, ex, np.ones(dims, bool), f)
    assert not m2[2, 1, 0] and not m2[3, 1, 0] and m2.sum() == np.prod(dims) - 2


def test_weights_prior_and_scale():
    m = models.gpuMCRMWI(TE, TR, FA)
    f = m.check_set_default(dict(weightPower=2), None)
    data, ex, s0, t1, r2, _ = _volume()
    w = m.compute_optimisation_weights(data, f)
    assert w.shape == data.shape[:3] + (TE.size, FA.size, 2)
    mag = np.abs(data)
    np.testing.assert_allclose(w[..., 0], np.minimum(mag ** 2 / mag[:, :, :, :1] ** 2, 1), rtol=1e-6)
    np.testing.assert_array_equal(w[..., 0], w[..., 1])
    mask = np.ones(data.shape[:3], bool)
    d2, m2, ex2, scale = m.prepare_data(data, mask, ex, f)
    _, m0, fitted = m.despot1(np.abs(data[:, :, :, 0, :]), m2, None)
    assert scale == pytest.approx(np.percentile(m0[fitted], 98, method="hazen"), rel=1e-9)
    np.testing.assert_allclose(d2, data / scale)
    m.updateProperty(f)
    f["modelParams"], f["lb"], f["ub"] = list(m.modelParams), list(m.lb), list(m.ub)
    p0 = m.estimate_prior(d2, m2, ex2)
    # mono-exponential single pool: the closed-form R2* is the trapezoid estimate of r2
    r2_trap, _ = m.R2star_trapezoidal(np.abs(d2).mean(axis=4), TE)
    np.testing.assert_allclose(p0["R2sIW"], np.clip(r2_trap - 3, m.lb[m.modelParams.index("R2sIW")] + 2,
                                                    m.ub[m.modelParams.index("R2sIW")] - 2).astype(np.float32),
                               rtol=1e-6)
    # S0 prior: DESPOT1 of the per-flip-angle S0 estimates, in the scaled units
    assert np.median(np.abs(p0["S0"] / (s0 / scale) - 1)) < 0.1
    assert p0["dfreqBKG"].shape == data.shape[:3] + (FA.size,)
