"""Host logic without a GPU: MATLAB layout helpers, option defaults, the
mapping of options onto the C configuration, and the posterior metrics."""
import numpy as np
import pytest

from gacelle_amd import _abi, models, utils
from gacelle_amd.askadam import askadam
from gacelle_amd.mcmc import mcmc

from .helpers import matlab_iqr as _matlab_iqr, matlab_mode_bins as _matlab_mode_bins


def test_find_mask_is_column_major():
    m = np.zeros((3, 2, 2), bool)
    m[1, 0, 0] = m[0, 1, 0] = m[2, 1, 1] = True
    # MATLAB: find on a 3x2x2 array, 1-based linear = i + 3 (j-1) + 6 (k-1)
    np.testing.assert_array_equal(utils.find_mask(m), [1, 3, 11])


def test_nd_gd_round_trip_and_complex_split():
    rng = np.random.default_rng(0)
    img = rng.standard_normal((4, 3, 2, 5)) + 1j * rng.standard_normal((4, 3, 2, 5))
    mask = rng.random((4, 3, 2)) < 0.6
    gd = utils.reshape_ND2GD(img, mask)
    assert gd.shape == (10, mask.sum())
    idx = utils.find_mask(mask)
    flat = img.reshape(24, 5, order="F")[idx]
    np.testing.assert_array_equal(gd[:5], flat.real.T)
    np.testing.assert_array_equal(gd[5:], flat.imag.T)
    back = utils.reshape_GD2ND(gd[:5], mask)
    np.testing.assert_array_equal(back[mask], img.real[mask])
    assert np.all(back[~mask] == 0)


def test_masking_preserve_round_trip():
    rng = np.random.default_rng(1)
    img = rng.standard_normal((3, 4, 2, 6, 2))
    mask = rng.random((3, 4, 2)) < 0.5
    gd = utils.masking_ND2GD_preserve(img, mask)
    assert gd.shape == (1, mask.sum(), 1, 6, 2)
    back = utils.undo_masking_ND2GD_preserve(gd, mask)
    np.testing.assert_array_equal(back[mask], img[mask])


def test_remove_img_naninf_drops_voxels():
    img = np.ones((2, 2, 1, 3))
    img[0, 1, 0, 2] = np.nan
    img[1, 1, 0, 0] = np.inf
    data, mask = utils.remove_img_naninf(img, np.ones((2, 2, 1)))
    assert np.all(np.isfinite(data)) and data[0, 1, 0, 2] == 0
    np.testing.assert_array_equal(mask[..., 0], [[True, False], [True, False]])


@pytest.mark.parametrize("x,p,ref", [
    ([1, 2, 3, 4], 50, 2.5), (list(range(1, 11)), 25, 3.0), (list(range(1, 11)), 1, 1.0),
    ([5.0], 75, 5.0), ([3, 1, 2], 50, 2.0), ([1, 2, 3, 4], 90, 4.0), ([1, 2, 3, 4], 30, 1.7),
])
def test_prctile_matlab_convention(x, p, ref):
    """MATLAB prctile places sorted sample k at 100 (k - 0.5)/n and
    interpolates linearly, clamping outside [50/n, 100 - 50/n]."""
    assert utils.prctile(x, p) == pytest.approx(ref)


def test_posterior_iqr_and_mode():
    a = np.array([[1.0, 2, 3, 4], [0.1, 0.1, 0.5, 0.9]])
    np.testing.assert_allclose(_matlab_iqr(a), [2.0, 0.6])
    # bins of width ~0.1 on [0, 1]: 0.1 occurs twice -> its bin centre
    mode = _matlab_mode_bins(a[1:], 0.0, 1.0, 11)
    assert mode[0] == pytest.approx(0.15, abs=1e-7)
    # ties resolve to the smallest bin, as MATLAB mode does
    mode = _matlab_mode_bins(np.array([[0.95, 0.05]]), 0.0, 1.0, 11)
    assert mode[0] == pytest.approx(0.05, abs=1e-7)


def test_askadam_defaults_follow_reference():
    """utils/askadam.m:759-815."""
    f = askadam.check_set_default_basic({})
    expect = dict(optimiser="adam", initialLearnRate=0.001, decayRate=0, lossFunction="L1", tol=1e-3,
                  iteration=1e4, convergenceValue=1e-6, patience=5, patienceConvergence=5, convergenceModel="ema",
                  convergenceWindow=20, emaDecay=0.95, robustConvergence=False, adamupdateGradDecay=0.9,
                  adamupdateSqGradDecay=0.999, adamupdateEpsilon=1e-8, parameterTransform="sigmoid",
                  convergenceStepTol=1e-6, convergenceGradTol=1e-6)
    for k, v in expect.items():
        assert f[k] == v, k
    f = askadam.check_set_default_basic(dict(Nepoch=77, patience=9))
    assert f["iteration"] == 77 and "Nepoch" not in f and f["patienceGrad"] == 9


def test_mcmc_defaults_follow_reference():
    """utils/mcmc.m check_set_default_basic."""
    f = mcmc.check_set_default_basic(dict(algorithm="GW", metric="mode"))
    assert f["algorithm"] == "ensemble" and f["metric"] == ["mode"] and f["Nbin"] == 1001
    assert (f["iteration"], f["thinning"], f["burnin"], f["Nwalker"], f["StepSize"]) == (2e5, 20, 0.1, 50, 2)


def test_make_cfg_maps_options_onto_the_c_struct():
    f = askadam.check_set_default_basic(dict(optimiser="rmsprop", lossFunction="huber", convergenceModel="linear",
                                             iteration=321, initialLearnRate=0.02, decayRate=0.01,
                                             parameterTransform="linear"))
    c = askadam().make_cfg(f, None)
    assert (c.iteration, c.optimiser, c.loss, c.convergence_model, c.transform) == (
        321, _abi.GC_OPT_RMSPROP, _abi.GC_LOSS_HUBER, _abi.GC_CONV_LINEAR, _abi.GC_TRANSFORM_LINEAR)
    assert (c.lr0, c.decay, c.rms_decay, c.rms_eps) == (0.02, 0.01, 0.9, 1e-8)


def test_r2star_model_flags_uniform_echo_spacing():
    md = models.gpuR2starMapping(np.arange(2, 48, 4) * 1e-3).gc_model(
        askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0], ub=[2, 200])), "askadam", None)
    assert md.flags & _abi.GC_FLAG_UNIFORM_TE and md.f[8] == pytest.approx(4e-3)
    md = models.gpuR2starMapping(np.array([2, 5, 9, 14]) * 1e-3).gc_model(
        askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0], ub=[2, 200])), "askadam", None)
    assert not md.flags & _abi.GC_FLAG_UNIFORM_TE


def test_mcmc_global_detection_and_errors():
    """utils/mcmc.m:54-72: a one-element field after the first marks global
    parameters; they need the ensemble sampler."""
    from gacelle_amd._abi import GacelleError

    te = np.arange(2, 48, 4) * 1e-3
    data = np.ones((2, 2, 1, 12), np.float32)
    mask = np.ones((2, 2, 1))
    f = dict(modelParams=["M0", "R2star", "noise"], lb=[0, 0, 0.001], ub=[2, 200, 0.1], algorithm="MH")
    pars0 = dict(M0=np.ones((2, 2, 1)), R2star=np.ones((2, 2, 1)), noise=np.float32(0.01))
    with pytest.raises(GacelleError, match="ensemble"):
        mcmc().optimisation(data, mask, None, pars0, f, models.gpuR2starMapping(te))
    bad = dict(M0=np.float32(1), R2star=np.ones((2, 2, 1)), noise=np.ones((2, 2, 1)))
    with pytest.raises(GacelleError, match="global"):
        mcmc().optimisation(data, mask, None, bad, dict(f, algorithm="GW"), models.gpuR2starMapping(te))
    g = mcmc.check_global_constant(dict(a=np.ones(3), b=np.float32(1), c=np.ones(1)), ["a", "b", "c"])
    assert list(g) == [False, True, True]


def test_whole_grid_parameters_with_a_regulariser_are_refused():
    """isOptimiseMemory = false with TV (utils/spatial_total_variation.m:32-48
    on whole-grid maps) is refused before any device work, not silently fitted
    in the masked form."""
    import pytest

    from gacelle_amd import GacelleError, _abi, models
    from gacelle_amd.askadam import askadam

    m = models.gpuR2starMapping(np.arange(2, 47, 4) * 1e-3)
    f = {"modelParams": ["M0", "R2star"], "lb": [0, 0.1], "ub": [2, 200], "lambda": [1e-3], "regmap": ["R2star"],
         "isOptimiseMemory": False}
    with pytest.raises(GacelleError) as e:
        askadam().optimisation(np.zeros((2, 2, 1, 12), np.float32), np.ones((2, 2, 1), bool), None,
                               {"M0": np.ones((2, 2, 1)), "R2star": np.ones((2, 2, 1))}, f, m)
    assert e.value.code == _abi.GC_ERR_UNSUPPORTED
