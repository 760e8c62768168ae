"""Model classes: host-side mirrors of the reference's model plugins.

Each class keeps the reference's parameter contract (modelParams, ub, lb,
startPoint, step), its fit/estimate glue and its FWD signature, and
describes its acquisition to the HIP library as a gc_model (the forward
model and its adjoint run in the library's fused kernels).

  gpuR2starMapping         R2star/gpuR2starMapping.m
  gpuJointR1R2starMapping  R1R2s/gpuJointR1R2starMapping.m
  gpuNEXI                  NEXI/gpuNEXI.m (+ gpuNEXIrice.m Rician variant)
"""
from __future__ import annotations

import copy
import ctypes

import numpy as np

from . import _abi, runtime, utils
from ._abi import GacelleError


class _ModelBase:
    modelParams: list
    ub: list
    lb: list
    startPoint: list
    step: list
    model_id = 0

    def _init_contract(self):
        # instance copies: updateProperty prunes them (handle semantics of the
        # reference, e.g. R2star/gpuR2starMapping.m:78-90)
        self.modelParams = list(type(self).modelParams)
        self.ub = list(type(self).ub)
        self.lb = list(type(self).lb)
        self.startPoint = list(type(self).startPoint)
        self.step = list(type(self).step)

    def updateProperty(self, fitting):
        if str(fitting.get("solver", "askadam")).lower() != "mcmc" and "noise" in self.modelParams:
            idx = self.modelParams.index("noise")
            for a in ("modelParams", "lb", "ub", "startPoint", "step"):
                getattr(self, a).pop(idx)
        return self

    # -- library description --------------------------------------------
    def _base_model(self, fitting) -> _abi.gc_model:
        md = _abi.gc_model()
        md.model_id = self.model_id
        names = list(fitting["modelParams"])
        lb = list(np.atleast_1d(fitting["lb"]).astype(float))
        ub = list(np.atleast_1d(fitting["ub"]).astype(float))
        slots = self.param_slots(fitting)
        P = slots[-1][1].stop if slots else 0
        md.n_par = P
        for name, sl in slots:
            k = names.index(name)
            for s in range(sl.start, sl.stop):
                md.lb[s] = lb[k]
                md.ub[s] = ub[k]
        return md

    def param_slots(self, fitting):
        """(name, slice of parameter slots) in modelParams order."""
        out, s = [], 0
        for name in fitting["modelParams"]:
            n = self._param_size(name)
            out.append((name, slice(s, s + n)))
            s += n
        return out

    def _param_size(self, name) -> int:
        return 1

    def extra_rows(self, extraData, mask, vox, nv):
        return None

    def FWD(self, pars, *args):
        """Forward signal [M, Nv] of G-D parameter maps (pars: dict of [1, Nv]
        arrays) computed by the library's gc_fwd."""
        fitting = {"modelParams": list(pars.keys()),
                   "lb": [self.lb[self.modelParams.index(k)] if k in self.modelParams else 0 for k in pars],
                   "ub": [self.ub[self.modelParams.index(k)] if k in self.modelParams else 1 for k in pars]}
        extra = args[0] if args and isinstance(args[0], dict) else None
        md = self.gc_model(fitting, "askadam", None)
        first = np.asarray(next(iter(pars.values())))
        nv = first.size
        theta = np.concatenate([np.asarray(pars[k], dtype=np.float32).reshape(-1, nv) for k in fitting["modelParams"]],
                               axis=0)
        ctx = runtime.context()
        th_d = ctx.to_device(theta)
        s_d = ctx.empty((md.n_meas, nv))
        ex = self.extra_rows(extra, np.ones(nv, bool), np.arange(nv, dtype=np.int64), nv)
        ex_d = ctx.to_device(ex) if ex is not None else None
        ctx.check(ctx.lib.gc_fwd(ctx.handle, ctypes.byref(md), runtime.ptr(th_d), runtime.ptr(ex_d),
                                 runtime.ptr(s_d), nv))
        return runtime.to_host(s_d)

    # -- shared glue ------------------------------------------------------
    def determine_x0(self, y, mask, fitting):
        """determine_x0 (e.g. R2star/gpuR2starMapping.m:257-290)."""
        dims = utils.dims3(mask)
        start = fitting.get("start", "prior")
        if isinstance(start, str):
            if start.lower() == "prior":
                x0 = self.estimate_prior(y, mask, fitting)
            else:
                x0 = utils.initialise_x0(dims, self.modelParams, self.startPoint)
        else:
            x0 = utils.initialise_x0(dims, self.modelParams, list(np.atleast_1d(start)))
        from .askadam import askadam as _ask

        return _ask.set_boundary(x0, fitting["ub"], fitting["lb"])

    def estimate_prior(self, y, mask, fitting):
        return utils.initialise_x0(utils.dims3(mask), self.modelParams, self.startPoint)

    def _solve(self, data, mask, w, pars0, fitting, *fwd_args):
        from .askadam import askadam as _ask
        from .mcmc import mcmc as _mcmc

        if str(fitting["solver"]).lower() == "mcmc":
            fitting = dict(fitting)
            fitting["xStepSize"] = list(self.step)
            return _mcmc().optimisation(data, mask, w, pars0, fitting, self, *fwd_args)
        return _ask().optimisation(data, mask, w, pars0, fitting, self, *fwd_args)


def uniform_spacing(te):
    """Spacing of equally spaced echo times, or None.  Spacings may differ by
    the float32 rounding of the echo times (up to 1e-5 of the spacing: an
    exponent error below 1e-5 * R * spacing)."""
    te = np.asarray(te, dtype=np.float64).ravel()
    if te.size < 3:
        return None
    d = np.diff(te)
    if np.all(np.abs(d - d[0]) <= 1e-5 * max(abs(d[0]), 1e-12)) and d[0] > 0:
        return float(np.float32(te[1]) - np.float32(te[0]))
    return None


def _fill_table(md, values, offset=0):
    for k, v in enumerate(values):
        md.table[offset + k] = float(v)
    return offset + len(values)


# ===========================================================================
class gpuR2starMapping(_ModelBase):
    """Mono-exponential R2* mapping, R2star/gpuR2starMapping.m."""

    modelParams = ["M0", "R2star", "noise"]
    ub = [2, 200, 0.1]
    lb = [0, 0.1, 0.001]
    startPoint = [1, 30, 0.05]
    step = [0.01, 1, 0.005]
    model_id = _abi.GC_MODEL_R2STAR
    B0 = 3
    seed = 48463

    def __init__(self, te):
        self.te = np.asarray(te, dtype=np.float32).ravel()
        self._init_contract()

    def gc_model(self, fitting, solver, nmeas):
        md = self._base_model(fitting)
        md.i[0] = self.te.size
        md.n_meas = self.te.size
        _fill_table(md, self.te)
        dte = uniform_spacing(self.te)
        if dte is not None:
            md.flags |= _abi.GC_FLAG_UNIFORM_TE
            md.f[8] = dte
        if nmeas is not None and nmeas != md.n_meas:
            raise GacelleError(_abi.GC_ERR_INPUT_MISMATCH, "The size of TE does not match with the 4th dimension of the image.")
        return md

    @staticmethod
    def check_set_default(fitting):
        from .askadam import askadam as _ask
        from .mcmc import mcmc as _mcmc

        fitting = dict(fitting or {})
        if "solver" not in fitting:
            fitting["solver"] = "askadam"
        f2 = _mcmc.check_set_default_basic(fitting) if str(fitting["solver"]).lower() == "mcmc" else \
            _ask.check_set_default_basic(fitting)
        for k, v in (("weightMethod", "1stecho"), ("isWeighted", False), ("weightPower", 2), ("start", "prior")):
            if k not in fitting:
                f2[k] = v
        return f2

    @staticmethod
    def R2star_trapezoidal(img, te):
        """Closed-form R2* (R2star/gpuR2starMapping.m:436-459)."""
        img = np.abs(np.asarray(img, dtype=np.float64))
        te = np.asarray(te, dtype=np.float64).ravel()
        n = img.shape[3]
        temp = 0
        for k in range(n - 1):
            temp = temp + 0.5 * (img[:, :, :, k] + img[:, :, :, k + 1]) * (te[k + 1] - te[k])
        with np.errstate(divide="ignore", invalid="ignore"):
            t2s = temp / (img[:, :, :, 0] - img[:, :, :, -1])
            r2s = 1.0 / t2s
            s0 = img[:, :, :, 0] * np.exp(r2s * te[0])
        return r2s, s0

    def estimate_prior(self, data, mask, fitting):
        """Closed-form starting point (R2star/gpuR2starMapping.m:293-316).
        The reference indexes lb(3)/ub(2) after updateProperty pruned
        'noise' (quirk Q1); this mirror uses the R2star bounds."""
        data = np.asarray(data)
        dims = utils.dims3(data)
        pars0 = {p: np.full(dims, self.startPoint[k], dtype=np.float32) for k, p in enumerate(self.modelParams)}
        r2s0, _ = self.R2star_trapezoidal(np.abs(data).reshape(dims + (data.shape[3], -1), order="F").mean(axis=4),
                                          self.te)
        k = self.modelParams.index("R2star")
        bad = ~np.isfinite(r2s0)
        r2s0[bad] = self.lb[k]
        r2s0[r2s0 < self.lb[k]] = self.lb[k]
        r2s0[r2s0 > self.ub[k]] = self.ub[k]
        pars0["M0"] = np.abs(data[:, :, :, 0]).astype(np.float32)
        pars0["R2star"] = r2s0.astype(np.float32)
        return pars0

    @staticmethod
    def compute_optimisation_weights(data, fitting):
        """R2star/gpuR2starMapping.m:507-529."""
        data = np.asarray(data)
        if fitting["isWeighted"]:
            if str(fitting["weightMethod"]).lower() == "norm":
                w = np.sqrt(np.abs(data))
            else:
                p = fitting["weightPower"]
                w = np.abs(data) ** p / np.abs(data[:, :, :, :1]) ** p
        else:
            return None
        w = np.where(w > 1, 1, w)
        w = np.where(w < 0, 0, w)
        return w.astype(np.float32)

    def fit(self, data, mask=None, fitting=None):
        """R2star/gpuR2starMapping.m:182-252."""
        data = np.asarray(data)
        dims = utils.dims3(data)
        if mask is None:
            mask = np.ones(dims, bool)
        fitting = self.check_set_default(fitting)
        self.updateProperty(fitting)
        fitting["modelParams"] = list(self.modelParams)
        if len(np.atleast_1d(fitting["ub"])) == 0:
            fitting["ub"] = list(self.ub)
        if len(np.atleast_1d(fitting["lb"])) == 0:
            fitting["lb"] = list(self.lb)
        pars0 = self.determine_x0(data, mask, fitting)
        w = self.compute_optimisation_weights(data, fitting)
        return self._solve(data, mask, w, pars0, fitting, fitting["solver"])

    def estimate(self, data, mask, fitting=None):
        """R2star/gpuR2starMapping.m:94-179 (without memory segmentation:
        a whole volume fits in one launch)."""
        fitting = self.check_set_default(fitting)
        data = np.asarray(data, dtype=np.float64)
        mask = np.asarray(mask) > 0
        scale = utils.prctile(np.abs(data[:, :, :, 0])[mask], 98)
        data = data / scale
        data, mask = utils.remove_img_naninf(data, mask)
        out = self.fit(data.astype(np.float32), mask, fitting)
        if str(fitting["solver"]).lower() == "mcmc":
            for k in fitting.get("metric", ["mean", "std"]):
                if k in out and "M0" in out[k]:
                    out[k]["M0"] = out[k]["M0"] * scale
            out["posterior"]["M0"] = out["posterior"]["M0"] * scale
        else:
            out["final"]["M0"] = out["final"]["M0"] * scale
            out["min"]["M0"] = out["min"]["M0"] * scale
        out["mask"] = mask
        return out


# ===========================================================================
class gpuJointR1R2starMapping(_ModelBase):
    """Joint VFA-SPGR x multi-echo R1-R2* mapping, R1R2s/gpuJointR1R2starMapping.m."""

    modelParams = ["M0", "R1", "R2star", "noise"]
    ub = [2, 10, 200, 0.1]
    lb = [0, 0.1, 0.1, 0.001]
    startPoint = [1, 1, 30, 0.05]
    step = [0.01, 0.01, 1, 0.005]
    model_id = _abi.GC_MODEL_JOINT_R1R2STAR
    seed = 48463

    def __init__(self, te, tr, fa):
        self.te = np.asarray(te, dtype=np.float32).ravel()
        self.tr = np.float32(np.asarray(tr).ravel()[0])
        self.fa = np.asarray(fa, dtype=np.float32).ravel()
        self._init_contract()
        self.replicate_fit_noise_quirk = True

    def gc_model(self, fitting, solver, nmeas):
        md = self._base_model(fitting)
        nte, nfa = self.te.size, self.fa.size
        md.i[0], md.i[1] = nte, nfa
        md.f[0] = float(self.tr)
        md.n_meas = nte * nfa
        off = _fill_table(md, self.te)
        _fill_table(md, np.deg2rad(self.fa.astype(np.float32)).astype(np.float32), off)
        md.n_extra = 1
        if nmeas is not None and nmeas != md.n_meas:
            raise GacelleError(_abi.GC_ERR_INPUT_MISMATCH, "The size of TE/FA does not match the image.")
        return md

    def extra_rows(self, extraData, mask, vox, nv):
        """extraData.b1 masked to G-D (:264); ones when absent (:466-468)."""
        if extraData is None or "b1" not in (extraData or {}):
            return np.ones((1, nv), np.float32)
        b1 = np.asarray(extraData["b1"], dtype=np.float32)
        if b1.size == nv:
            return b1.reshape(1, nv)
        return b1.ravel(order="F")[vox].reshape(1, nv)

    @staticmethod
    def check_set_default(fitting):
        return gpuR2starMapping.check_set_default(fitting)

    def estimate_prior(self, y, mask, fitting):
        return utils.initialise_x0(utils.dims3(mask), self.modelParams, self.startPoint)

    @staticmethod
    def compute_optimisation_weights(data, fitting):
        return gpuR2starMapping.compute_optimisation_weights(data, fitting)

    def fit(self, data, mask=None, fitting=None, extraData=None):
        """R1R2s/gpuJointR1R2starMapping.m:196-288.  The reference does not
        call updateProperty here (quirk Q2): with askadam, 'noise' stays a
        parameter slot that the model never reads.  Replicated when
        replicate_fit_noise_quirk is True."""
        data = np.asarray(data)
        dims = utils.dims3(data)
        if mask is None:
            mask = np.ones(dims, bool)
        fitting = self.check_set_default(fitting)
        if not self.replicate_fit_noise_quirk:
            self.updateProperty(fitting)
        fitting["modelParams"] = list(self.modelParams)
        if len(np.atleast_1d(fitting["ub"])) == 0:
            fitting["ub"] = list(self.ub)
        if len(np.atleast_1d(fitting["lb"])) == 0:
            fitting["lb"] = list(self.lb)
        pars0 = self.determine_x0(data, mask, fitting)
        w = self.compute_optimisation_weights(data, fitting)
        return self._solve(data, mask, w, pars0, fitting, extraData, fitting["solver"])


# ===========================================================================
def gauss_legendre_01(n: int):
    """n-point Gauss-Legendre nodes/weights on [0, 1]."""
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1) / 2, w / 2


def trapz_weights(x):
    """Weights of the composite trapezoid on nodes x (gacelle_trapz.m:217-226)."""
    x = np.asarray(x, dtype=np.float64)
    dx = np.diff(x)
    w = np.zeros_like(x)
    w[:-1] += dx / 2
    w[1:] += dx / 2
    return w


class gpuNEXI(_ModelBase):
    """NEXI Karger-exchange spherical-mean model, NEXI/gpuNEXI.m.

    quad='trapz14' reproduces the reference's 14-node trapezoid over
    x in [0, 1] (NEXI/gpuNEXI.m:950-963); quad='gl64' (or 'glN') uses
    N-point Gauss-Legendre.  rician=True applies the Rician first-moment
    correction of NEXI/gpuNEXIrice.m:652 with sigma = extraData['sigma'].
    """

    modelParams = ["fa", "Da", "De", "ra", "p2", "noise"]
    ub = [1, 3, 3, 1, 1, 0.1]
    lb = [float(np.finfo(np.float64).eps)] * 3 + [1 / 250, float(np.finfo(np.float64).eps), 0.01]
    startPoint = [0.4, 2, 1, 0.05, 0.2, 0.005]
    step = [0.05, 0.15, 0.15, 0.005, 0.05, 0.005]
    model_id = _abi.GC_MODEL_NEXI
    epsilon = 1e-8

    def __init__(self, b, Delta, Nav=None, quad="trapz14", rician=False, re_guard=None):
        b = np.asarray(b, dtype=np.float64).ravel()
        Delta = np.asarray(Delta, dtype=np.float64).ravel()
        # DWIutility.unique_shell_keepb0(..., discardB0=true): drop b=0, keep
        # one entry per (b, Delta) shell in sorted order
        keep = b > 0
        shells = sorted(set(zip(np.round(b[keep], 6), np.round(Delta[keep], 6))), key=lambda t: (t[1], t[0]))
        self.b = np.array([s[0] for s in shells], dtype=np.float32)
        self.Delta = np.array([s[1] for s in shells], dtype=np.float32)
        self.Nav = Nav
        self.quad = quad
        self.rician = bool(rician)
        self.re_guard = self.rician if re_guard is None else bool(re_guard)
        self._init_contract()

    def updateProperty(self, fitting):
        super().updateProperty(fitting)
        if int(fitting.get("lmax", 0)) == 0 and "p2" in self.modelParams:
            idx = self.modelParams.index("p2")
            for a in ("modelParams", "lb", "ub", "startPoint", "step"):
                getattr(self, a).pop(idx)
        return self

    def nodes(self):
        if self.quad.startswith("trapz"):
            n = int(self.quad[5:] or 14)
            x = np.linspace(0, 1, n).astype(np.float32)
            return x, trapz_weights(x.astype(np.float64)), _abi.GC_QUAD_TRAPZ
        n = int(self.quad[2:])
        x, w = gauss_legendre_01(n)
        return x, w, _abi.GC_QUAD_TABLE

    def gc_model(self, fitting, solver, nmeas):
        md = self._base_model(fitting)
        ns = self.b.size
        lmax = 2 if "p2" in fitting["modelParams"] else 0
        x, w, kind = self.nodes()
        md.i[0], md.i[1], md.i[2], md.i[3] = ns, lmax, x.size, kind
        md.f[0] = self.epsilon
        md.n_meas = ns * (2 if lmax == 2 else 1)
        off = _fill_table(md, self.b)
        off = _fill_table(md, self.Delta, off)
        off = _fill_table(md, x, off)
        off = _fill_table(md, w, off)
        _fill_table(md, np.cos(np.linspace(0, np.pi, 34)), off)
        flags = 0
        if self.rician:
            flags |= _abi.GC_FLAG_NEXI_RICIAN
            md.n_extra = 1
        if self.re_guard:
            flags |= _abi.GC_FLAG_NEXI_RE_GUARD
        md.flags = flags
        if nmeas is not None and nmeas != md.n_meas:
            raise GacelleError(_abi.GC_ERR_INPUT_MISMATCH, "The number of b-shells does not match the data.")
        return md

    def extra_rows(self, extraData, mask, vox, nv):
        if not self.rician:
            return None
        if extraData is None or "sigma" not in extraData:
            raise GacelleError(_abi.GC_ERR_ARG, "the Rician model needs extraData['sigma']")
        s = np.asarray(extraData["sigma"], dtype=np.float32)
        if s.size == 1:
            return np.full((1, nv), float(s), np.float32)
        if s.size == nv:
            return s.reshape(1, nv)
        return s.ravel(order="F")[vox].reshape(1, nv)

    @staticmethod
    def check_set_default(fitting):
        """NEXI/gpuNEXI.m:1101-1162 (volumetric, TV regularisation)."""
        from .askadam import askadam as _ask
        from .mcmc import mcmc as _mcmc

        fitting = dict(fitting or {})
        if "solver" not in fitting:
            fitting["solver"] = "askadam"
        if str(fitting["solver"]).lower() == "mcmc":
            f2 = _mcmc.check_set_default_basic(fitting)
            f2["lossFunction"] = "l2"
        else:
            f2 = _ask.check_set_default_basic(fitting)
            if "regmap" not in fitting:
                f2["regmap"] = ["fa"]
            if isinstance(f2["regmap"], str):
                f2["regmap"] = [f2["regmap"]]
            if "regularisationType" not in fitting:
                f2["regularisationType"] = "TV"
        if "lmax" not in fitting:
            f2["lmax"] = 0
        if "start" not in fitting:
            f2["start"] = "default"
        if "dataType" not in fitting:
            f2["dataType"] = "volumetric"
        return f2

    def fit(self, dwi, mask=None, fitting=None, extraData=None):
        """NEXI/gpuNEXI.m:290-450 (start 'default' or user values; the
        dictionary-likelihood start is not part of this path)."""
        dwi = np.asarray(dwi)
        dims = utils.dims3(dwi)
        if mask is None:
            mask = np.ones(dims, bool)
        fitting = self.check_set_default(fitting)
        self.updateProperty(fitting)
        fitting["modelParams"] = list(self.modelParams)
        if len(np.atleast_1d(fitting["ub"])) == 0:
            fitting["ub"] = list(self.ub)
        if len(np.atleast_1d(fitting["lb"])) == 0:
            fitting["lb"] = list(self.lb)
        pars0 = self.determine_x0(dwi, mask, fitting)
        return self._solve(dwi, mask, None, pars0, fitting, extraData)


gpuNEXIrice = gpuNEXI  # the Rician variant is gpuNEXI(..., rician=True)


# ===========================================================================
class gpuMCRMWI(_ModelBase):
    """Multi-compartment relaxometry myelin water imaging (3-pool complex
    GRE with VFA T1 weighting), MCRMWI/gpuMCRMWI.m.

    The device model covers the DIMWI-fitted path (isFitFreqMW,
    isFitFreqIW, isFitR2sEW and isFitIWF true) with isEPG = false: the
    T1 steady state is the Bloch-McConnell analytical solution when
    isFitExchange, else the non-exchanging Bloch model.  Complex data fit
    dfreqBKG (one map per flip angle) and dpini.  Four flip angles.
    """

    modelParams = ["S0", "MWF", "IWF", "R1IEW", "kIEWM", "R2sMW", "R2sIW", "R2sEW", "freqMW", "freqIW", "dfreqBKG",
                   "dpini", "noise"]
    ub = [2, 0.3, 1, 2, 10, 200, 50, 50, 0.25, 0.05, 0.4, np.pi / 2, 0.1]
    lb = [1e-8, 1e-8, 1e-8, 1 / 3, 1e-8, 50, 1 / 150e-3, 1 / 150e-3, -0.05, -0.1, -0.4, -np.pi / 2, 0.001]
    startPoint = [1, 0.1, 0.6, 1, 1, 100, 15, 21, 0.04, 1e-8, 1e-8, 1e-8, 0.005]
    step = [0.1, 0.015, 0.05, 0.1, 1, 10, 3.5, 3.5, 0.02, 0.01, 0.05, 0.075, 0.005]
    model_id = _abi.GC_MODEL_MCRMWI
    gyro = 42.57747892

    def __init__(self, te, tr, fa, fixed_params=None):
        self.te = np.asarray(te, dtype=np.float32).ravel()
        self.tr = float(np.float32(tr))
        self.fa = np.asarray(fa, dtype=np.float32).ravel()
        fp = fixed_params or {}
        self.B0 = float(fp.get("B0", 3.0))
        self.rho_mw = float(fp.get("rho_mw", 0.36 / 0.86))
        self.t1_mw = float(fp.get("t1_mw", 234e-3))
        self.freq_EW = float(fp.get("freq_EW", 0.0))
        self._init_contract()

    @staticmethod
    def check_set_default(fitting, data=None):
        """MCRMWI/gpuMCRMWI.m:1063-1107 (the ANN paths are not used)."""
        from .askadam import askadam as _ask
        from .mcmc import mcmc as _mcmc

        fitting = dict(fitting or {})
        fitting.setdefault("solver", "askadam")
        if str(fitting["solver"]).lower() == "mcmc":
            f2 = _mcmc.check_set_default_basic(fitting)
        else:
            f2 = _ask.check_set_default_basic(fitting)
            if "regmap" not in fitting:
                f2["regmap"] = "MWF"
        f2.setdefault("isWeighted", True)
        f2.setdefault("weightMethod", "1stecho")
        f2.setdefault("weightPower", 2)
        f2.setdefault("start", "prior")
        dimwi = dict(fitting.get("DIMWI", {}))
        for k in ("isFitFreqMW", "isFitFreqIW", "isFitR2sEW", "isFitIWF"):
            dimwi.setdefault(k, True)
        f2["DIMWI"] = dimwi
        f2.setdefault("isFitExchange", True)
        # the EPG-X MLP path (reference default) is not part of this library
        f2.setdefault("isEPG", False)
        f2.setdefault("isComplex", True)
        if data is not None and not np.iscomplexobj(data):
            f2["isComplex"] = False
        return f2

    def updateProperty(self, fitting):
        super().updateProperty(fitting)
        drop = []
        if not fitting.get("isComplex", True):
            drop += ["dfreqBKG", "dpini"]
        if not fitting.get("isFitExchange", True):
            drop += ["kIEWM"]
        for name in drop:
            if name in self.modelParams:
                idx = self.modelParams.index(name)
                for a in ("modelParams", "lb", "ub", "startPoint", "step"):
                    getattr(self, a).pop(idx)
        return self

    def _param_size(self, name) -> int:
        return self.fa.size if name == "dfreqBKG" else 1

    def _flags(self, fitting):
        f = 0
        if fitting.get("isComplex", True):
            f |= _abi.GC_FLAG_MCRMWI_COMPLEX
        if fitting.get("isFitExchange", True):
            f |= _abi.GC_FLAG_MCRMWI_EXCHANGE
        return f

    def gc_model(self, fitting, solver, nmeas):
        dimwi = fitting.get("DIMWI", {})
        if fitting.get("isEPG", False) or not all(dimwi.get(k, True) for k in
                                                  ("isFitFreqMW", "isFitFreqIW", "isFitR2sEW", "isFitIWF")):
            raise GacelleError(_abi.GC_ERR_UNSUPPORTED,
                               "gpuMCRMWI: the EPG-X and HCFM paths are not available (isEPG=false, DIMWI fitted)")
        md = self._base_model(fitting)
        md.flags = self._flags(fitting)
        nte, nfa = self.te.size, self.fa.size
        md.i[0], md.i[1] = nte, nfa
        md.n_meas = nte * nfa * (2 if md.flags & _abi.GC_FLAG_MCRMWI_COMPLEX else 1)
        md.n_extra = 3 + nfa
        md.f[0], md.f[1], md.f[2] = self.tr, self.B0, self.gyro
        md.f[3], md.f[4], md.f[5] = self.rho_mw, self.t1_mw, self.freq_EW
        off = _fill_table(md, self.te)
        _fill_table(md, np.deg2rad(self.fa.astype(np.float64)).astype(np.float32), off)
        dte = uniform_spacing(self.te)
        if dte is not None:
            md.flags |= _abi.GC_FLAG_UNIFORM_TE
            md.f[8] = dte
        if nmeas is not None and nmeas != md.n_meas:
            raise GacelleError(_abi.GC_ERR_INPUT_MISMATCH, "The size of TE/FA does not match the data.")
        return md

    def extra_rows(self, extraData, mask, vox, nv):
        """[B1, pini, ff, freqBKG x nfa] per voxel (extraData maps on the image
        grid, masked to G-D; defaults 1, 0, 1, 0)."""
        ex = extraData or {}
        nfa = self.fa.size
        rows = np.zeros((3 + nfa, nv), np.float32)

        def gd(key, default, k=None):
            if key not in ex or ex[key] is None:
                return np.full(nv, default, np.float32)
            a = np.asarray(ex[key], np.float32)
            if a.size == 1:
                return np.full(nv, float(a), np.float32)
            d = utils.dims3(mask)
            flat = a.reshape(d[0] * d[1] * d[2], -1, order="F")
            return flat[vox, 0 if k is None else k]

        rows[0] = gd("b1", 1.0)
        rows[1] = gd("pini", 0.0)
        rows[2] = gd("ff", 1.0)
        for k in range(nfa):
            rows[3 + k] = gd("freqBKG", 0.0, k)
        return rows
