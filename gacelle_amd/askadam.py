"""askadam: whole-volume gradient-descent estimation on the GPU.

Host-side mirror of utils/askadam.m.  ``askadam().optimisation`` keeps the
reference's signature and output structure; the whole iteration loop runs
in libgacelle_hip.so (gc_askadam_run): fused forward + hand-written adjoint
+ parameter transform + Adam/SGDM/RMSProp in one kernel per block of
iterations, with the stopping logic of the reference evaluated per
iteration.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _abi, runtime, utils
from ._abi import GacelleError


def _is_set(fitting: dict, key: str) -> bool:
    return key in fitting and fitting[key] is not None


class askadam:
    """Mirror of the askadam class (utils/askadam.m)."""

    epsilon = 1e-8

    # ------------------------------------------------------------------
    # defaults: utils/askadam.m:735-878
    @staticmethod
    def check_set_default_basic(fitting: dict | None) -> dict:
        fitting = dict(fitting or {})
        f2 = dict(fitting)
        f2["defaultRegularisation"] = True

        def d(key, val):
            if key not in fitting:
                f2[key] = val

        d("optimiser", "adam")
        d("initialLearnRate", 0.001)
        d("decayRate", 0)
        d("enableComplex", True)
        d("randomness", 0)
        d("parameterTransform", "sigmoid")
        opt = f2["optimiser"]
        if opt == "adam":
            d("adamupdateGradDecay", 0.9)
            d("adamupdateSqGradDecay", 0.999)
            d("adamupdateEpsilon", 1e-8)
        elif opt == "sgdm":
            d("sgdmupdateMomentum", 0.9)
        elif opt == "rmsprop":
            d("rmspropupdateSqGradDecay", 0.9)
            d("rmspropupdateEpsilon", 1e-8)
        d("lossFunction", "L1")
        d("tol", 1e-3)
        d("iteration", 1e4)
        d("convergenceValue", 1e-6)
        d("patience", 5)
        d("patienceConvergence", f2["patience"])
        if "Nepoch" in fitting:  # legacy field (:797)
            f2["iteration"] = fitting["Nepoch"]
            f2.pop("Nepoch", None)
        d("convergenceModel", "ema")
        d("convergenceWindow", 20)
        d("emaDecay", 0.95)
        d("robustConvergence", False)
        d("outlierThresholdMethod", "behaviour")
        d("outlierWeight", 0.1)
        d("weightUpdateInterval", 5)
        d("outlierCheckWindow", 5)
        d("outlierMinFlagDuration", 5)
        d("outlierVoxelThres", 0.01)
        d("outlierPopThres", 0.05)
        d("outlierInitThres", 0.05)
        d("outlierInitPopThres", 0.20)
        d("convergenceStepTol", 1e-6)
        d("convergenceGradTol", 1e-6)
        d("patienceStep", f2["patience"])
        d("patienceGrad", f2["patience"])
        d("lambda", [0])
        d("TVmode", "2D")
        d("regmap", [])
        d("voxelSize", [2, 2, 2])
        if not isinstance(f2["lambda"], (list, tuple)):
            f2["lambda"] = list(np.atleast_1d(f2["lambda"]).tolist())
        d("isOptimiseMemory", True)
        d("autoMemManage", True)
        d("segmentOverlap", 0)
        d("NSegmentUser", [])
        d("outputFilename", [])
        d("ub", [])
        d("lb", [])
        d("debug", False)
        d("isDisplay", 0)
        d("isMaskedOut", True)
        d("isSampleConsistency", False)
        d("isClipGradient", 0)
        d("maxGradientThres", 1)
        # options of this implementation (no reference counterpart)
        d("adamForm", "effective")        # 'effective' | 'kingma'
        d("iterationsPerLaunch", 0)       # 0 = automatic
        d("lossTrace", False)
        d("verbose", True)
        return f2

    # ------------------------------------------------------------------
    # transforms: utils/askadam.m:1063-1117 (host versions, numpy)
    @staticmethod
    def rescale01(img, lb, ub):
        return (img - lb) / (ub - lb)

    @staticmethod
    def unscale01(img_norm, lb, ub):
        return img_norm * (ub - lb) + lb

    @staticmethod
    def set_boundary(parameters: dict, ub, lb) -> dict:
        out = {}
        for k, name in enumerate(parameters):
            out[name] = np.minimum(np.maximum(parameters[name], lb[k]), ub[k])
        return out

    # ------------------------------------------------------------------
    def optimisation(self, data, mask, weights, parameters, fitting, userfuncCell, *varargin):
        """askadam().optimisation(data, mask, weights, parameters, fitting,
        userfuncCell, varargin) -- utils/askadam.m:121-365.

        userfuncCell is a model object of this package (or its bound FWD
        method, or a one-element list holding either); varargin are the
        FWD extra arguments of the reference call (e.g. extraData).
        """
        res = self.optimisation_multi([(data, mask, weights, parameters)], fitting, userfuncCell, *varargin)
        return res[0]

    def optimisation_multi(self, problems, fitting, userfuncCell, *varargin):
        """Batched askadam over independent volumes of uneven size: one
        fused launch per block of iterations, each volume with its own N,
        stopping state and outputs.  problems = [(data, mask, weights,
        parameters[, extraData]), ...]."""
        model = _resolve_model(userfuncCell)
        fitting = self.check_set_default_basic(fitting)
        lam = [float(x) for x in np.atleast_1d(fitting["lambda"])]
        if not fitting["isOptimiseMemory"] and lam and lam[0] > 0:
            # With isOptimiseMemory false the reference optimises whole N-D
            # maps: voxels outside the mask next to masked ones enter the TV
            # term and are learned (utils/spatial_total_variation.m:32-48).
            # Only the masked (G-D) form is implemented.
            raise GacelleError(_abi.GC_ERR_UNSUPPORTED, "isOptimiseMemory = false with a regulariser (whole-grid "
                                                        "parameter maps) is not supported; use the default true")
        ctx = runtime.context()
        lib = ctx.lib
        start = time.time()
        vols = (_abi.gc_volume * len(problems))()
        keep = []  # tensors that must outlive the call
        metas = []
        md = None
        for i, prob in enumerate(problems):
            data, mask, weights, parameters = prob[:4]
            extra_data = prob[4] if len(prob) > 4 else (varargin[0] if varargin else None)
            mask = np.asarray(mask) > 0
            vox = utils.find_mask(mask)
            y = _to_gd(data, mask, vox)
            nmeas, nv = y.shape
            w = None
            if weights is not None and np.size(weights) > 0:
                w = _to_gd(weights, mask, vox)
                if np.all(w == 1):
                    w = None
            mdi = model.gc_model(fitting, "askadam", nmeas)
            if hasattr(model, "model_for_extra"):
                mdi = model.model_for_extra(mdi, extra_data)
            # the acquisition part (network parameter pointers aside) is shared
            acq = bytes(mdi)[: _abi.gc_model.mlp.offset]
            if md is None:
                md_acq = acq
                md = model.bind_device(mdi, ctx, keep) if hasattr(model, "bind_device") else mdi
            elif md_acq != acq:
                raise GacelleError(_abi.GC_ERR_INPUT_MISMATCH, "all volumes of a batch must share the acquisition")
            extra = model.extra_rows(extra_data, mask, vox, nv)
            P = md.n_par
            theta0 = _initial_theta(parameters, fitting, mask, vox, nv)
            y_d = ctx.to_device(y)
            w_d = ctx.to_device(w) if w is not None else None
            ex_d = ctx.to_device(extra) if extra is not None else None
            z_d = ctx.empty((P, nv))
            if theta0 is None:
                # no starting point: z = randomness*randn (sigmoid) or U(0,1) (linear) (:394-401)
                if fitting["parameterTransform"] == "sigmoid":
                    z0 = fitting["randomness"] * np.random.randn(P, nv)
                else:
                    z0 = np.random.rand(P, nv)
                z_d.copy_(ctx.to_device(z0))
            else:
                th_d = ctx.to_device(theta0)
                tr = _transform_id(fitting)
                ctx.check(lib.gc_rescale_parameters(ctx.handle, ctypes.byref(md), tr, runtime.ptr(th_d),
                                                    runtime.ptr(z_d), nv))
                r = float(fitting["randomness"])
                if r != 0:
                    if tr == _abi.GC_TRANSFORM_SIGMOID:
                        noise = ctx.to_device(np.random.randn(P, nv))
                    else:
                        noise = ctx.to_device(np.random.rand(P, nv))
                    z_d.mul_(1 - r).add_(noise * r)
                keep.append(th_d)
            zmin_d = ctx.empty((P, nv))
            res_d = ctx.empty((nmeas, nv))
            resmin_d = ctx.empty((nmeas, nv))
            rl_d = ctx.empty((nv,))
            rlmin_d = ctx.empty((nv,))
            trace = None
            if fitting["lossTrace"]:
                trace = (ctypes.c_double * (int(fitting["iteration"]) + 1))()
            v = vols[i]
            v.nv = nv
            v.y_dev = runtime.ptr(y_d)
            v.w_dev = runtime.ptr(w_d)
            v.extra_dev = runtime.ptr(ex_d)
            vox_d = ctx.to_device(vox, np.int64)
            v.vox_dev = runtime.ptr(vox_d)
            d3 = utils.dims3(mask)
            for k in range(3):
                v.dims[k] = d3[k]
            v.z_dev = runtime.ptr(z_d)
            v.zmin_dev = runtime.ptr(zmin_d)
            v.resloss_dev = runtime.ptr(rl_d)
            v.resloss_min_dev = runtime.ptr(rlmin_d)
            v.residual_dev = runtime.ptr(res_d)
            v.residual_min_dev = runtime.ptr(resmin_d)
            v.loss_trace_host = ctypes.cast(trace, ctypes.POINTER(ctypes.c_double)) if trace is not None else None
            prior = _prior_rows(fitting, extra_data, mask, vox)
            if prior is not None:
                rows, v.reg_offset = prior
                pr_d = ctx.to_device(rows)
                v.prior_dev = runtime.ptr(pr_d)
                keep.append(pr_d)
            stv = _surface_tables(fitting, extra_data, mask)
            if stv is not None:
                sn_d, sl_d = ctx.to_device(stv[0], np.int32), ctx.to_device(stv[1])
                v.surf_nbr_dev, v.surf_len_dev = runtime.ptr(sn_d), runtime.ptr(sl_d)
                v.surf_max_nb = stv[0].shape[0]
                keep += [sn_d, sl_d]
            keep += [y_d, w_d, ex_d, vox_d, z_d, zmin_d, res_d, resmin_d, rl_d, rlmin_d]
            metas.append(dict(mask=mask, vox=vox, nv=nv, nmeas=nmeas, z=z_d, zmin=zmin_d, res=res_d, resmin=resmin_d,
                              rl=rl_d, rlmin=rlmin_d, trace=trace, theta0=theta0))
        cfg = self.make_cfg(fitting, model)
        if fitting.get("verbose", True):
            self.display_basic_fitting_parameters(fitting)
        ctx.check(lib.gc_askadam_run(ctx.handle, ctypes.byref(md), ctypes.byref(cfg), vols, len(problems)))
        outs = []
        for i, m in enumerate(metas):
            outs.append(self._assemble(ctx, md, fitting, model, vols[i], m))
        if fitting.get("verbose", True):
            v = vols[0]
            print(f"Final loss         =  {v.loss:e}")
            print(f"Final convergence  =  {v.convergence:e}")
            print(f"Final #iterations  =  {v.n_iteration}")
            print(f"Total Elapsed time =  {time.time() - start:.3f} s")
            if v.stop_reason:
                print(_abi.STOP_MESSAGES[v.stop_reason])
        return outs

    # ------------------------------------------------------------------
    def make_cfg(self, fitting: dict, model) -> _abi.gc_askadam_cfg:
        c = _abi.gc_askadam_cfg()
        c.iteration = int(fitting["iteration"])
        c.optimiser = {"adam": _abi.GC_OPT_ADAM, "sgdm": _abi.GC_OPT_SGDM, "rmsprop": _abi.GC_OPT_RMSPROP}[
            str(fitting["optimiser"]).lower()]
        c.loss = {"l1": _abi.GC_LOSS_L1, "l2": _abi.GC_LOSS_L2, "huber": _abi.GC_LOSS_HUBER, "mse": _abi.GC_LOSS_MSE}[
            str(fitting["lossFunction"]).lower()]
        c.transform = _transform_id(fitting)
        c.adam_form = _abi.GC_ADAM_KINGMA_BA if str(fitting["adamForm"]).lower().startswith("k") else \
            _abi.GC_ADAM_EFFECTIVE_LR
        c.convergence_model = _abi.GC_CONV_EMA if fitting["convergenceModel"] == "ema" else _abi.GC_CONV_LINEAR
        c.convergence_window = int(fitting["convergenceWindow"])
        c.patience_conv = int(fitting["patienceConvergence"])
        c.patience_step = int(fitting["patienceStep"])
        c.patience_grad = int(fitting["patienceGrad"])
        c.block_iters = int(fitting["iterationsPerLaunch"])
        c.tv_mode = _abi.GC_TV_3D if str(fitting["TVmode"]).upper() == "3D" else _abi.GC_TV_2D
        lam = [float(x) for x in fitting["lambda"]]
        regmap = fitting["regmap"]
        if isinstance(regmap, str):
            regmap = [regmap]
        if lam and lam[0] > 0:
            if len(regmap) < len(lam):
                raise GacelleError(_abi.GC_ERR_ARG, "fitting.regmap must name one map per lambda")
            c.n_reg = len(lam)
            names = list(fitting["modelParams"])
            slots = dict(model.param_slots(fitting)) if model is not None else None
            for k, name in enumerate(regmap[: len(lam)]):
                # the parameter slot of the map (multi-slot maps precede it)
                c.regmap[k] = slots[name].start if slots is not None else names.index(name)
                c.lambda_[k] = lam[k]
        c.reg_type = _abi.GC_REG_PRIOR if str(fitting.get("regularisationType", "TV")).lower() == "prior" \
            else (_abi.GC_REG_SURFACE_TV if _is_surface_tv(fitting) else _abi.GC_REG_TV)
        c.robust = int(bool(fitting["robustConvergence"]))
        c.lr0 = float(fitting["initialLearnRate"])
        c.decay = float(fitting["decayRate"])
        c.beta1 = float(fitting.get("adamupdateGradDecay", 0.9))
        c.beta2 = float(fitting.get("adamupdateSqGradDecay", 0.999))
        c.eps = float(fitting.get("adamupdateEpsilon", 1e-8))
        c.momentum = float(fitting.get("sgdmupdateMomentum", 0.9))
        c.rms_decay = float(fitting.get("rmspropupdateSqGradDecay", 0.9))
        c.rms_eps = float(fitting.get("rmspropupdateEpsilon", 1e-8))
        c.tol = float(fitting["tol"])
        c.convergence_value = float(fitting["convergenceValue"])
        c.ema_decay = float(fitting["emaDecay"])
        c.step_tol = float(fitting["convergenceStepTol"])
        c.grad_tol = float(fitting["convergenceGradTol"])
        vs = list(np.atleast_1d(fitting["voxelSize"]).astype(float)) + [1.0, 1.0, 1.0]
        for k in range(3):
            c.voxel_size[k] = vs[k]
        c.weight_update_interval = int(fitting["weightUpdateInterval"])
        c.outlier_check_window = int(fitting["outlierCheckWindow"])
        c.outlier_weight = float(fitting["outlierWeight"])
        c.outlier_voxel_thres = float(fitting["outlierVoxelThres"])
        c.outlier_pop_thres = float(fitting["outlierPopThres"])
        c.outlier_init_thres = float(fitting["outlierInitThres"])
        c.outlier_init_pop_thres = float(fitting["outlierInitPopThres"])
        return c

    def _assemble(self, ctx, md, fitting, model, v, m) -> dict:
        """Finalisation of utils/askadam.m:312-362."""
        lib = ctx.lib
        nv, P = m["nv"], md.n_par
        tr = _transform_id(fitting)
        th_f = ctx.empty((P, nv))
        th_m = ctx.empty((P, nv))
        ctx.check(lib.gc_unscale_parameters(ctx.handle, ctypes.byref(md), tr, runtime.ptr(m["z"]), runtime.ptr(th_f), nv))
        ctx.check(lib.gc_unscale_parameters(ctx.handle, ctypes.byref(md), tr, runtime.ptr(m["zmin"]), runtime.ptr(th_m),
                                            nv))
        th_f = runtime.to_host(th_f)
        th_m = runtime.to_host(th_m)
        mask = m["mask"]
        final, mn = {}, {}
        for name, sl in model.param_slots(fitting):
            a = th_f[sl]
            b = th_m[sl]
            final[name] = _to_image(a, mask, fitting, m["theta0"], sl)
            mn[name] = _to_image(b, mask, fitting, m["theta0"], sl)
        final["loss"] = v.loss
        final["loss_fidelity"] = np.float32(v.loss_fidelity)
        final["loss_reg"] = np.float32(v.loss_reg)
        final["resloss"] = utils.reshape_ND2image(runtime.to_host(m["rl"]), mask)
        final["residual"] = runtime.to_host(m["res"])
        final["Niteration"] = int(v.n_iteration)
        final["memoryUsage"] = 0.0
        final["stopReason"] = int(v.stop_reason)
        mn["loss"] = v.min_loss
        mn["loss_fidelity"] = np.float32(v.min_loss_fidelity)
        mn["loss_reg"] = np.float32(v.min_loss_reg)
        mn["resloss"] = utils.reshape_ND2image(runtime.to_host(m["rlmin"]), mask)
        mn["residual"] = runtime.to_host(m["resmin"])
        mn["Niteration"] = int(v.min_iteration)
        if m["trace"] is not None:
            final["lossTrace"] = np.array(m["trace"][: int(v.n_iteration) + 1])
        return {"final": final, "min": mn, "nLaunch": int(v.n_launch)}

    @staticmethod
    def display_basic_fitting_parameters(fitting: dict):
        print("============================")
        print("AskAdam algorithm parameters")
        print("============================")
        print(f"Optimiser                    = {fitting['optimiser']}")
        print(f"Initial learning rate        = {fitting['initialLearnRate']}")
        print(f"Learning rate decay rate     = {fitting['decayRate']}")
        print(f"Parameter transform          = {fitting['parameterTransform']}")
        print(f"Loss function                = {fitting['lossFunction']}")
        print(f"Max. iterations              = {int(fitting['iteration'])}")
        print(f"Convergence model            = {fitting['convergenceModel']}")


def _transform_id(fitting) -> int:
    return _abi.GC_TRANSFORM_SIGMOID if fitting["parameterTransform"] == "sigmoid" else _abi.GC_TRANSFORM_LINEAR


def _resolve_model(userfuncCell):
    f = userfuncCell
    if isinstance(f, (list, tuple)):
        if len(f) > 1:
            raise GacelleError(_abi.GC_ERR_UNSUPPORTED,
                               "user-supplied regularisers are not supported; use fitting.lambda/regmap (built-in TV)")
        f = f[0]
    if hasattr(f, "gc_model"):
        return f
    if hasattr(f, "__self__") and hasattr(f.__self__, "gc_model"):
        return f.__self__
    raise GacelleError(_abi.GC_ERR_UNSUPPORTED,
                       "the fused GPU path needs one of this package's model classes as the forward model")


def _is_surface_tv(fitting) -> bool:
    lam = [float(x) for x in np.atleast_1d(fitting.get("lambda", [0]))]
    return (str(fitting.get("dataType", "volumetric")).lower() == "surface"
            and str(fitting.get("regularisationType", "TV")).lower() == "tv" and bool(lam) and lam[0] > 0)


def _surface_tables(fitting, extra_data, mask):
    """The mesh TV regulariser of surface data (gpuNEXI.surface_total_variation,
    NEXI/gpuNEXI.m:1165-1254; checks of :416-429): extradata['neighbours']
    (1-based per-hemisphere indices, [1, Nvertex, Nhemi, maxNb]) and the edge
    lengths, extradata['dr'] as gpuNEXI.fit passes them (or
    extradata['edgeLength']), converted to the C ABI's grid-form tables; or
    None when the fit has no surface TV."""
    if not _is_surface_tv(fitting):
        return None
    ex = extra_data if isinstance(extra_data, dict) else {}
    el = ex.get("edgeLength", ex.get("dr"))
    if ex.get("neighbours") is None or el is None:
        raise GacelleError(_abi.GC_ERR_ARG, "GACELLE:missingSurfaceGeometry: fitting.dataType is 'surface' and "
                           "fitting.lambda > 0, but extradata.neighbours/extradata.dr were not provided to fit().")
    from .surface import surface_tv_tables

    return surface_tv_tables(mask, ex["neighbours"], el)


def _prior_rows(fitting, extra_data, mask, vox):
    """The normal prior of gpuNEXI (NEXI/gpuNEXI.m:1256-1303, checks of
    :333-349): mu / sigma maps of every regmap name from
    extradata['mu'][name] / extradata['sigma'][name] on the image grid.
    Returns ([mu rows; sigma rows] at the masked voxels, the constant the
    grid voxels outside the mask add to loss_reg), or None without a prior."""
    if str(fitting.get("regularisationType", "TV")).lower() != "prior":
        return None
    lam = [float(x) for x in fitting["lambda"]]
    if not lam or lam[0] <= 0:
        return None
    regmap = fitting["regmap"]
    regmap = [regmap] if isinstance(regmap, str) else list(regmap)
    ex = extra_data if isinstance(extra_data, dict) else {}
    mu, sd = ex.get("mu"), ex.get("sigma")
    if not isinstance(mu, dict) or not isinstance(sd, dict):
        raise GacelleError(_abi.GC_ERR_ARG, "GACELLE:missingPriorDistribution: fitting.regularisationType is 'prior' "
                           "and fitting.lambda > 0, but extradata.mu/extradata.sigma were not provided to fit().")
    miss_mu = [n for n in regmap if n not in mu]
    miss_sd = [n for n in regmap if n not in sd]
    if miss_mu or miss_sd:
        raise GacelleError(_abi.GC_ERR_ARG, "GACELLE:incompletePriorDistribution: extradata.mu/extradata.sigma must "
                           "have a field for every name in fitting.regmap. Missing from mu: {%s}. Missing from sigma: "
                           "{%s}." % (", ".join(miss_mu), ", ".join(miss_sd)))
    names = regmap[: len(lam)]
    ngrid = int(np.prod(utils.dims3(mask)))
    inside = np.zeros(ngrid, bool)
    inside[vox] = True
    rows = np.zeros((2 * len(names), vox.size), np.float32)
    offset = 0.0
    for k, n in enumerate(names):
        m = np.asarray(mu[n], np.float32).reshape(-1, order="F")
        s = np.asarray(sd[n], np.float32).reshape(-1, order="F")
        if m.size != ngrid or s.size != ngrid:
            raise GacelleError(_abi.GC_ERR_INPUT_MISMATCH, f"prior maps of '{n}' must match the mask grid")
        rows[k], rows[len(names) + k] = m[vox], s[vox]
        with np.errstate(divide="ignore", invalid="ignore"):
            zo = np.abs((np.float32(0) - m[~inside]) / s[~inside])
        offset += lam[k] / vox.size * float(zo[np.isfinite(zo)].astype(np.float64).sum())
    return rows, offset


def _to_gd(data, mask, vox) -> np.ndarray:
    """Data (G-D [M, Nvox] or N-D image) -> masked G-D [M, Nv] float32
    (utils/askadam.m:157-158)."""
    data = np.asarray(data)
    if data.ndim == 2 and (mask.ndim < 2 or mask.shape[0] == 1 or data.shape[1] == mask.size) and \
            data.shape[1] == np.asarray(mask).size:
        out = data[:, vox]
        if np.iscomplexobj(out):
            out = np.concatenate([out.real, out.imag], axis=0)
        return np.ascontiguousarray(out, dtype=np.float32)
    return np.ascontiguousarray(utils.reshape_ND2GD(data, mask), dtype=np.float32)


def _initial_theta(parameters, fitting, mask, vox, nv):
    """Starting maps -> [P, Nv] physical values, clamped as
    initialise_parameter does before the logit (:380-388)."""
    if parameters is None or (isinstance(parameters, dict) and len(parameters) == 0):
        return None
    names = list(fitting["modelParams"])
    rows = []
    for name in names:
        a = np.asarray(parameters[name], dtype=np.float64)
        if a.ndim >= 3 or (a.size == np.asarray(mask).size and a.size != nv):
            if a.ndim <= 3:
                rows.append(a.ravel(order="F")[vox].reshape(1, nv))
            else:
                d = utils.dims3(a)
                flat = a.reshape((d[0] * d[1] * d[2], -1), order="F")
                rows.append(flat[vox, :].T)
        elif a.size == 1:
            rows.append(np.full((1, nv), float(a)))
        else:
            a2 = a.reshape(-1, order="F")
            if a2.size == np.asarray(mask).size:
                a2 = a2[vox]
            rows.append(a2.reshape(1, nv))
    return np.ascontiguousarray(np.concatenate(rows, axis=0), dtype=np.float32)


def _to_image(a, mask, fitting, theta0, sl):
    """[k, Nv] -> image; masked-out voxels are zero (isMaskedOut) or keep the
    starting value."""
    a = np.asarray(a)
    if a.shape[0] == 1:
        img = utils.reshape_ND2image(a[0], mask)
    else:
        img = utils.reshape_ND2image(a.T, mask)
    return img.astype(np.float32)
