"""mcmc: Metropolis-Hastings and Goodman-Weare samplers on the GPU.

Host-side mirror of utils/mcmc.m.  Chains (repetitions) and walkers of all
voxels advance in one kernel launch (gc_mcmc_run); posterior mean and
standard deviation are reduced on the device.  Draws come from a native
counter-based stream (Philox4x32-10) or, in replay mode, from caller-
supplied arrays in the reference's draw order.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi, runtime, utils
from ._abi import GacelleError
from .askadam import _resolve_model, _to_gd


class mcmc:
    """Mirror of the mcmc class (utils/mcmc.m)."""

    # defaults: utils/mcmc.m:633-679
    @staticmethod
    def check_set_default_basic(fitting: dict | None) -> dict:
        fitting = dict(fitting or {})
        f2 = dict(fitting)

        def d(key, val):
            if key not in fitting:
                f2[key] = val

        d("iteration", 2e5)
        d("thinning", 20)
        d("metric", ["mean", "std"])
        d("burnin", 0.1)
        d("repetition", 1)
        d("outputFilename", [])
        d("algorithm", "MH")
        d("StepSize", 2)
        d("Nwalker", 50)
        d("ub", [])
        d("lb", [])
        d("startRange", 0.001)
        d("Ensembleupdate", "simultaneous")
        d("globalWarmup", 0)
        if isinstance(f2["metric"], str):
            f2["metric"] = [f2["metric"]]
        if any(m == "mode" for m in f2["metric"]):
            d("Nbin", 1001)
        d("autoMemManage", True)
        if str(f2["algorithm"]).lower() == "gw":
            f2["algorithm"] = "ensemble"
        # options of this implementation
        d("rngMode", "native")          # 'native' | 'replay'
        d("seed", 48463)
        d("storePosterior", None)      # None = store when it fits in 1 GiB
        d("verbose", True)
        return f2

    @staticmethod
    def get_number_burnin(fitting) -> int:
        """utils/mcmc.m:788-794."""
        if fitting["burnin"] < 1:
            return int(np.floor(fitting["iteration"] * fitting["burnin"]))
        return int(fitting["burnin"])

    # ------------------------------------------------------------------
    def optimisation(self, data, mask, weights, pars0, fitting, FWDfunc, *varargin):
        """mcmc().optimisation, utils/mcmc.m:18-80."""
        fitting = self.check_set_default_basic(fitting)
        mask = np.asarray(mask) > 0
        vox = utils.find_mask(mask)
        y = _to_gd(data, mask, vox)
        w = None
        if weights is not None and np.size(weights) > 0:
            w = _to_gd(weights, mask, vox)
        x0 = {}
        for name in fitting["modelParams"]:
            a = np.asarray(pars0[name], dtype=np.float32)
            if a.size == 1:
                raise GacelleError(_abi.GC_ERR_UNSUPPORTED,
                                   "global (shared) parameters (goodman_weare_wglobal_constant) are not supported")
            x0[name] = a.ravel(order="F")[vox] if a.size == mask.size else a.ravel()
        extra = varargin[0] if varargin and isinstance(varargin[0], dict) else None
        if str(fitting["algorithm"]).lower() == "mh":
            post = self.metropolis_hastings(y, x0, w, fitting, FWDfunc, extra, _mask=mask, _vox=vox)
        else:
            post = self.goodman_weare(y, x0, w, fitting, FWDfunc, extra, _mask=mask, _vox=vox)
        return self.res2out(post, fitting, mask)

    def metropolis_hastings(self, y, x0, weights, fitting, FWDfunc, *varargin, _mask=None, _vox=None):
        """utils/mcmc.m:82-190; y is G-D [Nmeas, Nv]."""
        return self._run(_abi.GC_MCMC_MH, y, x0, weights, fitting, FWDfunc, varargin, _mask, _vox)

    def goodman_weare(self, y, x0, weights, fitting, FWDfunc, *varargin, _mask=None, _vox=None):
        """utils/mcmc.m:192-383 (simultaneous or red-black update)."""
        return self._run(_abi.GC_MCMC_GW, y, x0, weights, fitting, FWDfunc, varargin, _mask, _vox)

    def _run(self, algo, y, x0, weights, fitting, FWDfunc, varargin, mask, vox):
        fitting = self.check_set_default_basic(fitting)
        model = _resolve_model(FWDfunc)
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float32))
        nmeas, nv = y.shape
        if mask is None:
            mask = np.ones((1, nv), bool)
            vox = np.arange(nv, dtype=np.int64)
        names = list(fitting["modelParams"])
        Q = len(names)
        if "noise" not in names or names.index("noise") != Q - 1:
            raise GacelleError(_abi.GC_ERR_UNSUPPORTED, "MCMC expects 'noise' as the last model parameter")
        md = model.gc_model(fitting, "mcmc", nmeas)
        extra_data = varargin[0] if varargin and isinstance(varargin[0], dict) else None
        extra = model.extra_rows(extra_data, mask, vox, nv)
        cfg = _abi.gc_mcmc_cfg()
        cfg.algorithm = algo
        cfg.iteration = int(fitting["iteration"])
        cfg.thinning = int(fitting["thinning"])
        cfg.n_burnin = self.get_number_burnin(fitting)
        cfg.repetition = int(fitting["repetition"])
        W = int(fitting["Nwalker"]) if algo == _abi.GC_MCMC_GW else 1
        cfg.n_walker = W
        cfg.ensemble_update = _abi.GC_ENSEMBLE_REDBLACK if str(fitting["Ensembleupdate"]).lower() == "redblack" \
            else _abi.GC_ENSEMBLE_SIMULTANEOUS
        replay = str(fitting["rngMode"]).lower() == "replay"
        cfg.rng_mode = _abi.GC_RNG_REPLAY if replay else _abi.GC_RNG_NATIVE
        cfg.noise_slot = Q - 1
        cfg.n_var = Q
        cfg.seed = int(fitting["seed"]) & 0xFFFFFFFFFFFFFFFF
        cfg.step_size = float(fitting["StepSize"])
        cfg.start_range = float(fitting["startRange"])
        steps = list(np.atleast_1d(fitting.get("xStepSize", [0] * Q)).astype(float))
        for q in range(Q):
            cfg.x_step[q] = steps[q] if q < len(steps) else 0.0
        R = cfg.repetition
        ns = 0 if cfg.iteration < cfg.n_burnin + 1 else (cfg.iteration - cfg.n_burnin - 1) // cfg.thinning + 1
        store = fitting["storePosterior"]
        post_elems = Q * R * ns * W * nv
        if store is None:
            store = post_elems * 4 <= (1 << 30)
        ctx = runtime.context()
        xs = np.stack([np.asarray(x0[n], dtype=np.float32).reshape(nv) for n in names])
        y_d = ctx.to_device(y)
        w_d = ctx.to_device(weights) if weights is not None else None
        ex_d = ctx.to_device(extra) if extra is not None else None
        x0_d = ctx.to_device(xs)
        io = _abi.gc_mcmc_io()
        io.nv = nv
        io.y_dev = runtime.ptr(y_d)
        io.w_dev = runtime.ptr(w_d)
        io.extra_dev = runtime.ptr(ex_d)
        io.x0_dev = runtime.ptr(x0_d)
        keep = []
        if replay:
            draws = fitting["replayDraws"]
            nd = ctx.to_device(np.asarray(draws["normals"], np.float32))
            ud = ctx.to_device(np.asarray(draws["uniforms"], np.float32))
            keep += [nd, ud]
            io.normals_dev = runtime.ptr(nd)
            io.uniforms_dev = runtime.ptr(ud)
            if algo == _abi.GC_MCMC_GW:
                pd = ctx.to_device(np.asarray(draws["partners"], np.int32), np.int32)
                keep.append(pd)
                io.partners_dev = runtime.ptr(pd)
        post_d = ctx.empty((Q, R, ns, W, nv)) if store and ns > 0 else None
        mean_d = ctx.empty((Q, nv))
        std_d = ctx.empty((Q, nv))
        acc_d = ctx.empty((nv,))
        xf_d = ctx.empty((R, W, Q, nv))
        trace_d = ctx.zeros((R, cfg.iteration, W, nv), np.uint8) if fitting.get("traceAccept") else None
        io.posterior_dev = runtime.ptr(post_d)
        io.mean_dev = runtime.ptr(mean_d)
        io.std_dev = runtime.ptr(std_d)
        io.accept_dev = runtime.ptr(acc_d)
        io.xfinal_dev = runtime.ptr(xf_d)
        io.trace_dev = runtime.ptr(trace_d)
        ctx.check(ctx.lib.gc_mcmc_run(ctx.handle, ctypes.byref(md), ctypes.byref(cfg), ctypes.byref(io)))
        res = {
            "names": names,
            "mean": runtime.to_host(mean_d),
            "std": runtime.to_host(std_d),
            "accept": runtime.to_host(acc_d),
            "xfinal": runtime.to_host(xf_d),
            "nSample": int(io.n_sample),
            "nStored": int(io.n_stored),
            "algorithm": algo,
        }
        if trace_d is not None:
            res["trace"] = runtime.to_host(trace_d)
        wanted = [m for m in ("median", "iqr", "mode") if m in [x.lower() for x in fitting["metric"]]]
        if wanted:
            if post_d is None:
                raise GacelleError(_abi.GC_ERR_ARG, f"metric {wanted} needs fitting.storePosterior = True")
            # res2out median / iqr / mode on the device (utils/mcmc.m:823-857)
            outs = {m: ctx.empty((Q, nv)) for m in wanted}
            lb = (ctypes.c_float * Q)(*[float(x) for x in np.atleast_1d(fitting["lb"])[:Q]])
            ub = (ctypes.c_float * Q)(*[float(x) for x in np.atleast_1d(fitting["ub"])[:Q]])
            ctx.check(ctx.lib.gc_posterior_stats(ctx.handle, runtime.ptr(post_d), Q, R * ns * W, nv, lb, ub,
                                                 int(fitting.get("Nbin", 1001)), runtime.ptr(outs.get("median")),
                                                 runtime.ptr(outs.get("iqr")), runtime.ptr(outs.get("mode"))))
            for m, t in outs.items():
                res[m] = runtime.to_host(t)
        posterior = {}
        if post_d is not None:
            p = runtime.to_host(post_d)  # [Q, R, Ns, W, nv]
            for q, n in enumerate(names):
                a = np.transpose(p[q], (3, 2, 1, 0))  # [nv, W, Ns, R]
                posterior[n] = a[:, 0, :, :] if algo == _abi.GC_MCMC_MH else a
        res["posterior"] = posterior
        return res

    # ------------------------------------------------------------------
    def res2out(self, res, fitting, mask) -> dict:
        """utils/mcmc.m:806-862.  Every metric comes from the device: mean /
        std from the sampler's reduction, median / iqr / mode from
        gc_posterior_stats over the stored posterior."""
        out = {"posterior": res["posterior"]}
        names = res["names"]
        for metric in fitting["metric"]:
            m = metric.lower()
            if m not in ("mean", "std", "median", "iqr", "mode"):
                raise GacelleError(_abi.GC_ERR_ARG, f"unknown metric {metric}")
            out[m] = {}
            for q, n in enumerate(names):
                out[m][n] = utils.reshape_ND2image(np.asarray(res[m][q], np.float32), mask)
        out["acceptance"] = utils.reshape_ND2image(res["accept"], mask)
        return out
