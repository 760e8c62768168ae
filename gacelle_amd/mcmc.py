"""mcmc: Metropolis-Hastings and Goodman-Weare samplers on the GPU.

Host-side mirror of utils/mcmc.m.  Chains (repetitions) and walkers of all
voxels advance in one kernel launch (gc_mcmc_run); posterior mean and
standard deviation are reduced on the device.  With global (shared)
parameters the ensemble advances one iteration per launch sequence, since
every global decision sums over all voxels.  Draws come from a native
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
            # scalar fields stay scalars (reshape_ND2GD_struct, utils/utils.m:166-178)
            x0[name] = a.ravel() if a.size == 1 else (a.ravel(order="F")[vox] if a.size == mask.size else a.ravel())
        # a field with one element while the first has N marks global
        # parameters (utils/mcmc.m:54-63)
        names = list(fitting["modelParams"])
        n_first = x0[names[0]].size
        is_global = any(x0[n].size != n_first and x0[n].size == 1 for n in names[1:])
        if not is_global and any(x0[n].size != n_first for n in names[1:]):
            raise GacelleError(_abi.GC_ERR_ARG, "pars0 fields differ in size and none is a global (one-element) "
                                                "field after the first model parameter")
        extra = varargin[0] if varargin and isinstance(varargin[0], dict) else None
        if str(fitting["algorithm"]).lower() == "mh":
            if is_global:
                raise GacelleError(_abi.GC_ERR_UNSUPPORTED, "global parameters need algorithm 'ensemble' (GW)")
            post = self.metropolis_hastings(y, x0, w, fitting, FWDfunc, extra, _mask=mask, _vox=vox)
        elif not is_global:
            post = self.goodman_weare(y, x0, w, fitting, FWDfunc, extra, _mask=mask, _vox=vox)
        else:
            post = self.goodman_weare_wglobal_constant(y, x0, w, fitting, FWDfunc, extra, _mask=mask, _vox=vox)
        return self.res2out(post, fitting, mask)

    def metropolis_hastings(self, y, x0, weights, fitting, FWDfunc, *varargin, _mask=None, _vox=None):
        """utils/mcmc.m:82-190; y is G-D [Nmeas, Nv]."""
        return self._run(_abi.GC_MCMC_MH, y, x0, weights, fitting, FWDfunc, varargin, _mask, _vox)

    def goodman_weare(self, y, x0, weights, fitting, FWDfunc, *varargin, _mask=None, _vox=None):
        """utils/mcmc.m:192-383 (simultaneous or red-black update)."""
        return self._run(_abi.GC_MCMC_GW, y, x0, weights, fitting, FWDfunc, varargin, _mask, _vox)

    def goodman_weare_wglobal_constant(self, y, x0, weights, fitting, FWDfunc, *varargin, _mask=None, _vox=None):
        """utils/mcmc.m:385-626: block GW with local and global (shared)
        parameters.  A field of x0 with one element along the voxel
        dimension is global (check_global_constant, :760-770)."""
        names = list(self.check_set_default_basic(fitting)["modelParams"])
        gmask = 0
        for q, n in enumerate(names):
            if np.asarray(x0[n]).size == 1:
                gmask |= 1 << q
        return self._run(_abi.GC_MCMC_GW, y, x0, weights, fitting, FWDfunc, varargin, _mask, _vox, gmask=gmask)

    @staticmethod
    def check_global_constant(x_struct, fields) -> np.ndarray:
        """utils/mcmc.m:760-770: True where a field has one element along the
        voxel dimension."""
        return np.array([np.asarray(x_struct[f]).reshape(-1).size <= 1 for f in fields])

    def _run(self, algo, y, x0, weights, fitting, FWDfunc, varargin, mask, vox, gmask=0):
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
        extra_data = varargin[0] if varargin and isinstance(varargin[0], dict) else None
        md = model.gc_model(fitting, "mcmc", nmeas)
        if hasattr(model, "model_for_extra"):
            md = model.model_for_extra(md, extra_data)
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
        cfg.global_mask = gmask
        cfg.global_warmup = int(fitting["globalWarmup"]) if gmask else 0
        is_g = [bool((gmask >> q) & 1) for q in range(Q)]
        loc = [q for q in range(Q) if not is_g[q]]
        glo = [q for q in range(Q) if is_g[q]]
        if gmask and (not loc or not glo):
            raise GacelleError(_abi.GC_ERR_ARG, f"Need at least one local and one global parameter "
                                                f"(NND={len(loc)}, NGlobal={len(glo)}). Use goodman_weare for "
                                                f"the all-local case.")
        if gmask and self.verbose_enabled(fitting):
            print(f"Ensemble update scheme: {fitting['Ensembleupdate']}")
        steps = list(np.atleast_1d(fitting.get("xStepSize", [0] * Q)).astype(float))
        for q in range(Q):
            cfg.x_step[q] = steps[q] if q < len(steps) else 0.0
        R = cfg.repetition
        ns = 0 if cfg.iteration < cfg.n_burnin + 1 else (cfg.iteration - cfg.n_burnin - 1) // cfg.thinning + 1
        store = fitting["storePosterior"]
        post_elems = len(loc) * R * ns * W * nv
        if store is None:
            store = post_elems * 4 <= (1 << 30)
        ctx = runtime.context()
        xs = np.stack([np.broadcast_to(np.asarray(x0[n], dtype=np.float32).reshape(-1), (nv,)) for n in names])
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
            if gmask:
                gud = ctx.to_device(np.asarray(draws["g_uniforms"], np.float32))
                gpd = ctx.to_device(np.asarray(draws["g_partners"], np.int32), np.int32)
                keep += [gud, gpd]
                io.g_uniforms_dev = runtime.ptr(gud)
                io.g_partners_dev = runtime.ptr(gpd)
        NL, NG = len(loc), len(glo)
        T = cfg.global_warmup + cfg.iteration
        post_d = ctx.empty((NL, R, ns, W, nv)) if store and ns > 0 else None
        gpost_d = ctx.empty((NG, R, ns, W)) if gmask and store and ns > 0 else None
        mean_d = ctx.empty((Q, nv))
        std_d = ctx.empty((Q, nv))
        acc_d = ctx.empty((nv,))
        xf_d = ctx.empty((R, W, NL, nv))
        gxf_d = ctx.empty((R, W, NG)) if gmask else None
        trace_d = ctx.zeros((R, cfg.iteration, W, nv), np.uint8) if fitting.get("traceAccept") else None
        gtrace_d = ctx.zeros((R, T, W), np.uint8) if gmask and fitting.get("traceAccept") else None
        io.posterior_dev = runtime.ptr(post_d)
        io.mean_dev = runtime.ptr(mean_d)
        io.std_dev = runtime.ptr(std_d)
        io.accept_dev = runtime.ptr(acc_d)
        io.xfinal_dev = runtime.ptr(xf_d)
        io.trace_dev = runtime.ptr(trace_d)
        io.g_posterior_dev = runtime.ptr(gpost_d)
        io.g_xfinal_dev = runtime.ptr(gxf_d)
        io.g_trace_dev = runtime.ptr(gtrace_d)
        ctx.check(ctx.lib.gc_mcmc_run(ctx.handle, ctypes.byref(md), ctypes.byref(cfg), ctypes.byref(io)))
        res = {
            "names": names,
            "is_global": is_g,
            "mean": runtime.to_host(mean_d),
            "std": runtime.to_host(std_d),
            "accept": runtime.to_host(acc_d),
            "xfinal": runtime.to_host(xf_d),
            "nSample": int(io.n_sample),
            "nStored": int(io.n_stored),
            "algorithm": algo,
            "kernel_ms": float(io.kernel_ms),
        }
        if gmask:
            res["g_xfinal"] = runtime.to_host(gxf_d)
        if trace_d is not None:
            res["trace"] = runtime.to_host(trace_d)
        if gtrace_d is not None:
            res["g_trace"] = runtime.to_host(gtrace_d)
        wanted = [m for m in ("median", "iqr", "mode") if m in [x.lower() for x in fitting["metric"]]]
        if wanted:
            if post_d is None:
                raise GacelleError(_abi.GC_ERR_ARG, f"metric {wanted} needs fitting.storePosterior = True")
            # res2out median / iqr / mode on the device (utils/mcmc.m:823-857)
            lball = np.atleast_1d(np.asarray(fitting["lb"], dtype=np.float64))
            uball = np.atleast_1d(np.asarray(fitting["ub"], dtype=np.float64))
            nbin = int(fitting.get("Nbin", 1001))
            for m in wanted:
                res[m] = np.zeros((Q, nv), np.float32)
            for rows, dev, n_vox in ((loc, post_d, nv), (glo, gpost_d, 1)):
                if not rows:
                    continue
                k = len(rows)
                outs = {m: ctx.empty((k, n_vox)) for m in wanted}
                lb = (ctypes.c_float * k)(*[float(lball[q]) for q in rows])
                ub = (ctypes.c_float * k)(*[float(uball[q]) for q in rows])
                ctx.check(ctx.lib.gc_posterior_stats(ctx.handle, runtime.ptr(dev), k, R * ns * W, n_vox, lb, ub, nbin,
                                                     runtime.ptr(outs.get("median")), runtime.ptr(outs.get("iqr")),
                                                     runtime.ptr(outs.get("mode"))))
                for m, t in outs.items():
                    res[m][rows] = runtime.to_host(t)
        posterior = {}
        if post_d is not None:
            p = runtime.to_host(post_d)  # [NND, R, Ns, W, nv]
            for j, q in enumerate(loc):
                a = np.transpose(p[j], (3, 2, 1, 0))  # [nv, W, Ns, R]
                posterior[names[q]] = a[:, 0, :, :] if algo == _abi.GC_MCMC_MH else a
        if gpost_d is not None:
            gp = runtime.to_host(gpost_d)  # [NG, R, Ns, W]
            for i, q in enumerate(glo):
                posterior[names[q]] = np.transpose(gp[i], (2, 1, 0))[None]  # [1, W, Ns, R]
        res["posterior"] = posterior
        return res

    @staticmethod
    def verbose_enabled(fitting) -> bool:
        return bool(fitting.get("verbose", True))

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
                if res.get("is_global", [False] * len(names))[q]:
                    # a global parameter's statistic is one value (reshape_ND2image of a scalar)
                    out[m][n] = np.float32(res[m][q][0])
                else:
                    out[m][n] = utils.reshape_ND2image(np.asarray(res[m][q], np.float32), mask)
        out["acceptance"] = utils.reshape_ND2image(res["accept"], mask)
        return out
