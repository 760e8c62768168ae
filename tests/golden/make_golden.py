"""Write the golden fixtures of tests/golden/*.npz.

The reference (MATLAB) holds no golden vectors, known-answer tests or
fixtures for this path, and cannot run here.  These fixtures are therefore
produced by the FP64 oracle (oracle/), after the oracle itself has been
pinned by tests/test_oracle.py (closed forms, independent integrations,
finite differences, known-answer vectors).  They freeze small, seeded
cases — forward models, adjoints, askadam trajectories with stopping,
replayed MH and GW chains, the TV operator — so that the oracle and the HIP
path are both checked against committed numbers.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gacelle_amd import _abi, models  # noqa: E402
from gacelle_amd.askadam import askadam  # noqa: E402
from gacelle_amd.mcmc import mcmc  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests import golden_io  # noqa: E402

OUT = golden_io.GOLDEN_DIR
TE12 = np.arange(2, 47, 4) * 1e-3
TE_J = np.arange(2, 25, 2) * 1e-3
FA_J = np.array([3.0, 8.0, 15.0, 25.0])
TR_J = 30e-3


def save(name, **arrays):
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
    print("wrote", name, sum(np.asarray(v).nbytes for v in arrays.values()), "bytes")


def ask_fitting(names, lb, ub, **kw):
    return askadam.check_set_default_basic(dict(modelParams=names, lb=lb, ub=ub, **kw))


def r2s_md(uniform=True, solver="askadam"):
    te = TE12 if uniform else np.array([2, 5, 9, 14, 20, 27, 35, 44]) * 1e-3
    m = models.gpuR2starMapping(te)
    if solver == "askadam":
        f = ask_fitting(["M0", "R2star"], [0, 0.1], [2, 200])
    else:
        f = mcmc.check_set_default_basic(dict(modelParams=["M0", "R2star", "noise"], lb=[0, 0.1, 0.001],
                                              ub=[2, 200, 0.1]))
    return m.gc_model(f, solver, te.size)


def joint_md():
    m = models.gpuJointR1R2starMapping(TE_J, TR_J, FA_J)
    f = ask_fitting(["M0", "R1", "R2star"], [0, 0.1, 0.1], [2, 10, 200])
    return m.gc_model(f, "askadam", None)


def nexi_md(quad, lmax=0, rician=False, solver="askadam"):
    b = np.repeat(np.arange(1, 7, dtype=float), 2)
    D = np.tile([20.0, 30.0], 6)
    m = models.gpuNEXI(b, D, quad=quad, rician=rician)
    names = ["fa", "Da", "De", "ra"] + (["p2"] if lmax == 2 else [])
    if solver == "askadam":
        f = ask_fitting(names, m.lb[: len(names)], m.ub[: len(names)])
    else:
        names = names + ["noise"]
        f = mcmc.check_set_default_basic(dict(modelParams=names, lb=list(m.lb[: len(names) - 1]) + [0.01],
                                              ub=list(m.ub[: len(names) - 1]) + [0.1]))
    return m.gc_model(f, solver, None)


def forward_fixtures(rng):
    nv = 32
    cases = {
        "fwd_r2star12_uniform": (r2s_md(True), np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)]), None),
        "fwd_r2star8": (r2s_md(False), np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)]), None),
        "fwd_joint12x4": (joint_md(), np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv),
                                                rng.uniform(10, 60, nv)]), rng.uniform(0.8, 1.2, (1, nv))),
        "fwd_nexi_trapz14": (nexi_md("trapz14"), np.stack([rng.uniform(0.2, 0.8, nv), rng.uniform(1.5, 3, nv),
                                                            rng.uniform(0.5, 1.5, nv), rng.uniform(0.01, 0.3, nv)]),
                             None),
        "fwd_nexi_gl64_lmax2": (nexi_md("gl64", lmax=2), np.stack([rng.uniform(0.2, 0.8, nv),
                                                                   rng.uniform(1.5, 3, nv), rng.uniform(0.5, 1.5, nv),
                                                                   rng.uniform(0.01, 0.3, nv),
                                                                   rng.uniform(0.1, 0.6, nv)]), None),
        "fwd_nexi_rician": (nexi_md("trapz14", rician=True), np.stack([rng.uniform(0.2, 0.8, nv),
                                                                        rng.uniform(1.5, 3, nv),
                                                                        rng.uniform(0.5, 1.5, nv),
                                                                        rng.uniform(0.01, 0.3, nv)]),
                            np.full((1, nv), 0.02)),
    }
    for name, (md, th, ex) in cases.items():
        th = th.astype(np.float32)
        ex = None if ex is None else ex.astype(np.float32)
        s, j = orc.fwd(md, th, ex)
        arrays = dict(theta=th, s=s, jac=j, **golden_io.pack("model", md))
        if ex is not None:
            arrays["extra"] = ex
        save(name, **arrays)


def adjoint_fixtures(rng):
    nv = 32
    md = r2s_md(True)
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    z = rng.uniform(-2, 2, (2, nv)).astype(np.float32)
    for loss in ("L1", "L2", "huber", "mse"):
        lid = {"L1": _abi.GC_LOSS_L1, "L2": _abi.GC_LOSS_L2, "huber": _abi.GC_LOSS_HUBER, "mse": _abi.GC_LOSS_MSE}[loss]
        # d mean(loss) / d z: the oracle's dual-number Jacobian of the forward
        # model, the elementwise loss derivative and the sigmoid chain rule
        # (the composition of askadam.model_gradient, utils/askadam.m:76-105)
        lb = np.array([md.lb[0], md.lb[1]])[:, None]
        ub = np.array([md.ub[0], md.ub[1]])[:, None]
        zz = z.astype(np.float64)
        sg = 1 / (1 + np.exp(-zz))
        thz = lb + (ub - lb) * sg
        s, jac = orc.fwd(md, thz)
        r = s - y.astype(np.float64)
        if loss == "L1":
            l, dl = np.abs(r), np.sign(r)
        elif loss == "L2":
            l, dl = r * r, 2 * r
        elif loss == "huber":
            a = np.abs(r)
            l = np.where(a <= 1, 0.5 * r * r, a - 0.5)
            dl = np.where(a <= 1, r, np.sign(r))
        else:
            l, dl = 0.5 * r * r, r
        scale = 1.0 / r.size
        grad = np.einsum("mv,mpv->pv", dl, jac) * (ub - lb) * sg * (1 - sg) * scale
        save(f"adj_r2star12_{loss}", z=z, y=y, loss_sum=np.sum(l), grad=grad, loss_id=np.int32(lid),
             scale=np.float64(scale), **golden_io.pack("model", md))


def askadam_fixtures(rng):
    # R2*, L1, defaults with stopping on (tol, EMA convergence, step and gradient norms)
    nv = 64
    md = r2s_md(True)
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    f = ask_fitting(["M0", "R2star"], [0, 0.1], [2, 200], iteration=400, initialLearnRate=0.01,
                    convergenceValue=1e-4, tol=1e-6)
    cfg = askadam().make_cfg(f, None)
    z0 = np.zeros((2, nv), np.float32)
    o = orc.askadam(md, cfg, y, z0, trace=True)
    save("askadam_r2star12_l1_ema", y=y, z0=z0, z_final=o["z_final"], z_min=o["z_min"], trace=o["trace"],
         n_iteration=np.int32(o["n_iteration"]), min_iteration=np.int32(o["min_iteration"]),
         stop_reason=np.int32(o["stop_reason"]), min_loss=np.float64(o["min_loss"]),
         **golden_io.pack("model", md), **golden_io.pack("cfg", cfg))
    # R2*, L2 on noiseless data with a loss-tolerance stop placed where the
    # trace falls by more than 1% in one iteration (an FP32/FP64-robust stop)
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = orc.fwd_value(md, th).astype(np.float32)
    f = ask_fitting(["M0", "R2star"], [0, 0.1], [2, 200], iteration=600, lossFunction="L2",
                    initialLearnRate=0.02, tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0)
    cfg = askadam().make_cfg(f, None)
    tr = orc.askadam(md, cfg, y, z0, trace=True)["trace"]
    run_min = np.minimum.accumulate(tr)
    cand = [k for k in range(250, 550) if tr[k + 1] < 0.98 * run_min[k]]
    k = cand[0]
    cfg.tol = float(np.sqrt(run_min[k] * tr[k + 1]))
    o = orc.askadam(md, cfg, y, z0, trace=True)
    assert o["n_iteration"] == k + 1 and o["stop_reason"] == _abi.GC_STOP_TOL
    save("askadam_r2star12_l2_tol", y=y, z0=z0, z_final=o["z_final"], z_min=o["z_min"], trace=o["trace"],
         n_iteration=np.int32(o["n_iteration"]), min_iteration=np.int32(o["min_iteration"]),
         stop_reason=np.int32(o["stop_reason"]), min_loss=np.float64(o["min_loss"]),
         **golden_io.pack("model", md), **golden_io.pack("cfg", cfg))
    # joint R1-R2*, L2, fixed 200 iterations
    md = joint_md()
    thj = np.stack([rng.uniform(0.5, 2, nv), rng.uniform(0.5, 2, nv), rng.uniform(10, 60, nv)])
    b1 = rng.uniform(0.8, 1.2, (1, nv))
    y = (orc.fwd_value(md, thj, b1) + 0.005 * rng.standard_normal((48, nv))).astype(np.float32)
    f = ask_fitting(["M0", "R1", "R2star"], [0, 0.1, 0.1], [2, 10, 200], iteration=200, lossFunction="L2",
                    initialLearnRate=0.01, tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0)
    cfg = askadam().make_cfg(f, None)
    z0 = np.zeros((3, nv), np.float32)
    o = orc.askadam(md, cfg, y, z0, extra=b1.astype(np.float32), trace=True)
    save("askadam_joint12x4_l2", y=y, z0=z0, extra=b1.astype(np.float32), z_final=o["z_final"], z_min=o["z_min"],
         trace=o["trace"], n_iteration=np.int32(o["n_iteration"]), min_iteration=np.int32(o["min_iteration"]),
         stop_reason=np.int32(o["stop_reason"]), min_loss=np.float64(o["min_loss"]),
         **golden_io.pack("model", md), **golden_io.pack("cfg", cfg))


def mcmc_cfg(algo, iteration, thinning, burnin, rep, W, steps, redblack=False):
    c = _abi.gc_mcmc_cfg()
    c.algorithm, c.iteration, c.thinning, c.n_burnin = algo, iteration, thinning, burnin
    c.repetition, c.n_walker = rep, W
    c.ensemble_update = _abi.GC_ENSEMBLE_REDBLACK if redblack else _abi.GC_ENSEMBLE_SIMULTANEOUS
    c.rng_mode = _abi.GC_RNG_REPLAY
    c.seed, c.step_size, c.start_range = 48463, 2.0, 1e-3
    for q, s in enumerate(steps):
        c.x_step[q] = s
    return c


def mcmc_fixtures(rng):
    # MH replay, R2* + noise
    nv, it, R = 16, 60, 2
    md = r2s_md(True, "mcmc")
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv), np.full(nv, 0.02)])
    y = (orc.fwd_value(md, th) + 0.02 * rng.standard_normal((12, nv))).astype(np.float32)
    cfg = mcmc_cfg(_abi.GC_MCMC_MH, it, 5, 10, R, 1, (0.01, 1.0, 0.005))
    cfg.noise_slot, cfg.n_var = 2, 3
    normals = rng.standard_normal((R, it, nv, 3)).astype(np.float32)
    uniforms = rng.uniform(0, 1, (R, it, nv)).astype(np.float32)
    x0 = th.astype(np.float32)
    o = orc.mcmc(md, cfg, y, x0, normals=normals, uniforms=uniforms, trace=True, margin=True)
    save("mh_replay_r2star12", y=y, x0=x0, normals=normals, uniforms=uniforms, posterior=o["posterior"],
         mean=o["mean"], std=o["std"], accept=o["accept"], trace=o["trace"], margin=o["margin"],
         **golden_io.pack("model", md), **golden_io.pack("cfg", cfg))
    # GW replay (simultaneous and red-black), NEXI trapz14 + noise
    nv, it, R, W = 8, 30, 1, 8
    md = nexi_md("trapz14", solver="mcmc")
    th = np.stack([rng.uniform(0.3, 0.7, nv), rng.uniform(1.5, 3, nv), rng.uniform(0.5, 1.5, nv),
                   rng.uniform(0.02, 0.2, nv), np.full(nv, 0.02)])
    y = (orc.fwd_value(md, th) + 0.02 * rng.standard_normal((12, nv))).astype(np.float32)
    for redblack in (False, True):
        cfg = mcmc_cfg(_abi.GC_MCMC_GW, it, 2, 4, R, W, (), redblack)
        cfg.noise_slot, cfg.n_var = 4, 5
        normals = rng.standard_normal((W, nv, 5)).astype(np.float32)  # [W][nv][Q]
        uniforms = rng.uniform(0, 1, (R, it, 2, W, nv)).astype(np.float32)
        partners = np.zeros((R, it, W), np.int32)
        for r in range(R):
            for k in range(it):
                if redblack:
                    h = W // 2
                    partners[r, k, :h] = rng.integers(h, W, h)
                    partners[r, k, h:] = rng.integers(0, h, W - h)
                else:
                    while True:
                        p = rng.permutation(W)
                        if np.all(p != np.arange(W)):
                            break
                    partners[r, k] = p
        x0 = th.astype(np.float32)
        o = orc.mcmc(md, cfg, y, x0, normals=normals, uniforms=uniforms, partners=partners, trace=True, margin=True)
        save(f"gw_replay_nexi_{'redblack' if redblack else 'simultaneous'}", y=y, x0=x0, normals=normals,
             uniforms=uniforms, partners=partners, posterior=o["posterior"], mean=o["mean"], std=o["std"],
             accept=o["accept"], trace=o["trace"], margin=o["margin"],
             **golden_io.pack("model", md), **golden_io.pack("cfg", cfg))


def gw_c4_fixtures(rng):
    """Goodman-Weare replay at the NEXI benchmark settings: 64-point
    Gauss-Legendre quadrature, the Rician mean correction with its sigma row,
    16 walkers, simultaneous and red-black updates, Rician data."""
    nv, it, R, W = 12, 30, 1, 16
    md = nexi_md("gl64", rician=True, solver="mcmc")
    fa = rng.uniform(0.2, 0.8, nv)
    th = np.stack([fa, rng.uniform(1.5, 3, nv), rng.uniform(0.5, 1.5, nv), (1 - fa) / rng.uniform(2, 50, nv),
                   np.full(nv, 0.02)])
    sigma = np.full((1, nv), 0.02, np.float32)
    md0 = nexi_md("gl64", solver="mcmc")
    s = orc.fwd_value(md0, th)
    y = np.sqrt((s + 0.02 * rng.standard_normal(s.shape)) ** 2 + (0.02 * rng.standard_normal(s.shape)) ** 2)
    y = y.astype(np.float32)
    for redblack in (False, True):
        cfg = mcmc_cfg(_abi.GC_MCMC_GW, it, 2, 4, R, W, (), redblack)
        cfg.noise_slot, cfg.n_var = 4, 5
        normals = rng.standard_normal((W, nv, 5)).astype(np.float32)
        uniforms = rng.uniform(0, 1, (R, it, 2, W, nv)).astype(np.float32)
        partners = np.zeros((R, it, W), np.int32)
        h = W // 2
        for r in range(R):
            for k in range(it):
                if redblack:
                    partners[r, k, :h] = rng.integers(h, W, h)
                    partners[r, k, h:] = rng.integers(0, h, W - h)
                else:
                    while True:
                        p = rng.permutation(W)
                        if np.all(p != np.arange(W)):
                            break
                    partners[r, k] = p
        x0 = th.astype(np.float32)
        o = orc.mcmc(md, cfg, y, x0, extra=sigma, normals=normals, uniforms=uniforms, partners=partners, trace=True,
                     margin=True)
        save(f"gw_replay_nexi_gl64_rician_w16_{'redblack' if redblack else 'simultaneous'}", y=y, x0=x0,
             extra=sigma, normals=normals, uniforms=uniforms, partners=partners, posterior=o["posterior"],
             mean=o["mean"], std=o["std"], accept=o["accept"], trace=o["trace"], margin=o["margin"],
             **golden_io.pack("model", md), **golden_io.pack("cfg", cfg))


def tv_fixture(rng):
    dims = (7, 6, 5)
    mask = (rng.random(dims) < 0.75).astype(np.uint8)
    img = (rng.standard_normal(dims) * mask).astype(np.float32)
    n = int(mask.sum())
    for mode, name in ((_abi.GC_TV_2D, "2d"), (_abi.GC_TV_3D, "3d")):
        loss, g = orc.tv(img.astype(np.float64), mask, mode, (1.5, 1.0, 2.0), 0.3, n)
        save(f"tv_{name}", img=img, mask=mask, mode=np.int32(mode), vsize=np.array([1.5, 1.0, 2.0]),
             lam=np.float64(0.3), nmasked=np.int64(n), loss=np.float64(loss), grad=g)


def main():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(20260101)
    forward_fixtures(rng)
    adjoint_fixtures(rng)
    askadam_fixtures(rng)
    mcmc_fixtures(rng)
    tv_fixture(rng)
    gw_c4_fixtures(np.random.default_rng(20261204))


if __name__ == "__main__":
    main()
