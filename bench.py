"""Benchmark: voxels fitted/s (askadam) and MCMC samples/s on one MI355X.

Workload (one "step" = one whole-volume askadam fit): mono-exponential R2*
mapping of a 256x256x128 volume with a Bernoulli(0.7) mask (~5.87M voxels),
12 echoes TE = 2:4:46 ms, S0 ~ U(0.5,1.5), R2* ~ U(10,60) /s, Gaussian noise
at SNR 100; askadam 2000 iterations, L1 loss, Adam lr 1e-3, sigmoid
transform, stopping criteria off (fixed iteration count).  Metropolis-
Hastings: 4 chains (repetitions) x 5000 iterations per voxel, native RNG.

Reported beside the fused path, in the same run:
  * the plain per-op GPU tape (PyTorch eager autograd + Adam), same fit;
  * the CPU path (the oracle's FP32 restatement, OpenMP over host cores)
    on a bounded voxel sample;
  * the roofline position of the fused kernel from HIP-event kernel time.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

TE = np.arange(2, 47, 4) * 1e-3
FLOP_PER_VOXEL_ITER = 140.0      # 9M + 16P, M = 12 echoes, P = 2
TRANS_PER_VOXEL_ITER = 18.0      # M exp + P (exp + sqrt + div)
PEAK_FP32_TFLOPS = 157.3         # MI355X published FP32 vector (packed)
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # half of it; v_fma_f32 alone measures 77.3 on this GPU
PEAK_HBM_GBS = 8000.0
PEAK_FP64_MFMA_TFLOPS = 78.6     # MI355X published FP64 matrix


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "k_askadam_r2star12_5873183vox_traffic.md")


def measured_traffic():
    """HBM bytes per k_askadam launch on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated against a copy)."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", text)
    return int(m.group(1)) if m else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_volume(dims, frac, seed):
    rng = np.random.default_rng(seed)
    mask = rng.random(dims) < frac
    nv = int(mask.sum())
    s0 = rng.uniform(0.5, 1.5, nv).astype(np.float32)
    r2 = rng.uniform(10, 60, nv).astype(np.float32)
    sigma = float(s0.mean()) / 100.0
    return mask, nv, s0, r2, sigma


def synth_device(s0, r2, sigma, seed, device):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    te = torch.tensor(TE, dtype=torch.float32, device=device).reshape(-1, 1)
    S0 = torch.from_numpy(s0).to(device)
    R2 = torch.from_numpy(r2).to(device)
    y = S0.unsqueeze(0) * torch.exp(-te * R2.unsqueeze(0))
    y += sigma * torch.randn(y.shape, generator=g, device=device, dtype=torch.float32)
    return y.contiguous()


def r2s_model(lb, ub, n_par):
    """The mono-exponential model as the Python mirror builds it (equal
    echo spacing selects the exponential-recurrence kernel)."""
    from gacelle_amd import models
    from gacelle_amd.askadam import askadam
    from gacelle_amd.mcmc import mcmc

    names = ["M0", "R2star", "noise"][:n_par]
    if n_par == 2:
        f, solver = askadam.check_set_default_basic(dict(modelParams=names, lb=list(lb), ub=list(ub))), "askadam"
    else:
        f, solver = mcmc.check_set_default_basic(dict(modelParams=names, lb=list(lb), ub=list(ub))), "mcmc"
    return models.gpuR2starMapping(TE).gc_model(f, solver, TE.size)


def askadam_cfg(iterations):
    from gacelle_amd.askadam import askadam

    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200],
                                             iteration=iterations, initialLearnRate=1e-3, lossFunction="L1",
                                             tol=0, convergenceValue=-1, convergenceStepTol=0,
                                             convergenceGradTol=0, verbose=False))
    return askadam().make_cfg(f, None)


def logit_z0(nv, md, device):
    import torch

    z = np.zeros((2, 1), np.float64)
    for p, v in enumerate((1.0, 30.0)):
        t = (v - md.lb[p]) / (md.ub[p] - md.lb[p])
        z[p] = np.log(t / (1 - t))
    return torch.from_numpy(np.repeat(z, nv, axis=1).astype(np.float32)).to(device).contiguous()


def run_askadam(ctx, md, cfg, y, z0, n_mean=0, signs=None):
    """One whole fit (gc_askadam_run).  n_mean: N of the loss mean (0 = this
    volume's voxels); signs: a device int32 buffer [iteration, ceil(M/16), nv]
    that receives the residual sign trace."""
    from gacelle_amd import _abi, runtime

    nv = y.shape[1]
    z = z0.clone()
    zmin = ctx.empty((2, nv))
    rl = ctx.empty((nv,))
    v = (_abi.gc_volume * 1)()
    v[0].nv = nv
    v[0].y_dev = runtime.ptr(y)
    v[0].z_dev = runtime.ptr(z)
    v[0].zmin_dev = runtime.ptr(zmin)
    v[0].resloss_dev = runtime.ptr(rl)
    v[0].n_mean = int(n_mean)
    v[0].sign_trace_dev = runtime.ptr(signs)
    ctx.check(ctx.lib.gc_askadam_run(ctx.handle, ctypes.byref(md), ctypes.byref(cfg), v, 1))
    return z, v[0]


def subset_for_oracle(ctx, md, cfg, y, z0, zf, n=10000, seed=7):
    """A voxel subset of a whole fit, fitted again alone with the whole fit's
    N (gc_volume.n_mean) and its residual sign trace recorded: with stopping
    off those voxels follow exactly their trajectories in the whole fit,
    which is checked bit for bit.  Returns host arrays for the oracle check
    of the CPU leg."""
    import torch

    nv = y.shape[1]
    idx = torch.from_numpy(np.sort(np.random.default_rng(seed).choice(nv, min(n, nv), replace=False))).to(y.device)
    ys, zs = y[:, idx].contiguous(), z0[:, idx].contiguous()
    signs = torch.zeros((int(cfg.iteration), (md.n_meas + 15) // 16, idx.numel()), dtype=torch.int32, device=y.device)
    zsub, _ = run_askadam(ctx, md, cfg, ys, zs, n_mean=nv, signs=signs)
    same = bool(torch.equal(zsub, zf[:, idx]))
    return dict(y=ys.cpu().numpy(), z0=zs.cpu().numpy(), z=zsub.cpu().numpy(),
                signs=signs.cpu().numpy().view(np.uint32), n_mean=nv, same_as_whole_fit=same)


def mh_cfg(iterations, reps, seed=48463, replay=False):
    from gacelle_amd import _abi

    c = _abi.gc_mcmc_cfg()
    c.algorithm = _abi.GC_MCMC_MH
    c.iteration = iterations
    c.thinning = 20
    c.n_burnin = int(np.floor(iterations * 0.1))
    c.repetition = reps
    c.n_walker = 1
    c.rng_mode = _abi.GC_RNG_REPLAY if replay else _abi.GC_RNG_NATIVE
    c.noise_slot = 2
    c.n_var = 3
    c.seed = seed
    for q, s in enumerate((0.01, 1.0, 0.005)):
        c.x_step[q] = s
    return c


def run_mh(ctx, md, y, x0, iterations, reps, seed=48463, normals=None, uniforms=None):
    """gc_mcmc_run, MH; with normals / uniforms (device tensors in the
    reference's draw layout) in replay mode."""
    from gacelle_amd import _abi, runtime

    nv = y.shape[1]
    c = mh_cfg(iterations, reps, seed, replay=normals is not None)
    io = _abi.gc_mcmc_io()
    mean = ctx.empty((3, nv))
    std = ctx.empty((3, nv))
    io.nv = nv
    io.y_dev = runtime.ptr(y)
    io.x0_dev = runtime.ptr(x0)
    io.mean_dev = runtime.ptr(mean)
    io.std_dev = runtime.ptr(std)
    io.normals_dev = runtime.ptr(normals)
    io.uniforms_dev = runtime.ptr(uniforms)
    ctx.check(ctx.lib.gc_mcmc_run(ctx.handle, ctypes.byref(md), ctypes.byref(c), ctypes.byref(io)))
    return mean, std, io


def omp_threads():
    """Host threads for the CPU legs: the process's CPU share (OMP_NUM_THREADS
    when set, else every core in the affinity mask), and the host's core count."""
    ncores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else ncores
    return min(n, ncores), ncores


def theta_of(z, md):
    out = []
    for p in range(2):
        out.append(md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-z[p].astype(np.float64))))
    return np.stack(out)


def device_fwd(ctx, md, theta, extra=None):
    """Noiseless signal of the library's own forward model (gc_fwd)."""
    import torch

    from gacelle_amd import runtime

    nv = theta.shape[1]
    s = torch.empty((md.n_meas, nv), dtype=torch.float32, device=theta.device)
    ctx.check(ctx.lib.gc_fwd(ctx.handle, ctypes.byref(md), runtime.ptr(theta), runtime.ptr(extra), runtime.ptr(s), nv))
    return s


def timed(fn):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def bench_c1(ctx, device, iterations, mh_iterations=5000):
    """The CPU-runnable configuration: 100x100x1 fully masked (10,000 voxels),
    the same mono-exponential model, askadam (L1, lr 1e-3) and MH (5000
    iterations, one chain); the CPU path runs the whole volume."""
    import torch

    dims = (100, 100, 1)
    rng = np.random.default_rng(5438973)
    nv = 10000
    s0 = rng.uniform(0.5, 1.5, nv).astype(np.float32)
    r2 = rng.uniform(10, 60, nv).astype(np.float32)
    sigma = float(s0.mean()) / 100.0
    y = synth_device(s0, r2, sigma, 5438973, device)
    md = r2s_model([0, 0.1], [2, 200], 2)
    cfg = askadam_cfg(iterations)
    z0 = logit_z0(nv, md, device)
    run_askadam(ctx, md, cfg, y, z0)  # warm-up
    t_ask, (zf, vol) = timed(lambda: run_askadam(ctx, md, cfg, y, z0))
    sub = subset_for_oracle(ctx, md, cfg, y, z0, zf, n=nv)
    md3 = r2s_model([0, 0.1, 0.001], [2, 200, 0.1], 3)
    x0 = torch.stack([torch.from_numpy(s0), torch.from_numpy(r2),
                      torch.full((nv,), sigma, dtype=torch.float32)]).to(device).contiguous()
    run_mh(ctx, md3, y, x0, 50, 1)
    t_mh, (_, _, io) = timed(lambda: run_mh(ctx, md3, y, x0, mh_iterations, 1))
    yh, x0h = y.cpu().numpy(), x0.cpu().numpy()

    def cpu_path(nthreads, ncores):
        from oracle import oracle as orc

        t0 = time.perf_counter()
        orc.askadam(md, cfg, yh, sub["z0"], precision=1, nthreads=nthreads)
        t_a = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.mcmc(md3, mh_cfg(mh_iterations, 1), yh, x0h, store=False, precision=1, nthreads=nthreads)
        t_m = time.perf_counter() - t0
        return dict(kind="port", cores=nthreads, host_cores=ncores, askadam_voxels_per_s=nv / t_a,
                    mh_samples_per_s=nv * mh_iterations / t_m,
                    sample=f"the whole volume, oracle FP32 restatement on {nthreads} threads: askadam {t_a:.2f} s, "
                           f"MH {t_m:.2f} s")

    return dict(workload=f"mono-exponential R2* 100x100x1 (10000 voxels), 12 echoes 2:4:46 ms, SNR 100: askadam "
                         f"{iterations} iterations L1/Adam lr 1e-3, MH {mh_iterations} iterations x 1 chain",
                askadam_seconds=t_ask, voxels_per_s=nv / t_ask, askadam_kernel_ms=vol.kernel_ms,
                mh_seconds=t_mh, mh_samples_per_s=nv * mh_iterations / t_mh, mh_kernel_ms=io.kernel_ms,
                _oracle_check=lambda nthreads: oracle_check(md, cfg, sub, ["M0", "R2star"], nthreads),
                _cpu_baseline=cpu_path)


def bench_c3(ctx, device, iterations):
    """Joint R1-R2* VFA-SPGR, 4 flip angles x 12 echoes, TR 30 ms, three fully
    masked volumes of 0.4M / 0.9M / 1.6M voxels fitted in one batched launch."""
    import torch

    from gacelle_amd import _abi, models, runtime
    from gacelle_amd.askadam import askadam

    te = np.arange(2, 25, 2) * 1e-3
    m = models.gpuJointR1R2starMapping(te, 30e-3, np.array([3.0, 8.0, 15.0, 25.0]))
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R1", "R2star"], lb=[0, 0.1, 0.1], ub=[2, 10, 200],
                                             iteration=iterations, tol=0, convergenceValue=-1,
                                             convergenceStepTol=0, convergenceGradTol=0, verbose=False))
    md = m.gc_model(f, "askadam", None)
    cfg = askadam().make_cfg(f, m)
    g = torch.Generator(device=device)
    g.manual_seed(3)
    vols = []
    for nx, ny, nz in ((100, 100, 40), (150, 150, 40), (200, 200, 40)):
        nv = nx * ny * nz
        th = torch.stack([torch.rand(nv, generator=g, device=device) * 1.5 + 0.5,
                          torch.rand(nv, generator=g, device=device) * 1.5 + 0.5,
                          torch.rand(nv, generator=g, device=device) * 50 + 10]).contiguous()
        b1 = (torch.rand((1, nv), generator=g, device=device) * 0.4 + 0.8).contiguous()
        s = device_fwd(ctx, md, th, b1)
        sigma = float(th[0].mean()) / 80.0
        n1 = torch.randn(s.shape, generator=g, device=device)
        n2 = torch.randn(s.shape, generator=g, device=device)
        y = torch.sqrt((s + sigma * n1) ** 2 + (sigma * n2) ** 2).contiguous()
        # start at M0 = 1, R1 = 1 /s, R2* = 30 /s (logit space), as the R2* workload
        zc = [float(np.log(t / (1 - t))) for t in ((1.0 - 0) / 2, (1.0 - 0.1) / 9.9, (30.0 - 0.1) / 199.9)]
        z0 = torch.tensor(zc, dtype=torch.float32, device=device)[:, None].repeat(1, nv).contiguous()
        vols.append((y, b1, z0, th))

    def run(c=cfg):
        arr = (_abi.gc_volume * 3)()
        keep = []
        for i, (y, b1, z0, _) in enumerate(vols):
            z = z0.clone()
            arr[i].nv = y.shape[1]
            arr[i].y_dev = runtime.ptr(y)
            arr[i].extra_dev = runtime.ptr(b1)
            arr[i].z_dev = runtime.ptr(z)
            keep.append(z)
        ctx.check(ctx.lib.gc_askadam_run(ctx.handle, ctypes.byref(md), ctypes.byref(c), arr, 3))
        return arr, keep

    run()  # warm-up
    t, (arr, zs) = timed(run)
    nv_total = sum(v[0].shape[1] for v in vols)

    def gt_err(zlist):
        err = {}
        for p, name in enumerate(("M0", "R1", "R2star")):
            rel = []
            for (y, b1, z0, th), z in zip(vols, zlist):
                est = md.lb[p] + (md.ub[p] - md.lb[p]) * torch.sigmoid(z[p])
                rel.append(torch.abs(est - th[p]) / th[p])
            err[name] = float(torch.median(torch.cat(rel)))
        return err

    err = gt_err(zs)
    # the same fit converged (lr 1e-2, where the FP64 oracle reaches the error
    # of a 10x longer lr 1e-3 run): the error floor that the Rician noise at
    # SNR 80 sets for this protocol
    _, zs_c = run(askadam().make_cfg(dict(f, initialLearnRate=1e-2), m))
    floor = gt_err(zs_c)
    del zs_c
    # a 4000-voxel piece of the first volume fitted alone with that volume's N
    # and its sign trace, for the oracle check of the CPU leg
    y0, b10, z00, _ = vols[0]
    nv0 = y0.shape[1]
    idx = torch.from_numpy(np.sort(np.random.default_rng(9).choice(nv0, 4000, replace=False))).to(device)
    ys, bs, zs0 = y0[:, idx].contiguous(), b10[:, idx].contiguous(), z00[:, idx].contiguous()
    sg = torch.zeros((iterations, (md.n_meas + 15) // 16, idx.numel()), dtype=torch.int32, device=device)
    a1 = (_abi.gc_volume * 1)()
    zsub = zs0.clone()
    a1[0].nv = idx.numel()
    a1[0].y_dev, a1[0].extra_dev, a1[0].z_dev = runtime.ptr(ys), runtime.ptr(bs), runtime.ptr(zsub)
    a1[0].n_mean, a1[0].sign_trace_dev = nv0, runtime.ptr(sg)
    ctx.check(ctx.lib.gc_askadam_run(ctx.handle, ctypes.byref(md), ctypes.byref(cfg), a1, 1))
    sub = dict(y=ys.cpu().numpy(), z0=zs0.cpu().numpy(), z=zsub.cpu().numpy(), signs=sg.cpu().numpy().view(np.uint32),
               n_mean=nv0, same_as_whole_fit=bool(torch.equal(zsub, zs[0][:, idx])))
    b1h = bs.cpu().numpy()
    return dict(workload="joint R1-R2* 4 FA x 12 TE (TE 2:2:24 ms, TR 30 ms), Rician SNR 80, volumes 0.4M/0.9M/1.6M "
                         f"voxels in one batched launch, {iterations} iterations L1/Adam",
                voxels=nv_total, seconds=t, voxels_per_s=nv_total / t,
                voxel_iterations_per_s=nv_total * iterations / t, kernel_ms=arr[0].kernel_ms,
                launches=arr[0].kernel_launches, median_rel_err=err, median_rel_err_converged_lr1e2=floor,
                _oracle_check=lambda nthreads: oracle_check(md, cfg, sub, ["M0", "R1", "R2star"], nthreads,
                                                            extra=b1h))


def bench_c4(ctx, device, iterations, nv=500_000, walkers=16):
    """NEXI Karger exchange, 6 shells x 2 diffusion times, 64-point
    Gauss-Legendre quadrature, Rician SNR 50; Goodman-Weare with 16 walkers."""
    import torch

    from gacelle_amd import _abi, models, runtime
    from gacelle_amd.mcmc import mcmc

    b = np.repeat(np.arange(1, 7, dtype=float), 2)
    D = np.tile([20.0, 30.0], 6)
    m = models.gpuNEXI(b, D, quad="gl64", rician=True)
    names = ["fa", "Da", "De", "ra", "noise"]
    f = mcmc.check_set_default_basic(dict(modelParams=names, lb=list(m.lb[:4]) + [0.01], ub=list(m.ub[:4]) + [0.1]))
    md = m.gc_model(f, "mcmc", None)
    g = torch.Generator(device=device)
    g.manual_seed(4)
    fa = torch.rand(nv, generator=g, device=device) * 0.6 + 0.2
    tex = torch.rand(nv, generator=g, device=device) * 48 + 2
    th = torch.stack([fa, torch.rand(nv, generator=g, device=device) * 1.5 + 1.5,
                      torch.rand(nv, generator=g, device=device) + 0.5, (1 - fa) / tex,
                      torch.full((nv,), 0.02, device=device)]).contiguous()
    sigma = torch.full((1, nv), 0.02, device=device)
    # noiseless spherical mean without the Rician correction, then Rician noise
    md0 = models.gpuNEXI(b, D, quad="gl64").gc_model(
        mcmc.check_set_default_basic(dict(modelParams=names[:4], lb=list(m.lb[:4]), ub=list(m.ub[:4]))), "mcmc", None)
    md0.n_par = 5
    s = device_fwd(ctx, md0, th, None)
    y = torch.sqrt((s + 0.02 * torch.randn(s.shape, generator=g, device=device)) ** 2 +
                   (0.02 * torch.randn(s.shape, generator=g, device=device)) ** 2).contiguous()
    c = _abi.gc_mcmc_cfg()
    c.algorithm, c.iteration, c.thinning = _abi.GC_MCMC_GW, iterations, 20
    c.n_burnin = int(iterations * 0.1)
    c.repetition, c.n_walker = 1, walkers
    c.ensemble_update = _abi.GC_ENSEMBLE_SIMULTANEOUS
    c.rng_mode, c.noise_slot, c.n_var, c.seed = _abi.GC_RNG_NATIVE, 4, 5, 48463
    c.step_size, c.start_range = 2.0, 1e-3
    mean = torch.empty((5, nv), device=device)
    std = torch.empty((5, nv), device=device)

    def run(cfg):
        io = _abi.gc_mcmc_io()
        io.nv = nv
        io.y_dev = runtime.ptr(y)
        io.extra_dev = runtime.ptr(sigma)
        io.x0_dev = runtime.ptr(th)
        io.mean_dev = runtime.ptr(mean)
        io.std_dev = runtime.ptr(std)
        ctx.check(ctx.lib.gc_mcmc_run(ctx.handle, ctypes.byref(md), ctypes.byref(cfg), ctypes.byref(io)))
        return io

    warm = _abi.gc_mcmc_cfg.from_buffer_copy(c)
    warm.iteration, warm.n_burnin = 20, 2
    run(warm)
    t, io = timed(lambda: run(c))
    steps = float(nv) * walkers * iterations
    err = {n: float(torch.median(torch.abs(mean[q] - th[q]) / th[q])) for q, n in enumerate(names[:4])}
    # how wide the posterior is: a parameter whose posterior SD is comparable to
    # (or larger than) its true value is not identified by these data, and its
    # posterior mean sits about one SD from the truth by construction
    width = {n: float(torch.median(std[q] / th[q])) for q, n in enumerate(names[:4])}
    zdev = {n: float(torch.median((mean[q] - th[q]) / std[q].clamp_min(1e-12))) for q, n in enumerate(names[:4])}
    return dict(workload=f"NEXI (6 shells x 2 Delta, GL64, Rician mean) Goodman-Weare, {nv} voxels x {walkers} walkers x "
                         f"{iterations} steps (simultaneous update, native stream)",
                seconds=t, walker_steps_per_s=steps / t, kernel_ms=io.kernel_ms, block_threads=io.block_threads,
                transcendentals_per_s=steps * 3072 / t, posterior_mean_median_rel_err=err,
                posterior_sd_over_true_median=width, posterior_mean_minus_true_over_sd_median=zdev)


def bench_c5(ctx, device, iterations, dims=(128, 128, 64), epg=False):
    """MCR-MWI complex 3-pool GRE (BM exchange), 24 echoes x 4 flip angles,
    complex Gaussian SNR 150, askadam with 3D TV on MWF (one launch per
    iteration through the step kernel).  epg: the EPG-X MLP T1 path on top
    (the k_epgx pre-pass before every step kernel)."""
    import torch

    from gacelle_amd import _abi, models, runtime, utils
    from gacelle_amd.askadam import askadam

    te = np.arange(2, 49, 2) * 1e-3
    m = models.gpuMCRMWI(te, 38e-3, np.array([5.0, 10.0, 20.0, 50.0]))
    # epg: the reference's default T1 path with its bundled trained networks
    # (11-20-30-40-45-60-65-75-1 / -2); otherwise the BM steady state
    f = m.check_set_default(dict(isComplex=True, isFitExchange=True, isEPG=bool(epg)))
    m.updateProperty(f)
    f.update(modelParams=list(m.modelParams), lb=list(m.lb), ub=list(m.ub), iteration=iterations, tol=0,
             convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0, TVmode="3D", voxelSize=[1, 1, 1],
             verbose=False, **{"lambda": [1e-3]})
    f = askadam.check_set_default_basic(f)
    md = m.gc_model(f, "askadam", None)
    keep = []
    if epg:
        md = m.bind_device(md, ctx, keep)
    cfg = askadam().make_cfg(f, m)
    nx, ny, nz = dims
    nv = nx * ny * nz
    g = torch.Generator(device=device)
    g.manual_seed(5)
    # piecewise-smooth ground truth: smooth fields with a step along x
    xs = torch.linspace(0, 1, nx, device=device)[:, None, None].expand(nx, ny, nz)
    ys = torch.linspace(0, 1, ny, device=device)[None, :, None].expand(nx, ny, nz)
    zs = torch.linspace(0, 1, nz, device=device)[None, None, :].expand(nx, ny, nz)
    step = (xs > 0.5).float()

    def field(lo, hi, ph):
        v = 0.5 + 0.25 * torch.sin(2 * np.pi * (xs + ph)) * torch.cos(2 * np.pi * (ys - ph)) + 0.2 * step - 0.1
        return (lo + (hi - lo) * v.clamp(0, 1)).permute(2, 1, 0).reshape(-1)  # column-major voxel order

    maps = {"S0": field(0.8, 1.2, 0.1), "MWF": field(0.05, 0.2, 0.2), "IWF": field(0.4, 0.8, 0.3),
            "R1IEW": field(0.8, 1.4, 0.4), "kIEWM": field(1.0, 3.0, 0.5), "R2sMW": field(80, 150, 0.6),
            "R2sIW": field(15, 30, 0.7), "R2sEW": field(15, 30, 0.8), "freqMW": field(0.02, 0.1, 0.9),
            "freqIW": field(-0.05, 0.0, 0.15)}
    rows = [maps[n] for n in m.modelParams if n in maps]
    rows += [torch.zeros(nv, device=device)] * 4 + [torch.zeros(nv, device=device)]  # dfreqBKG x 4, dpini
    th = torch.stack(rows).contiguous()
    ex = torch.zeros((7, nv), device=device)
    ex[0] = 1.0
    ex[2] = 1.0
    ex[3:] = torch.randn((4, nv), generator=g, device=device) * 0.02
    s = device_fwd(ctx, md, th, ex)
    sigma = float(maps["S0"].mean()) / 150.0
    y = (s + sigma * torch.randn(s.shape, generator=g, device=device)).contiguous()
    mask = np.ones(dims, bool)
    vox = torch.from_numpy(utils.find_mask(mask)).to(device)
    z0 = torch.zeros((md.n_par, nv), device=device)

    def run(c):
        z = z0.clone()
        arr = (_abi.gc_volume * 1)()
        arr[0].nv = nv
        arr[0].y_dev = runtime.ptr(y)
        arr[0].extra_dev = runtime.ptr(ex)
        arr[0].vox_dev = runtime.ptr(vox)
        for k in range(3):
            arr[0].dims[k] = dims[k]
        arr[0].z_dev = runtime.ptr(z)
        ctx.check(ctx.lib.gc_askadam_run(ctx.handle, ctypes.byref(md), ctypes.byref(c), arr, 1))
        return arr[0]

    warm = _abi.gc_askadam_cfg.from_buffer_copy(cfg)
    warm.iteration = 5
    run(warm)
    t, v = timed(lambda: run(cfg))
    vi = float(nv) * iterations
    bytes_per_vi = 4 * md.n_meas + 2 * 3 * md.n_par * 4 + 6 * 4  # y, (z, m, v) r/w, TV neighbour reads
    path = "EPG-X MLP (the reference's trained networks 11-20-30-40-45-60-65-75-1 / -2) on BM" if epg else "BM"
    extra_out = {}
    if epg:
        # per (voxel, flip angle): the forward pass of each network (sum over
        # its layers of w_l w_(l-1) FMAs) and one reverse pass per output
        # (the same sum without the last layer, whose row is used directly)
        hid = [11, 20, 30, 40, 45, 60, 65, 75]
        fwd = lambda w: sum(w[l] * w[l - 1] for l in range(1, len(w)))
        per_sample = fwd(hid + [1]) + fwd(hid) + fwd(hid + [2]) + 2 * fwd(hid)
        extra_out = dict(epgx_flop_per_voxel_iteration=2 * 4 * per_sample)
    return dict(workload=f"MCR-MWI complex {path} 24 TE x 4 FA, {dims[0]}x{dims[1]}x{dims[2]} ({nv} voxels), 3D TV "
                         f"on MWF lambda 1e-3, {iterations} iterations L1/Adam, one step launch per iteration",
                seconds=t, voxel_iterations_per_s=vi / t, kernel_ms=v.kernel_ms, launches=v.kernel_launches,
                avg_launch_ms=v.kernel_ms / max(v.kernel_launches, 1),
                hbm_gbs_algorithmic=vi * bytes_per_vi / (v.kernel_ms * 1e-3) / 1e9,
                loss=v.loss, loss_reg=v.loss_reg, **extra_out)


def bench_tv_small(ctx, device, iterations, dims=(100, 100, 1)):
    """Launch-bound step form: the 100x100 mono-exponential slice with 2D TV
    on R2* (one k_askadam_step per iteration, 10,000 voxels), with the
    iteration blocks replayed from captured hipGraphs and, beside it, with
    one launch per iteration."""
    import torch

    from gacelle_amd import _abi, runtime, utils
    from gacelle_amd.askadam import askadam

    nx, ny, nz = dims
    nv = nx * ny * nz
    g = torch.Generator(device=device)
    g.manual_seed(11)
    s0 = (torch.rand(nv, generator=g, device=device) + 0.5)
    r2 = torch.rand(nv, generator=g, device=device) * 50 + 10
    te = torch.from_numpy(TE.astype(np.float32)).to(device)
    y = (s0[None] * torch.exp(-te[:, None] * r2[None]))
    y = (y + float(s0.mean()) / 100 * torch.randn(y.shape, generator=g, device=device)).contiguous()
    f = askadam.check_set_default_basic(dict(modelParams=["M0", "R2star"], lb=[0, 0.1], ub=[2, 200],
                                             iteration=iterations, initialLearnRate=1e-3, lossFunction="L1",
                                             tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0,
                                             regmap=["R2star"], TVmode="2D", voxelSize=[1, 1, 1], verbose=False,
                                             **{"lambda": [1e-3]}))
    md = r2s_model([0, 0.1], [2, 200], 2)
    cfg = askadam().make_cfg(f, None)
    vox = torch.from_numpy(utils.find_mask(np.ones(dims, bool))).to(device)
    z0 = logit_z0(nv, md, device)

    def run(c):
        z = z0.clone()
        arr = (_abi.gc_volume * 1)()
        arr[0].nv = nv
        arr[0].y_dev = runtime.ptr(y)
        arr[0].vox_dev = runtime.ptr(vox)
        for k in range(3):
            arr[0].dims[k] = dims[k]
        arr[0].z_dev = runtime.ptr(z)
        ctx.check(ctx.lib.gc_askadam_run(ctx.handle, ctypes.byref(md), ctypes.byref(c), arr, 1))
        return arr[0], z

    warm = _abi.gc_askadam_cfg.from_buffer_copy(cfg)
    warm.iteration = 70
    out = {}
    for name, opt in (("graphs", 1), ("direct", 0)):
        ctx.set_option(_abi.GC_OPT_GRAPHS, opt)
        run(warm)
        t, (v, z) = timed(lambda: run(cfg))
        out[name] = dict(seconds=t, us_per_iteration=t / iterations * 1e6, kernel_ms=v.kernel_ms, loss=v.loss,
                         z=z)
    ctx.set_option(_abi.GC_OPT_GRAPHS, 1)
    same = bool(torch.equal(out["graphs"].pop("z"), out["direct"].pop("z")))
    return dict(workload=f"mono-exponential R2* {nx}x{ny}x{nz} ({nv} voxels), 2D TV on R2* lambda 1e-3, "
                         f"{iterations} iterations L1/Adam, one k_askadam_step per iteration",
                graphs=out["graphs"], direct=out["direct"], speedup=out["direct"]["seconds"] / out["graphs"]["seconds"],
                bit_identical=same, voxels_per_s=nv / out["graphs"]["seconds"])


def bench_dict(ctx, device, steps, nv=1_000_000, nd=10_000):
    """NEXI likelihood start (NEXI/gpuNEXI.m:675-849): the reference's
    10,000-entry dictionary (4 b-values x 2 diffusion times; lmax 0 and 2)
    searched for 1M voxels of scaled, noisy dictionary signals; the FP64
    oracle on a 200,000-voxel sample beside it (cpu_agreement: share of that
    sample where both pick the same entry)."""
    import torch

    from gacelle_amd import models, runtime
    from oracle import oracle as orc

    b = np.array([0, 1, 2.5, 4, 6] * 2, float)
    D = np.array([20] * 5 + [30] * 5, float)
    out = {}
    for lmax in (0, 2):
        m = models.gpuNEXI(b, D)
        x, S = m.traindata(nd, lmax)
        M = S.shape[0]
        g = torch.Generator(device=device)
        g.manual_seed(17 + lmax)
        Sd = torch.from_numpy(np.ascontiguousarray(S, np.float32)).to(device)
        k = torch.randint(0, nd, (nv,), generator=g, device=device)
        y = (Sd[:, k] * (torch.rand(nv, generator=g, device=device) * 0.2 + 0.8)
             + 0.01 * torch.randn((M, nv), generator=g, device=device)).contiguous()
        best = torch.empty(nv, dtype=torch.int32, device=device)
        wt = np.kron(m.Nav.reshape(-1, 1), 1.0 / (2 * np.arange(0, lmax + 1, 2) + 1)).ravel(order="F")
        w = None if np.all(wt == 1) else (ctypes.c_float * M)(*wt.astype(np.float32))

        def call():
            ctx.check(ctx.lib.gc_dict_search(ctx.handle, runtime.ptr(y), nv, M, runtime.ptr(Sd), nd, w,
                                             runtime.ptr(best), None))

        call()
        ts = []
        for _ in range(max(steps, 1)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            call()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        t = float(np.median(ts))
        # MFMA products: a_j (and 2 b_j when weighted), 2 M flop per voxel-entry each
        flop = 2.0 * M * nv * nd * (1 if w is None else 2)
        ncpu = 200_000
        yc = y[:, :ncpu].cpu().numpy()
        nthreads = min(len(os.sched_getaffinity(0)), 16)
        t0 = time.perf_counter()
        bo, _ = orc.dict_search(yc, S, None if w is None else wt, threads=nthreads)
        t_cpu = time.perf_counter() - t0
        agree = float(np.mean(bo == best[:ncpu].cpu().numpy()))
        out[f"lmax{lmax}"] = dict(
            workload=f"{nd} entries x {nv} voxels, {M} measurements per voxel"
                     + (", weights 1 and 1/5" if w is not None else ""),
            seconds=t, voxels_per_s=nv / t, voxel_entries_per_s=nv * nd / t,
            mfma_tflops=flop / t * 1e-12, frac_fp64_mfma_peak=flop / t * 1e-12 / PEAK_FP64_MFMA_TFLOPS,
            cpu_agreement=agree,
            cpu_baseline=dict(voxels_per_s=ncpu / t_cpu, cores=nthreads, kind="port",
                              sample=f"oracle FP64 search of {ncpu} voxels, OpenMP {nthreads} threads: "
                                     f"{t_cpu:.2f} s"))
        log(f"dict lmax{lmax}: {t * 1e3:.2f} ms per call ({nv / t:.3e} voxels/s), CPU {ncpu / t_cpu:.3e} voxels/s")
        del y, Sd, best
    return out


def cpu_baseline_askadam(s0, r2, sigma, iterations, nthreads, nsample, seed=7):
    """The CPU path: the oracle's FP32 restatement, OpenMP over host cores."""
    from oracle import oracle as orc

    rng = np.random.default_rng(seed)
    idx = np.arange(nsample)
    S = s0[idx, None].T * np.exp(-np.outer(TE, r2[idx]))
    y = (S + sigma * rng.standard_normal(S.shape)).astype(np.float32)
    md = r2s_model([0, 0.1], [2, 200], 2)
    cfg = askadam_cfg(iterations)
    z0 = np.zeros((2, nsample), np.float32)
    for p, v in enumerate((1.0, 30.0)):
        t = (v - md.lb[p]) / (md.ub[p] - md.lb[p])
        z0[p] = np.log(t / (1 - t))
    t0 = time.perf_counter()
    orc.askadam(md, cfg, y, z0, precision=1, nthreads=nthreads)
    return time.perf_counter() - t0


def cpu_baseline_mh(s0, r2, sigma, iterations, reps, nthreads, nsample, seed=8):
    from gacelle_amd import _abi
    from oracle import oracle as orc

    rng = np.random.default_rng(seed)
    S = s0[:nsample, None].T * np.exp(-np.outer(TE, r2[:nsample]))
    y = (S + sigma * rng.standard_normal(S.shape)).astype(np.float32)
    md = r2s_model([0, 0.1, 0.001], [2, 200, 0.1], 3)
    c = _abi.gc_mcmc_cfg()
    c.algorithm, c.iteration, c.thinning = _abi.GC_MCMC_MH, iterations, 20
    c.n_burnin, c.repetition, c.n_walker, c.rng_mode = int(iterations * 0.1), reps, 1, _abi.GC_RNG_NATIVE
    c.noise_slot, c.n_var, c.seed = 2, 3, 48463
    for q, s in enumerate((0.01, 1.0, 0.005)):
        c.x_step[q] = s
    x0 = np.stack([s0[:nsample], r2[:nsample], np.full(nsample, sigma, np.float32)]).astype(np.float32)
    t0 = time.perf_counter()
    orc.mcmc(md, c, y, x0, store=False, precision=1, nthreads=nthreads)
    return time.perf_counter() - t0


def _sigmoid_theta(z, md, P):
    z = np.asarray(z, np.float64)
    return np.stack([md.lb[p] + (md.ub[p] - md.lb[p]) / (1 + np.exp(-z[p])) for p in range(P)])


def oracle_check(md, cfg, sub, names, nthreads, extra=None):
    """Checker run of the CPU leg: the device fit of a subset (see
    subset_for_oracle) against the FP64 oracle following the device's residual
    signs, and against the oracle with its own signs.  Per-parameter relative
    errors (median, 99.9th percentile, max)."""
    from oracle import oracle as orc

    P = len(names)
    kw = dict(extra=extra, n_mean=sub["n_mean"], nthreads=nthreads)
    t0 = time.perf_counter()
    matched = orc.askadam(md, cfg, sub["y"], sub["z0"], sign_trace=sub["signs"], **kw)
    own = orc.askadam(md, cfg, sub["y"], sub["z0"], **kw)
    th = _sigmoid_theta(sub["z"], md, P)
    out = {"voxels": int(sub["y"].shape[1]), "subset_equals_whole_fit": sub["same_as_whole_fit"],
           "oracle_seconds": time.perf_counter() - t0}
    for key, ref in (("matched_signs", matched), ("own_signs", own)):
        tr = _sigmoid_theta(ref["z_final"], md, P)
        e = np.abs(th - tr) / np.abs(tr)
        out[key] = {n: {"median": float(np.median(e[p])), "p999": float(np.quantile(e[p], 0.999)),
                        "max": float(e[p].max())} for p, n in enumerate(names)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iterations", type=int, default=2000)
    ap.add_argument("--mh-iterations", type=int, default=5000)
    ap.add_argument("--mh-chains", type=int, default=4)
    ap.add_argument("--dims", type=int, nargs=3, default=[256, 256, 128])
    ap.add_argument("--skip-tape", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-mh", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="run a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--configs", default="c1,c3,c4,c5,c5epg,tvsmall,dict",
                    help="further workloads to measure (c1, c3, c4, c5, c5epg, tvsmall, dict or none)")
    ap.add_argument("--c3-iterations", type=int, default=2000)
    ap.add_argument("--c4-iterations", type=int, default=5000)
    ap.add_argument("--c5-iterations", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported by this build "
              "(the path shards over voxels; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch

    from gacelle_amd import runtime
    from gacelle_amd.tape import askadam_r2star_tape

    device = torch.device("cuda", 0)
    ctx = runtime.context(0)
    dims = tuple(args.dims)
    mask, nv, s0, r2, sigma = make_volume(dims, 0.7, 23439)
    log(f"volume {dims}, {nv} masked voxels")
    y = synth_device(s0, r2, sigma, 5438973, device)
    md = r2s_model([0, 0.1], [2, 200], 2)
    cfg = askadam_cfg(args.iterations)
    z0 = logit_z0(nv, md, device)

    if args.calib_copy:
        a = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=device)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB write")
        del a, b

    # ---- fused askadam --------------------------------------------------
    for w in range(args.warmup):
        run_askadam(ctx, md, cfg, y, z0)
        log(f"warmup {w + 1}/{args.warmup} done")
    times, kms, kl = [], [], []
    zf = vol = None
    torch.cuda.synchronize()
    for s in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        zf, vol = run_askadam(ctx, md, cfg, y, z0)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        kms.append(vol.kernel_ms)
        kl.append(vol.kernel_launches)
        log(f"askadam step {s + 1}/{args.steps}: {times[-1] * 1e3:.2f} ms (kernel {vol.kernel_ms:.2f} ms, "
            f"{vol.kernel_launches} launches, K={vol.iters_per_launch}, {vol.block_threads} threads per block)")
    t_step = float(np.median(times))
    vox_per_s = nv / t_step
    zf_h = zf.cpu().numpy()
    th = theta_of(zf_h, md)
    err_fused = dict(M0=float(np.median(np.abs(th[0] - s0) / s0)), R2star=float(np.median(np.abs(th[1] - r2) / r2)))
    fused_loss = float(vol.loss)

    # roofline of the fused kernel: algorithmic FLOP per voxel-iteration
    k_ms = float(np.median(kms))
    units = float(nv) * args.iterations
    achieved_tflops = units * FLOP_PER_VOXEL_ITER / (k_ms * 1e-3) / 1e12
    trans_rate = units * TRANS_PER_VOXEL_ITER / (k_ms * 1e-3)
    launches = int(np.median(kl))
    # a 10^4-voxel subset of this fit for the oracle check of the CPU leg
    sub = None if args.skip_cpu else subset_for_oracle(ctx, md, cfg, y, z0, zf)

    # ---- plain per-op tape -----------------------------------------------
    tape = None
    if not args.skip_tape:
        te_t = torch.tensor(TE, dtype=torch.float32, device=device)
        torch.cuda.synchronize()
        askadam_r2star_tape(y, te_t, [0.0, 0.1], [2.0, 200.0], z0, 5)  # warm-up (allocator, kernels)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        zt, tape_loss = askadam_r2star_tape(y, te_t, [0.0, 0.1], [2.0, 200.0], z0, args.iterations)
        torch.cuda.synchronize()
        t_tape = time.perf_counter() - t0
        tht = theta_of(zt.cpu().numpy(), md)
        tape = dict(voxels_per_s=nv / t_tape, seconds=t_tape, final_loss=tape_loss,
                    speedup_fused_vs_tape=t_tape / t_step,
                    loss_rel_diff=abs(tape_loss - fused_loss) / abs(tape_loss),
                    median_rel_err=dict(M0=float(np.median(np.abs(tht[0] - s0) / s0)),
                                        R2star=float(np.median(np.abs(tht[1] - r2) / r2))))
        log(f"tape: {t_tape:.2f} s, loss {tape_loss:.6e} vs fused {fused_loss:.6e}")
        del zt

    # ---- MH -----------------------------------------------------------------
    mh = None
    if not args.skip_mh:
        md3 = r2s_model([0, 0.1, 0.001], [2, 200, 0.1], 3)
        x0 = torch.stack([torch.from_numpy(s0), torch.from_numpy(r2),
                          torch.full((nv,), sigma, dtype=torch.float32)]).to(device).contiguous()
        run_mh(ctx, md3, y, x0, 50, args.mh_chains)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mean, std, io = run_mh(ctx, md3, y, x0, args.mh_iterations, args.mh_chains)
        torch.cuda.synchronize()
        t_mh = time.perf_counter() - t0
        steps_total = float(nv) * args.mh_chains * args.mh_iterations
        mh = dict(samples_per_s=steps_total / t_mh, retained_samples_per_s=float(nv) * args.mh_chains *
                  io.n_sample / t_mh, seconds=t_mh, kernel_ms=io.kernel_ms, block_threads=io.block_threads,
                  chains=args.mh_chains,
                  iterations=args.mh_iterations, retained_per_chain=int(io.n_sample))
        m = mean.cpu().numpy()
        mh["posterior_mean_median_rel_err"] = dict(M0=float(np.median(np.abs(m[0] - s0) / s0)),
                                                   R2star=float(np.median(np.abs(m[1] - r2) / r2)))
        log(f"MH: {t_mh:.2f} s, {mh['samples_per_s']:.3e} chain-steps/s")
        # replay mode (the reference's draw layout, a caller-supplied stream) against
        # the native stream on a 10^4-voxel subset: the replay draws alone are
        # 4 x 5000 x 10^4 x (3 + 1) floats = 3.2 GB
        nsub = min(10000, nv)
        ys, xs = y[:, :nsub].contiguous(), x0[:, :nsub].contiguous()
        gen = torch.Generator(device=device)
        gen.manual_seed(11)
        normals = torch.randn((args.mh_chains, args.mh_iterations, nsub, 3), generator=gen, device=device)
        uniforms = torch.rand((args.mh_chains, args.mh_iterations, nsub), generator=gen, device=device)
        run_mh(ctx, md3, ys, xs, args.mh_iterations, args.mh_chains, normals=normals, uniforms=uniforms)  # warm-up
        t_rep, (_, _, io_r) = timed(lambda: run_mh(ctx, md3, ys, xs, args.mh_iterations, args.mh_chains,
                                                   normals=normals, uniforms=uniforms))
        t_nat, (_, _, io_n) = timed(lambda: run_mh(ctx, md3, ys, xs, args.mh_iterations, args.mh_chains))
        steps_sub = float(nsub) * args.mh_chains * args.mh_iterations
        mh["replay_vs_native_subset"] = dict(
            voxels=nsub, replay_samples_per_s=steps_sub / t_rep, native_samples_per_s=steps_sub / t_nat,
            replay_kernel_ms=io_r.kernel_ms, native_kernel_ms=io_n.kernel_ms,
            replay_draw_bytes=int(normals.numel() * 4 + uniforms.numel() * 4))
        log(f"MH replay vs native ({nsub} voxels): {t_rep * 1e3:.1f} ms vs {t_nat * 1e3:.1f} ms")
        del normals, uniforms
        # the plain per-op GPU path of the reference's MH (one kernel per
        # operation, repetitions in turn) on the first 500,000 voxels, beside
        # the fused sampler's result at the same voxels
        if not args.skip_tape:
            from gacelle_amd.tape import mh_r2star_tape

            ntp = min(500_000, nv)
            te_t = torch.tensor(TE, dtype=torch.float32, device=device)
            gen_t = torch.Generator(device=device)
            gen_t.manual_seed(12)
            yt, xt = y[:, :ntp].contiguous(), x0[:, :ntp].contiguous()
            mh_r2star_tape(yt, te_t, xt, [0, 0.1, 0.001], [2, 200, 0.1], [0.01, 1, 0.005], 10, 1, 20, 1, gen_t)
            nburn = int(args.mh_iterations * 0.1)
            t_tp, mt = timed(lambda: mh_r2star_tape(yt, te_t, xt, [0, 0.1, 0.001], [2, 200, 0.1], [0.01, 1, 0.005],
                                                     args.mh_iterations, args.mh_chains, 20, nburn, gen_t))
            mt = mt.cpu().numpy()
            steps_tp = float(ntp) * args.mh_chains * args.mh_iterations
            mh["tape_subset"] = dict(
                voxels=ntp, samples_per_s=steps_tp / t_tp, seconds=t_tp,
                speedup_fused_vs_tape=mh["samples_per_s"] / (steps_tp / t_tp),
                posterior_mean_median_rel_err=dict(M0=float(np.median(np.abs(mt[0] - s0[:ntp]) / s0[:ntp])),
                                                   R2star=float(np.median(np.abs(mt[1] - r2[:ntp]) / r2[:ntp]))),
                fused_posterior_mean_median_rel_err=dict(
                    M0=float(np.median(np.abs(m[0, :ntp] - s0[:ntp]) / s0[:ntp])),
                    R2star=float(np.median(np.abs(m[1, :ntp] - r2[:ntp]) / r2[:ntp]))))
            log(f"MH tape ({ntp} voxels): {t_tp:.2f} s, {steps_tp / t_tp:.3e} chain-steps/s")

    # ---- further workloads, reported beside the line ---------------------------
    extra = {}
    wanted = [c.strip() for c in args.configs.split(",") if c.strip() and c.strip() != "none"]
    for name, fn, it in (("c1", bench_c1, args.iterations), ("c3", bench_c3, args.c3_iterations),
                         ("c4", bench_c4, args.c4_iterations),
                         ("c5", bench_c5, args.c5_iterations),
                         ("c5epg", lambda c, d, it: bench_c5(c, d, it, epg=True), args.c5_iterations // 4),
                         ("tvsmall", bench_tv_small, args.iterations),
                         ("dict", bench_dict, args.steps)):
        if name in wanted:
            extra[name] = fn(ctx, device, it)
            log(f"{name}: {json.dumps({k: v for k, v in extra[name].items() if not k.startswith('_')})}")
            torch.cuda.empty_cache()

    # ---- CPU path -------------------------------------------------------------
    cpu = None
    oracle_rel_err = None
    if not args.skip_cpu:
        nthreads, ncores = omp_threads()
        # about 10 s of host work on 16 threads
        nsample = 16000 * nthreads
        t_cpu = cpu_baseline_askadam(s0, r2, sigma, args.iterations, nthreads, nsample)
        cpu = dict(value=nsample / t_cpu, unit="voxels/s", cores=nthreads, host_cores=ncores, kind="port",
                   sample=f"askadam {args.iterations} iterations on {nsample} voxels of the same workload "
                          f"(oracle FP32 restatement, OpenMP {nthreads} threads = the process's CPU share, "
                          f"OMP_NUM_THREADS, of {ncores} host cores in its affinity mask): {t_cpu:.1f} s")
        log(f"CPU askadam: {t_cpu:.2f} s for {nsample} voxels")
        if mh is not None:
            ns_mh = 4000 * nthreads
            t_cmh = cpu_baseline_mh(s0, r2, sigma, args.mh_iterations, args.mh_chains, nthreads, ns_mh)
            cpu["mh_samples_per_s"] = ns_mh * args.mh_chains * args.mh_iterations / t_cmh
            cpu["mh_sample"] = f"MH {args.mh_chains}x{args.mh_iterations} on {ns_mh} voxels: {t_cmh:.1f} s"
            log(f"CPU MH: {t_cmh:.2f} s for {ns_mh} voxels")
        # checker runs of the FP64 oracle on device results (not timed as a baseline)
        if sub is not None:
            oracle_rel_err = oracle_check(md, cfg, sub, ["M0", "R2star"], nthreads)
            log(f"oracle check (headline subset): {json.dumps(oracle_rel_err)}")
        for name, w in extra.items():
            chk = w.pop("_oracle_check", None)
            if chk is not None:
                w["oracle_rel_err"] = chk(nthreads)
                log(f"oracle check ({name}): {json.dumps(w['oracle_rel_err'])}")
            cb = w.pop("_cpu_baseline", None)
            if cb is not None:
                w["cpu_baseline"] = cb(nthreads, ncores)
                log(f"CPU path ({name}): {json.dumps(w['cpu_baseline'])}")
    for w in extra.values():
        w.pop("_oracle_check", None)
        w.pop("_cpu_baseline", None)

    traffic = measured_traffic() if tuple(args.dims) == (256, 256, 128) and args.iterations == 2000 else None
    out = {
        "metric": "voxels fitted/s (askadam) + MCMC samples/s vs plain GPU path at equal error",
        "value": vox_per_s,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"mono-exponential R2* askadam, {dims[0]}x{dims[1]}x{dims[2]} Bernoulli(0.7) mask "
                        f"({nv} voxels), 12 echoes 2:4:46 ms, {args.iterations} iterations L1/Adam lr 1e-3; "
                        f"MH {args.mh_chains} chains x {args.mh_iterations} iterations",
            "voxels": nv,
            "iterations": args.iterations,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, os.path.dirname(os.path.abspath(__file__)))
            if traffic is not None else None,
            "hbm_frac": (traffic / (k_ms / max(launches, 1) * 1e-3) / (PEAK_HBM_GBS * 1e9))
            if traffic is not None else None,
            "bound_note": "FP32 vector ALU and transcendental issue: no dense contraction on this path (no MFMA) "
                          "and O(1) HBM bytes per voxel per launch; peak = published FP32 157.3 TFLOP/s",
            "kernel": "k_askadam (fused forward+adjoint+Adam, K iterations per launch)",
            "kernel_ms_total": k_ms,
            "launches": launches,
            "avg_launch_ms": k_ms / max(launches, 1),
            "frac_of_nonpacked_peak": achieved_tflops / PEAK_FP32_NONPACKED_TFLOPS,
            "algorithmic_transcendentals_per_s": trans_rate,
            "flop_per_unit": FLOP_PER_VOXEL_ITER,
            "unit_of_work": "voxel-iteration",
        },
        "cpu_baseline": cpu,
        "fused": {"final_loss": fused_loss, "median_rel_err": err_fused, "oracle_rel_err": oracle_rel_err,
                  "iterations_per_launch": int(vol.iters_per_launch), "threads_per_block": int(vol.block_threads)},
        "tape": tape,
        "mcmc": mh,
        "workloads": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
