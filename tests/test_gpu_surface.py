"""GPU parity of the mesh total-variation regulariser of surface data
(gpuNEXI.surface_total_variation) inside askadam, through the C ABI, against
the FP64 oracle; and gpuNEXI.estimate on surface data read from FreeSurfer
triangle surfaces."""
import os

import numpy as np
import pytest

from gacelle_amd import _abi, models, surface
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from . import gpu_abi
from .helpers import TE12, logit_start, r2s_setup
from .test_oracle_surface import torus_grid

pytestmark = pytest.mark.gpu


def _surface_problem(iteration, optimiser, lr, noisy_start, lam=0.002, seed=4):
    """Two hemispheres of a 9 x 6 periodic triangulated grid (54 vertices
    each, degree 6), random edge lengths, 75% of the vertices masked."""
    rng = np.random.default_rng(seed)
    nvert, nhemi = 54, 2
    faces = torus_grid(9, 6)
    lists = surface.mesh_neighbours(faces, nvert)
    width = max(x.size for x in lists) + 1                   # one self-padding slot
    nb = np.zeros((1, nvert, nhemi, width))
    el = np.zeros((1, nvert, nhemi, width))
    for h in range(nhemi):
        for i, x in enumerate(lists):
            nb[0, i, h, :] = i + 1
            nb[0, i, h, : x.size] = x + 1
            el[0, i, h, : x.size] = rng.uniform(0.6, 1.8, x.size)
            el[0, i, h, x.size:] = surface.EDGE_LENGTH_SENTINEL
    mask = rng.random((1, nvert, nhemi)) < 0.75
    idx, ln = surface.surface_tv_tables(mask, nb, el)
    vox = np.flatnonzero(mask.ravel(order="F")).astype(np.int64)
    nv = vox.size
    _, f, md = r2s_setup(dict(iteration=iteration, lossFunction="L2", optimiser=optimiser, initialLearnRate=lr,
                              tol=0, convergenceValue=-1, convergenceStepTol=0, convergenceGradTol=0,
                              regmap=["R2star"], dataType="surface", **{"lambda": [lam]}))
    cfg = askadam().make_cfg(f, models.gpuR2starMapping(TE12))
    assert cfg.reg_type == _abi.GC_REG_SURFACE_TV
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = (orc.fwd_value(md, th) + 0.01 * rng.standard_normal((12, nv))).astype(np.float32)
    if noisy_start:
        z0 = logit_start(np.stack([rng.uniform(0.8, 1.2, nv), rng.uniform(20, 40, nv)]), md)
    else:
        z0 = logit_start(np.stack([np.ones(nv), np.full(nv, 30.0)]), md)
    dims = (1, nvert, nhemi)
    return md, cfg, y, z0, vox, dims, idx, ln


def _run_gpu(md, cfg, y, z0, vox, dims, idx, ln, trace=True):
    return gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, vox=vox, dims=dims, surf_nbr=idx, surf_len=ln)],
                           trace=trace)[0]


def test_surface_tv_step_matches_oracle():
    """One SGDM step (z1 = z0 - lr g) from a start with no tied neighbours:
    the mesh TV value and its adjoint (own and reverse terms) match FP64."""
    md, cfg, y, z0, vox, dims, idx, ln = _surface_problem(2, "sgdm", 50.0, True)
    got = _run_gpu(md, cfg, y, z0, vox, dims, idx, ln)
    ref = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, trace=True, surf_nbr=idx, surf_len=ln)
    np.testing.assert_allclose(got["trace"], ref["trace"], rtol=1e-5)
    assert got["loss_reg"] == pytest.approx(ref["loss_reg"], rel=1e-5)
    assert ref["loss_reg"] > 0.1 * ref["loss_fidelity"]          # the regulariser is not negligible
    dz_g = got["z_final"].astype(np.float64) - z0
    dz_o = ref["z_final"] - z0
    for p in range(2):
        err = np.abs(dz_g[p] - dz_o[p]) / np.abs(dz_o[p]).max()
        assert err.max() < 1e-4, (p, err.max())


def test_surface_tv_trajectory_within_fp32_spread():
    """150 Adam iterations: within 3x of the oracle's own FP32 restatement's
    distance from FP64 (the TV subgradient flips as neighbours cross)."""
    md, cfg, y, z0, vox, dims, idx, ln = _surface_problem(150, "adam", 0.01, False)
    got = _run_gpu(md, cfg, y, z0, vox, dims, idx, ln)
    ref = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, trace=True, surf_nbr=idx, surf_len=ln)
    f32 = orc.askadam(md, cfg, y, z0, vox=vox, dims=dims, trace=True, surf_nbr=idx, surf_len=ln, precision=1)
    assert got["n_iteration"] == ref["n_iteration"] == 150
    spread = np.abs(f32["trace"] - ref["trace"]).max() / ref["trace"].max()
    assert np.abs(got["trace"] - ref["trace"]).max() / ref["trace"].max() <= 3 * spread + 1e-6


def test_surface_tv_needs_its_tables():
    md, cfg, y, z0, vox, dims, idx, ln = _surface_problem(2, "sgdm", 1.0, True)
    with pytest.raises(_abi.GacelleError, match="surface TV needs"):
        gpu_abi.askadam(md, cfg, [dict(y=y, z0=z0, vox=vox, dims=dims)])
    bad = idx.copy()
    bad[0, 0] = dims[1] * dims[2]
    with pytest.raises(_abi.GacelleError, match="outside the grid"):
        _run_gpu(md, cfg, y, z0, vox, dims, bad, ln)


def _write_mesh(d, nx, ny):
    """lh/rh white and pial surfaces of a triangulated nx x ny grid on a
    cylinder (pial 2 mm outward)."""
    faces = torus_grid(nx, ny)
    ang = 2 * np.pi * (np.arange(nx * ny) % nx) / nx
    zz = (np.arange(nx * ny) // nx) * 1.0
    for h, r in (("lh", 20.0), ("rh", 25.0)):
        for name, rr in (("white", r), ("pial", r + 2.0)):
            v = np.stack([rr * np.cos(ang), rr * np.sin(ang), zz], axis=1)
            surface.write_surf(os.path.join(d, f"{h}.{name}"), v, faces)
    return nx * ny


def test_nexi_estimate_on_surface_data(tmp_path):
    """gpuNEXI.estimate with dataType 'surface': the mesh is read from
    FreeSurfer surfaces, the fit with mesh TV on fa runs through the
    library, the regulariser lowers the roughness of fa along the mesh, and
    the exact per-hemisphere split (NSegmentUser = 2) lands each
    hemisphere at its place with fits close to the joint one."""
    nvert = _write_mesh(str(tmp_path), 12, 8)
    b = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    D = np.full(6, 20.0)
    m = models.gpuNEXI(b, D)
    m0 = models.gpuNEXI(b, D)
    m0.updateProperty(m0.check_set_default(dict(lmax=0)))
    md = models.gpuNEXI(b, D).gc_model(dict(modelParams=m0.modelParams, lb=m0.lb, ub=m0.ub), "askadam", None)
    rng = np.random.default_rng(8)
    fa = rng.uniform(0.3, 0.7, (nvert, 2))
    th = np.stack([fa, rng.uniform(1.8, 2.8, (nvert, 2)), rng.uniform(0.6, 1.2, (nvert, 2)),
                   (1 - fa) / rng.uniform(5, 40, (nvert, 2))])
    s = orc.fwd_value(md, th.reshape(4, -1, order="F").astype(np.float64))      # [M, nvert*2]
    y = (s + 0.02 * rng.standard_normal(s.shape)).astype(np.float32)
    data = y.T.reshape(1, nvert, 2, -1, order="F")
    mask = np.ones((1, nvert, 2), bool)
    ex = dict(surf_dir=str(tmp_path), hemisphere=["lh", "rh"])
    rough = {}
    outs = {}
    for lam in (0.0, 0.05):
        fit = dict(iteration=200, lossFunction="L2", verbose=False, tol=0, convergenceValue=-1,
                   convergenceStepTol=0, convergenceGradTol=0, dataType="surface", regmap=["fa"],
                   initialLearnRate=0.01, **{"lambda": [lam]})
        out = m.estimate(data, mask, ex, fit)
        fa = np.asarray(out["final"]["fa"]).reshape(nvert, 2, order="F")
        nb, el = m.get_surface_geometry(str(tmp_path), 0.5, ["lh", "rh"])
        r = 0.0
        for h in range(2):
            for k in range(nb.shape[3]):
                j = nb[0, :, h, k].astype(int) - 1
                r += np.abs(fa[:, h] - fa[j, h]).sum()
        rough[lam] = r
        outs[lam] = fa
    assert rough[0.05] < 0.8 * rough[0.0]
    fit = dict(iteration=200, lossFunction="L2", verbose=False, tol=0, convergenceValue=-1, convergenceStepTol=0,
               convergenceGradTol=0, dataType="surface", regmap=["fa"], NSegmentUser=2, initialLearnRate=0.01,
               **{"lambda": [0.05]})
    split = np.asarray(m.estimate(data, mask, ex, fit)["final"]["fa"]).reshape(nvert, 2, order="F")
    assert np.median(np.abs(split - outs[0.05]) / outs[0.05]) < 0.02
    with pytest.raises(_abi.GacelleError, match="invalidSurfaceSegmentation"):
        m.estimate(data, mask, ex, dict(fit, NSegmentUser=3))
    with pytest.raises(_abi.GacelleError, match="missingSurfDir"):
        m.estimate(data, mask, {"hemisphere": ["lh", "rh"]}, dict(fit, NSegmentUser=1))
