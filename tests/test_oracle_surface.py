"""CPU tests of the surface pieces: the FreeSurfer triangle-surface reader,
the mesh adjacency of utils.get_surface_neighbours, the grid-form tables of
the C ABI, and the oracle's restatement of gpuNEXI.surface_total_variation
(pinned against the volumetric TV restatement, central finite differences
and closed forms)."""
import os

import numpy as np
import pytest

from gacelle_amd import _abi, surface
from gacelle_amd.askadam import askadam
from oracle import oracle as orc

from .helpers import logit_start, r2s_setup


def octahedron(scale=1.0):
    v = scale * np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    return v, f


def torus_grid(nx, ny):
    """Triangulated periodic nx x ny grid whose edges include the 4-neighbour
    edges of the grid (plus one diagonal per cell)."""
    idx = lambda x, y: (x % nx) + nx * (y % ny)
    faces = []
    for y in range(ny):
        for x in range(nx):
            faces.append([idx(x, y), idx(x + 1, y), idx(x + 1, y + 1)])
            faces.append([idx(x, y), idx(x + 1, y + 1), idx(x, y + 1)])
    return np.array(faces)


def test_read_surf_round_trip_and_header(tmp_path):
    v, f = octahedron(2.5)
    p = os.path.join(tmp_path, "lh.white")
    surface.write_surf(p, v, f)
    raw = open(p, "rb").read()
    assert raw[:3] == b"\xff\xff\xfe"
    v2, f2 = surface.read_surf(p)
    np.testing.assert_array_equal(v2, v.astype(np.float32))
    np.testing.assert_array_equal(f2, f)
    open(p, "wb").write(raw[:-4])
    with pytest.raises(_abi.GacelleError, match="truncated"):
        surface.read_surf(p)


def test_mesh_neighbours_are_sorted_face_sharing_vertices():
    v, f = octahedron()
    nb = surface.mesh_neighbours(f, 6)
    # each octahedron vertex touches every vertex but its antipode
    for i in range(6):
        anti = i ^ 1
        assert list(nb[i]) == [j for j in range(6) if j not in (i, anti)]
    # a vertex in no face has an empty list
    nb7 = surface.mesh_neighbours(f, 7)
    assert nb7[6].size == 0


def _write_hemi(d, h, v, f, pial_scale):
    surface.write_surf(os.path.join(d, f"{h}.white"), v, f)
    surface.write_surf(os.path.join(d, f"{h}.pial"), pial_scale * v, f)


def test_get_surface_neighbours_layout_depth_and_padding(tmp_path):
    """Edge lengths at depth 0.5 between white (unit octahedron) and pial
    (scaled 3x) are 2*sqrt(2); the ragged hemisphere (a pyramid whose apex
    has 4 neighbours and base corners 3) is padded with the vertex's own
    index and the 1e6 sentinel, the wider hemisphere sets maxNb."""
    v, f = octahedron()
    _write_hemi(tmp_path, "lh", v, f, 3.0)
    # rh: square pyramid on 6 vertices (vertex 5 isolated) -> max degree 4
    pv = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, 1], [5, 5, 5]], float)
    pf = np.array([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]])
    _write_hemi(tmp_path, "rh", pv, pf, 1.0)
    nb, dr, el = surface.get_surface_neighbours(str(tmp_path), 0.5, ["lh", "rh"])
    assert nb.shape == (1, 6, 2, 4) and dr.shape == (1, 6, 2, 4, 3) and el.shape == (1, 6, 2, 4)
    np.testing.assert_allclose(el[0, :, 0, :], 2 * np.sqrt(2), rtol=1e-6)
    np.testing.assert_array_equal(nb[0, 0, 0], [3, 4, 5, 6])          # 1-based, sorted
    # pyramid base corner 0: neighbours 1, 3, 4 (0-based) then self padding
    np.testing.assert_array_equal(nb[0, 0, 1], [2, 4, 5, 1])
    assert el[0, 0, 1, 3] == surface.EDGE_LENGTH_SENTINEL
    np.testing.assert_allclose(el[0, 0, 1, :3], [np.sqrt(2), np.sqrt(2), np.sqrt(2)], rtol=1e-6)
    # isolated vertex: all padding
    np.testing.assert_array_equal(nb[0, 5, 1], [6, 6, 6, 6])
    assert np.all(el[0, 5, 1] == surface.EDGE_LENGTH_SENTINEL)
    with pytest.raises(_abi.GacelleError, match="invalidHemisphere"):
        surface.get_surface_neighbours(str(tmp_path), 0.5, ["xx"])
    with pytest.raises(_abi.GacelleError, match="surfDirNotFound"):
        surface.get_surface_neighbours(str(tmp_path / "nope"), 0.5, ["lh"])


def test_hemisphere_vertex_counts_must_match(tmp_path):
    v, f = octahedron()
    _write_hemi(tmp_path, "lh", v, f, 2.0)
    _write_hemi(tmp_path, "rh", np.vstack([v, [[9, 9, 9]]]), f, 2.0)
    with pytest.raises(_abi.GacelleError, match="hemisphereVertexMismatch"):
        surface.get_surface_neighbours(str(tmp_path), 0.5, ["lh", "rh"])


def test_tables_add_the_hemisphere_offset():
    nvert = 6
    nb = np.zeros((1, nvert, 2, 2))
    nb[0, :, :, 0] = (np.arange(nvert) + 1)[:, None]
    nb[0, :, :, 1] = ((np.arange(nvert) + 1) % nvert + 1)[:, None]
    el = np.ones_like(nb)
    idx, ln = surface.surface_tv_tables(np.ones((1, nvert, 2), bool), nb, el)
    assert idx.shape == (2, 2 * nvert) and ln.shape == (2, 2 * nvert)
    np.testing.assert_array_equal(idx[0], np.arange(2 * nvert))          # self
    np.testing.assert_array_equal(idx[1, :nvert], (np.arange(nvert) + 1) % nvert)
    np.testing.assert_array_equal(idx[1, nvert:], nvert + (np.arange(nvert) + 1) % nvert)
    with pytest.raises(_abi.GacelleError, match="surfaceShapeMismatch"):
        surface.surface_tv_tables(np.ones((2, nvert, 2), bool), nb, el)


def _grid_tables(nbl, ngrid):
    width = max(len(x) for x in nbl)
    idx = np.tile(np.arange(ngrid)[None, :], (width, 1)).astype(np.int32)
    for i, x in enumerate(nbl):
        idx[: len(x), i] = x
    return idx


def test_surface_tv_on_a_torus_grid_is_twice_the_volumetric_2d_tv():
    """On the full-mask periodic grid with the 4 grid neighbours and unit
    edges every edge is counted from both ends: twice the volumetric
    forward-difference TV ('2D', unit voxels), value and gradient."""
    nx, ny = 7, 5
    rng = np.random.default_rng(1)
    img = rng.standard_normal((nx, ny, 1))
    mask = np.ones((nx, ny, 1), bool)
    n = nx * ny
    nbl = []
    for y in range(ny):
        for x in range(nx):
            nbl.append([((x + 1) % nx) + nx * y, ((x - 1) % nx) + nx * y, x + nx * ((y + 1) % ny),
                        x + nx * ((y - 1) % ny)])
    idx = _grid_tables(nbl, n)
    el = np.ones(idx.shape, np.float32)
    lam = 0.3
    ls, gs = orc.surface_tv(img.ravel(order="F"), mask.ravel(order="F"), idx, el, lam, n)
    lv, gv = orc.tv(img, mask, _abi.GC_TV_2D, [1, 1, 1], lam, n)
    assert ls == pytest.approx(2 * lv, rel=1e-12)
    np.testing.assert_allclose(gs, 2 * gv.ravel(order="F"), rtol=1e-12, atol=1e-15)


def test_surface_tv_gradient_matches_central_differences_with_a_partial_mask():
    rng = np.random.default_rng(2)
    n = 40
    f = torus_grid(8, 5)
    nbl = [list(x) for x in surface.mesh_neighbours(f, n)]
    idx = _grid_tables(nbl, n)
    el = rng.uniform(0.5, 2.0, idx.shape).astype(np.float32)
    mask = rng.random(n) < 0.7
    img = np.where(mask, rng.standard_normal(n), 0.0)
    nm = int(mask.sum())
    _, g = orc.surface_tv(img, mask, idx, el, 0.7, nm)
    h = 1e-6
    for i in np.flatnonzero(mask):
        a, b = img.copy(), img.copy()
        a[i] += h
        b[i] -= h
        fd = (orc.surface_tv(a, mask, idx, el, 0.7, nm, grad=False) -
              orc.surface_tv(b, mask, idx, el, 0.7, nm, grad=False)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-9), i


def test_surface_tv_closed_form_and_zeroed_entries():
    """Path 0-1-2 with vertex 2 outside the mask (theta = 0 there):
    vertex 0 lists 1 (length 2) and itself (padding, length 0 -> 0/0
    zeroed); vertex 1 lists 0 (length 2) and 2 (length 4)."""
    img = np.array([1.0, 3.0, 0.0])
    mask = np.array([1, 1, 0], np.uint8)
    idx = np.array([[1, 0, 1], [0, 2, 2]], np.int32)
    el = np.array([[2.0, 2.0, 1.0], [0.0, 4.0, 0.0]], np.float32)
    lam = 0.5
    loss, g = orc.surface_tv(img, mask, idx, el, lam, 2)
    # |1-3|/2 + |3-1|/2 + |3-0|/4
    assert loss == pytest.approx(lam / 2 * (1 + 1 + 0.75))
    assert g[0] == pytest.approx(lam / 2 * (-0.5 - 0.5))
    assert g[1] == pytest.approx(lam / 2 * (0.5 + 0.5 + 0.25))


def test_oracle_askadam_surface_regulariser_value_at_the_start():
    """loss_reg of the oracle's askadam with GC_REG_SURFACE_TV is the
    surface TV of the starting parameter map (theta zero outside the mask)."""
    rng = np.random.default_rng(3)
    nvert = 30
    faces = torus_grid(6, 5)
    nbl = [list(x) for x in surface.mesh_neighbours(faces, nvert)]
    idx = _grid_tables(nbl, nvert)
    el = rng.uniform(0.8, 1.5, idx.shape).astype(np.float32)
    mask = rng.random(nvert) < 0.8
    vox = np.flatnonzero(mask)
    nv = vox.size
    _, f, md = r2s_setup(dict(iteration=1, lossFunction="L2", tol=0, convergenceValue=-1, convergenceStepTol=0,
                              convergenceGradTol=0, regmap=["R2star"], dataType="surface", **{"lambda": [0.01]}))
    cfg = askadam().make_cfg(f, None)
    assert cfg.reg_type == _abi.GC_REG_SURFACE_TV
    th = np.stack([rng.uniform(0.5, 1.5, nv), rng.uniform(10, 60, nv)])
    y = orc.fwd_value(md, th).astype(np.float32)
    z0 = logit_start(th, md)
    out = orc.askadam(md, cfg, y, z0, vox=vox, dims=(1, nvert, 1), trace=True, surf_nbr=idx, surf_len=el)
    lb, ub = md.lb[1], md.ub[1]
    t = lb + (ub - lb) / (1 + np.exp(-z0[1].astype(np.float64)))
    img = np.zeros(nvert)
    img[vox] = t
    ref = orc.surface_tv(img, mask, idx, el, 0.01, nv, grad=False)
    # one epoch: evaluated at the start, no update on the last epoch
    assert out["n_iteration"] == 1
    assert out["loss_reg"] == pytest.approx(ref, rel=1e-9)
    assert out["trace"][0] == pytest.approx(out["loss_fidelity"] + ref, rel=1e-9)
