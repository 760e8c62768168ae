"""Surface (cortical mesh) data for the mesh total-variation regulariser.

- FreeSurfer triangle surfaces (``lh.white``, ``lh.pial``, ...): the binary
  format that FreeSurfer's ``read_surf`` reads (a 3-byte magic number
  0xFFFFFE, a creator line ended by two newlines, then big-endian int32
  vertex and face counts, float32 xyz per vertex and int32 vertex triples
  per face).
- ``get_surface_neighbours``: utils/utils.m:1615-1830, the per-vertex mesh
  adjacency on the surface at a cortical depth between white and pial, its
  edge vectors and edge lengths in the [1, Nvertex, Nhemi, maxNeighbours]
  layout.
- ``surface_tv_tables``: the neighbours and edge lengths as the C ABI takes
  them (gc_volume.surf_nbr_dev / surf_len_dev): global 0-based vertex
  indices, [maxNeighbours, Nvertex*Nhemi], the conversion that
  surface_total_variation makes with its hemisphere offset
  (NEXI/gpuNEXI.m:1229-1233).
"""
from __future__ import annotations

import os
import warnings

import numpy as np

from . import _abi
from ._abi import GacelleError

_TRIANGLE_MAGIC = 0xFFFFFE
EDGE_LENGTH_SENTINEL = 1e6  # utils/utils.m:1826


def read_surf(path: str) -> tuple[np.ndarray, np.ndarray]:
    """FreeSurfer triangle surface -> (vertices [N, 3] float64, faces [F, 3]
    int64, 0-based as in the file)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    if len(raw) < 3 or int.from_bytes(raw[:3], "big") != _TRIANGLE_MAGIC:
        raise GacelleError(_abi.GC_ERR_ARG, f"{path}: not a FreeSurfer triangle surface")
    end = raw.find(b"\n\n", 3)
    if end < 0:
        raise GacelleError(_abi.GC_ERR_ARG, f"{path}: truncated header")
    pos = end + 2
    nv, nf = (int(x) for x in np.frombuffer(raw, ">i4", 2, pos))
    pos += 8
    need = pos + 12 * nv + 12 * nf
    if nv < 0 or nf < 0 or len(raw) < need:
        raise GacelleError(_abi.GC_ERR_ARG, f"{path}: truncated surface ({nv} vertices, {nf} faces)")
    verts = np.frombuffer(raw, ">f4", 3 * nv, pos).reshape(nv, 3).astype(np.float64)
    pos += 12 * nv
    faces = np.frombuffer(raw, ">i4", 3 * nf, pos).reshape(nf, 3).astype(np.int64)
    if nf and (faces.min() < 0 or faces.max() >= nv):
        raise GacelleError(_abi.GC_ERR_ARG, f"{path}: face index outside the vertex list")
    return verts, faces


def write_surf(path: str, vertices, faces, created_by: str = "gacelle_amd") -> None:
    """Write a FreeSurfer triangle surface (faces 0-based)."""
    v = np.asarray(vertices, np.float64).reshape(-1, 3)
    f = np.asarray(faces, np.int64).reshape(-1, 3)
    with open(path, "wb") as fh:
        fh.write(_TRIANGLE_MAGIC.to_bytes(3, "big"))
        fh.write(f"created by {created_by}\n\n".encode())
        fh.write(np.array([v.shape[0], f.shape[0]], ">i4").tobytes())
        fh.write(v.astype(">f4").tobytes())
        fh.write(f.astype(">i4").tobytes())


def mesh_neighbours(faces: np.ndarray, nvert: int) -> list[np.ndarray]:
    """Sorted, unique 0-based neighbours of every vertex: the vertices that
    share a face with it (utils/utils.m:1709-1716, where each face adds its
    other two vertices to a vertex's list through unique())."""
    f = np.asarray(faces, np.int64)
    if f.size == 0:
        return [np.zeros(0, np.int64) for _ in range(nvert)]
    a = np.concatenate([f[:, 0], f[:, 0], f[:, 1], f[:, 1], f[:, 2], f[:, 2]])
    b = np.concatenate([f[:, 1], f[:, 2], f[:, 2], f[:, 0], f[:, 0], f[:, 1]])
    pairs = np.unique(np.stack([a, b], axis=1), axis=0)       # sorted by (vertex, neighbour)
    starts = np.searchsorted(pairs[:, 0], np.arange(nvert + 1))
    return [pairs[starts[i]:starts[i + 1], 1] for i in range(nvert)]


def get_surface_neighbours(surf_dir: str, depth: float = 0.5, hemisphere=("lh", "rh")):
    """utils/utils.m:1615-1830: (neighbours, dr, edgeLength).

    neighbours [1, Nvertex, Nhemi, maxNb] holds 1-based per-hemisphere
    vertex indices (float64, as MATLAB), padded with the vertex's own index;
    dr [1, Nvertex, Nhemi, maxNb, 3] the edge vectors on the surface at
    ``depth`` between white (0) and pial (1); edgeLength = |dr|, with zero
    lengths (the padding, and any degenerate edge) set to 1e6."""
    if depth is None:
        depth = 0.5
    if hemisphere is None or len(hemisphere) == 0:
        hemisphere = ("lh", "rh")
    if isinstance(hemisphere, str):
        hemisphere = (hemisphere,)
    hemisphere = list(hemisphere)
    if any(h not in ("lh", "rh") for h in hemisphere):
        raise GacelleError(_abi.GC_ERR_ARG, "GACELLE:invalidHemisphere: hemisphere must be a non-empty list from "
                                            "{'lh','rh'}, got {%s}" % ", ".join(repr(h) for h in hemisphere))
    if not os.path.isdir(surf_dir):
        raise GacelleError(_abi.GC_ERR_ARG, f"GACELLE:surfDirNotFound: surf_dir does not exist: {surf_dir}")
    per_nb, per_dr, per_pad, counts, widths = [], [], [], [], []
    for h in hemisphere:
        vw, faces = read_surf(os.path.join(surf_dir, f"{h}.white"))
        vp, _ = read_surf(os.path.join(surf_dir, f"{h}.pial"))
        if vp.shape != vw.shape:
            raise GacelleError(_abi.GC_ERR_ARG, f"{h}.white and {h}.pial have different vertex counts")
        vmid = vw + depth * (vp - vw)
        n = vw.shape[0]
        lists = mesh_neighbours(faces, n)
        width = max((x.size for x in lists), default=0)
        nb = np.tile(np.arange(n)[:, None], (1, width))      # self-index padding (0-based here)
        pad = np.ones((n, width), bool)
        for i, x in enumerate(lists):
            nb[i, : x.size] = x
            pad[i, : x.size] = False
        per_nb.append(nb)
        per_dr.append(vmid[nb] - vmid[:, None, :])            # [n, width, 3]
        per_pad.append(pad)
        counts.append(n)
        widths.append(width)
    if len(set(counts)) > 1:
        raise GacelleError(_abi.GC_ERR_ARG, "GACELLE:hemisphereVertexMismatch: requested hemispheres have "
                                            f"different vertex counts ({counts}) and cannot be concatenated")
    n = counts[0]
    width = max(widths)
    nh = len(hemisphere)
    neighbours = np.empty((1, n, nh, width))
    dr = np.zeros((1, n, nh, width, 3))
    is_pad = np.ones((1, n, nh, width), bool)
    for k in range(nh):
        w = widths[k]
        nbk = np.tile(np.arange(n)[:, None], (1, width))        # cross-hemisphere padding: self index
        nbk[:, :w] = per_nb[k]
        neighbours[0, :, k, :] = nbk + 1                        # 1-based as MATLAB
        dr[0, :, k, :w, :] = per_dr[k]
        is_pad[0, :, k, :w] = per_pad[k]
    edge = np.sqrt((dr ** 2).sum(axis=4))
    degenerate = (edge == 0) & ~is_pad
    if degenerate.any():
        warnings.warn(f"GACELLE:degenerateSurfaceEdge: {int(degenerate.sum())} non-padding mesh edge(s) have zero "
                      "length (two distinct connected vertices at coincident positions)")
    edge[edge == 0] = EDGE_LENGTH_SENTINEL
    return neighbours, dr, edge


def surface_tv_tables(mask, neighbours, edge_length) -> tuple[np.ndarray, np.ndarray]:
    """Grid-form tables for gc_volume.surf_nbr_dev / surf_len_dev: global
    0-based vertex indices (per-hemisphere 1-based index - 1 + hemisphere *
    Nvertex, NEXI/gpuNEXI.m:1229-1233) and edge lengths, each
    [maxNb, Nvertex*Nhemi] with the vertex grid index (column-major
    [1, Nvertex, Nhemi]) fastest."""
    mask = np.asarray(mask)
    if mask.ndim < 2 or mask.shape[0] != 1:
        raise GacelleError(_abi.GC_ERR_ARG, "GACELLE:surfaceShapeMismatch: surface mode expects a [1, Nvertex, "
                                            f"Nhemi] mask, got {list(mask.shape)}")
    nvert = mask.shape[1]
    nhemi = mask.shape[2] if mask.ndim > 2 else 1
    nb = np.asarray(neighbours, np.float64)
    el = np.asarray(edge_length, np.float64)
    if nb.ndim == 3:
        nb = nb[..., None]
        el = el[..., None]
    if nb.shape[:3] != (1, nvert, nhemi) or el.shape != nb.shape:
        raise GacelleError(_abi.GC_ERR_INPUT_MISMATCH, "GACELLE:surfaceGeometryMismatch: neighbours / edge lengths "
                           f"{list(nb.shape)} / {list(el.shape)} do not match the mask [1, {nvert}, {nhemi}]")
    width = nb.shape[3]
    local = nb[0] - 1                                             # [nvert, nhemi, width], 0-based
    if np.any(local != np.round(local)) or local.min(initial=0) < 0 or local.max(initial=0) >= nvert:
        raise GacelleError(_abi.GC_ERR_ARG, "surface neighbour indices must be integers in 1..Nvertex")
    glob = local + (np.arange(nhemi) * nvert)[None, :, None]
    idx = np.ascontiguousarray(glob.reshape(nvert * nhemi, width, order="F").T.astype(np.int32))
    lens = np.ascontiguousarray(el[0].reshape(nvert * nhemi, width, order="F").T.astype(np.float32))
    return idx, lens
