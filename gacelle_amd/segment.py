"""Segmented estimate(): the volume split along z into voxel-balanced slabs
(utils/utils.m:592-1003) and the outputs reassembled (utils/utils.m:1348-1540).

The reference measures GPU memory by fitting probe volumes under
nvidia-smi and extrapolating MB per voxel.  Here the footprint is
analytic: the device buffers the library and this mirror allocate per
voxel for one fit (`fit_bytes`), against the free device memory (or
fitting['memoryBudget'] in bytes).  A 288 GB device holds every practical
volume in one piece, so the split normally comes from
fitting['NSegmentUser'] (with fitting['segmentOverlap'] halo slices),
exactly as in the reference.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi
from ._abi import GacelleError

FIXED_BYTES = 256 << 20  # context, model tables, pinned staging, allocator slack


def fit_bytes(nmeas, npar, n_extra=0, weighted=False, solver="askadam", n_sample=0, n_chain=1):
    """Device bytes per masked voxel of one fit.  askadam: data, weights,
    extra rows, three (z, m, v) state sets, two step-form ping-pong sets,
    two z_last buffers, final / min maps, residuals and per-voxel losses,
    the voxel index, and the block partials (one float per 256 voxels and
    iteration, up to 256 iterations per launch).  mcmc: data, weights,
    extra rows, chain state and the stored posterior (n_sample retained
    samples of every chain)."""
    f = 4
    b = nmeas * f + (nmeas * f if weighted else 0) + n_extra * f + 8
    if str(solver).lower() == "mcmc":
        return b + n_chain * npar * f * (3 + n_sample) + 2 * npar * f
    partials = 256 * f // 256  # one float per 256-voxel block and iteration, 256 iterations
    return b + (9 + 6 + 2 + 3) * npar * f + 2 * nmeas * f + 2 * f + partials


def free_device_bytes():
    try:
        import torch

        if torch.cuda.is_available():
            return int(torch.cuda.mem_get_info()[0])
    except Exception:
        pass
    return None


def build_balanced_boundaries(mask, nvox_per_seg, nseg_min=1):
    """Owned slice ranges (0-based z indices) with about equal voxel counts
    (utils/utils.m:868-943)."""
    density = np.asarray(mask, bool).reshape(mask.shape[0] * mask.shape[1], -1, order="F").sum(axis=0)
    nz = density.size
    total = int(density.sum())
    nseg = max(int(math.ceil(total / max(nvox_per_seg, 1))), int(nseg_min))
    target = total / nseg
    cum = np.cumsum(density)
    out, start = [], 0
    for ks in range(1, nseg + 1):
        if ks < nseg:
            t = ks * target
            below = np.flatnonzero(cum < t)
            above = np.flatnonzero(cum >= t)
            if below.size == 0:
                end = start
            elif above.size == 0:
                end = nz - 1
            else:
                b, a = below[-1], above[0]
                vb = cum[b] - (ks - 1) * target
                va = cum[a] - (ks - 1) * target
                end = b if abs(vb - target) <= abs(va - target) else a
            end = max(end, start)
        else:
            end = nz - 1
        if start <= end and start < nz:
            out.append(np.arange(start, min(end, nz - 1) + 1))
        start = end + 1
    return out


def expand_segments_with_halo(owned, h, nz):
    """fit / owned / local slice ranges (utils/utils.m:945-985)."""
    seg = []
    for ks, o in enumerate(owned):
        lo = h if ks > 0 else 0
        hi = h if ks < len(owned) - 1 else 0
        fit = np.arange(max(0, o[0] - lo), min(nz - 1, o[-1] + hi) + 1)
        seg.append(dict(owned=o, fit=fit, local=o - fit[0]))
    return seg


def find_optimal_segment_3D(mask, fitting, bytes_per_voxel):
    """utils/utils.m:592-826 with an analytic footprint: one segment when the
    fit fits the free device memory and fitting['NSegmentUser'] <= 1;
    otherwise voxel-balanced slabs, at least NSegmentUser of them, each
    with fitting['segmentOverlap'] halo slices on its inner sides."""
    mask = np.asarray(mask, bool)
    if mask.ndim < 3:
        mask = mask.reshape(mask.shape + (1,) * (3 - mask.ndim))
    nz = mask.shape[2]
    h = int(fitting.get("segmentOverlap") or 0)
    nseg_user = fitting.get("NSegmentUser")
    nseg_user = 1 if nseg_user is None or np.size(nseg_user) == 0 or nseg_user < 1 else int(round(nseg_user))
    nvox = int(mask.sum())
    avail = fitting.get("memoryBudget")
    avail = free_device_bytes() if avail is None else int(avail)
    need = FIXED_BYTES + bytes_per_voxel * nvox
    whole = [dict(owned=np.arange(nz), fit=np.arange(nz), local=np.arange(nz))]
    if avail is not None and avail < need:
        per_fit = (avail - FIXED_BYTES) // max(bytes_per_voxel, 1)
        if per_fit <= 0:
            raise GacelleError(_abi.GC_ERR_NOMEM, "GACELLE:memoryError: insufficient GPU memory even for a single "
                               "segment")
        dens = mask.reshape(-1, nz, order="F").sum(axis=0)
        mean_slice = float(dens[dens > 0].mean()) if np.any(dens > 0) else 0.0
        owned_per = max(1, int(per_fit - 2 * h * mean_slice))
        nseg_mem = max(1, int(math.ceil(nvox / owned_per)))
        owned = build_balanced_boundaries(mask, owned_per, max(nseg_mem, nseg_user))
        return expand_segments_with_halo(owned, h, nz)
    if nseg_user > 1:
        owned = build_balanced_boundaries(mask, int(math.ceil(nvox / nseg_user)), nseg_user)
        return expand_segments_with_halo(owned, h, nz)
    return whole


def slice_segment(data, mask, fit, extra=None):
    """Data, mask and extraData maps restricted to the slices of a segment
    (the model classes' slice_segment)."""
    data = np.asarray(data)
    nz = np.asarray(mask).shape[2] if np.asarray(mask).ndim >= 3 else 1
    ds = data[:, :, fit]
    ms = np.asarray(mask)[:, :, fit]
    es = None
    if extra is not None:
        es = {}
        for k, v in extra.items():
            if isinstance(v, dict):
                es[k] = {kk: (np.asarray(vv)[:, :, fit] if np.ndim(vv) >= 3 and np.shape(vv)[2] == nz else vv)
                         for kk, vv in v.items()}
            elif v is not None and np.ndim(v) >= 3 and np.shape(v)[2] == nz:
                es[k] = np.asarray(v)[:, :, fit]
            else:
                es[k] = v
    return ds, ms, es


def _crop(val, seg):
    if isinstance(val, (str, bool)) or val is None or np.isscalar(val):
        return val
    a = np.asarray(val)
    if a.ndim >= 3 and a.shape[2] == seg["fit"].size:
        return a[:, :, seg["local"]]
    return val


def crop_segment_output(out, seg):
    """utils/utils.m:1486-1540: maps cropped to the owned slices."""
    res = {}
    for k, v in out.items():
        res[k] = {kk: _crop(vv, seg) for kk, vv in v.items()} if isinstance(v, dict) else _crop(v, seg)
    return res


def _restore(existing, new, owned, nz, is_first):
    if isinstance(new, (str, bool)) or new is None:
        return new if is_first else existing
    if np.isscalar(new) or np.ndim(new) == 0:
        return [new] if is_first else list(existing) + [new]
    a = np.asarray(new)
    if a.ndim >= 3 and a.shape[2] == owned.size:
        if is_first:
            shape = list(a.shape)
            shape[2] = nz
            existing = np.zeros(shape, a.dtype)
        existing[:, :, owned] = a
        return existing
    if a.ndim == 2:
        return a if is_first else np.concatenate([existing, a], axis=1)
    return a if is_first else existing


def restore_segment_structure(out, out_seg, owned, k, nz):
    """utils/utils.m:1348-1482: maps placed at their owned slices of the full
    grid; scalars collected per segment; G-D matrices concatenated along
    voxels."""
    first = k == 0
    out = {} if first else out
    for key, v in out_seg.items():
        if isinstance(v, dict):
            cur = out.setdefault(key, {})
            for kk, vv in v.items():
                cur[kk] = _restore(cur.get(kk), vv, owned, nz, first)
        else:
            out[key] = _restore(out.get(key), v, owned, nz, first)
    return out


def run_segments(fit_one, data, mask, extra, fitting, bytes_per_voxel):
    """estimate()'s loop over segments: fit_one(data, mask, extra) per
    segment, outputs cropped to the owned slices and reassembled."""
    mask = np.asarray(mask)
    m3 = mask.reshape(mask.shape + (1,) * (3 - mask.ndim)) if mask.ndim < 3 else mask
    nz = m3.shape[2]
    segs = find_optimal_segment_3D(m3, fitting, bytes_per_voxel)
    if len(segs) == 1:
        return fit_one(data, mask, extra)
    out = None
    for k, sg in enumerate(segs):
        ds, ms, es = slice_segment(data, m3, sg["fit"], extra)
        o = fit_one(ds, ms, es)
        out = restore_segment_structure(out, crop_segment_output(o, sg), sg["owned"], k, nz)
    return out


def surface_segments(fitting, nhemi):
    """Surface data segment exactly per hemisphere or not at all
    (NEXI/gpuNEXI.m:537-546): fitting['NSegmentUser'] of 1 (or unset) or
    Nhemi; the mesh graphs of the hemispheres are independent, so the split
    needs no halo."""
    n = fitting.get("NSegmentUser")
    n = 1 if n is None or np.size(n) == 0 else int(round(float(np.asarray(n).ravel()[0])))
    if n not in (1, nhemi):
        raise GacelleError(_abi.GC_ERR_ARG, "GACELLE:invalidSurfaceSegmentation: Surface segmentation is exact-only "
                           f"and limited to fitting.NSegmentUser = 1 (no split) or {nhemi} (exact per-hemisphere "
                           f"split); got {n}.")
    if n == 1 or nhemi == 1:
        return None
    return [dict(owned=np.array([h]), fit=np.array([h]), local=np.array([0])) for h in range(nhemi)]


def run_surface_segments(fit_one, data, mask, extra, fitting):
    """estimate() on surface data: one fit, or one per hemisphere with the
    outputs placed back at their hemisphere (dim 3)."""
    mask = np.asarray(mask)
    m3 = mask.reshape(mask.shape + (1,) * (3 - mask.ndim)) if mask.ndim < 3 else mask
    segs = surface_segments(fitting, m3.shape[2])
    if segs is None:
        return fit_one(data, mask, extra)
    out = None
    for k, sg in enumerate(segs):
        ds, ms, es = slice_segment(data, m3, sg["fit"], extra)
        out = restore_segment_structure(out, fit_one(ds, ms, es), sg["owned"], k, m3.shape[2])
    return out
