"""Layout helpers with MATLAB (column-major) index semantics.

Host-side counterparts of utils/utils.m:85-378.  Arrays are numpy arrays
indexed like the MATLAB arrays they stand for (first index fastest in
memory order 'F'); voxel order is that of find(mask).
"""
from __future__ import annotations

import numpy as np


def dims3(a) -> tuple[int, int, int]:
    """size(a, 1:3) (trailing singleton dimensions added)."""
    s = tuple(np.shape(a)) + (1, 1, 1)
    return int(s[0]), int(s[1]), int(s[2])


def find_mask(mask) -> np.ndarray:
    """find(mask > 0) as 0-based int64 column-major linear indices."""
    m = np.asarray(mask)
    return np.flatnonzero(m.ravel(order="F") > 0).astype(np.int64)


def vectorise_NDto2D(data, mask=None) -> tuple[np.ndarray, np.ndarray]:
    """utils/utils.m:215-238: [x,y,z,meas...] -> [Nvox_masked, Nmeas]; complex
    data are split as [real, imag] along the measurement axis."""
    data = np.asarray(data)
    d = dims3(data)
    nvox = d[0] * d[1] * d[2]
    nmeas = int(np.prod(data.shape[3:])) if data.ndim > 3 else 1
    flat = np.reshape(data, (nvox, nmeas), order="F")
    if mask is None or np.size(mask) == 0:
        idx = np.arange(nvox, dtype=np.int64)
    elif np.size(mask) == nvox:
        idx = find_mask(mask)
    else:
        idx = np.asarray(mask, dtype=np.int64)
    out = flat[idx, :]
    if np.iscomplexobj(out):
        out = np.concatenate([out.real, out.imag], axis=1)
    return out, idx


def reshape_ND2GD(data, mask=None) -> np.ndarray:
    """utils/utils.m:157-163: N-D image -> G-D [Nmeas, Nvox]."""
    out, _ = vectorise_NDto2D(data, mask)
    return np.ascontiguousarray(out.T)


def reshape_ND2image(dist, mask) -> np.ndarray:
    """utils/utils.m:299-315: [Nvox_masked, extra...] -> [x,y,z,extra...] with
    zeros outside the mask."""
    dist = np.asarray(dist)
    if dist.size == 1:
        return dist
    d = dims3(mask)
    extra = dist.shape[1:]
    idx = find_mask(mask)
    img = np.zeros((d[0] * d[1] * d[2],) + extra, dtype=dist.dtype)
    img[idx, ...] = dist
    return np.reshape(img, d + extra, order="F")


def reshape_GD2ND(data, mask) -> np.ndarray:
    """utils/utils.m:194-198: G-D [Nmeas, Nvox] -> image [x,y,z,Nmeas]."""
    return reshape_ND2image(np.asarray(data).T, mask)


def masking_ND2GD_preserve(data, mask) -> np.ndarray:
    """utils/utils.m:89-109: [x,y,z,a,b,...] -> [1, Nvox_masked, 1, a, b, ...]."""
    data = np.asarray(data)
    d = dims3(data)
    rest = data.shape[3:]
    idx = find_mask(mask) if np.size(mask) == d[0] * d[1] * d[2] else np.asarray(mask, dtype=np.int64)
    flat = np.reshape(data, (d[0] * d[1] * d[2],) + rest, order="F")
    out = flat[idx, ...]
    return np.reshape(out, (1, idx.size, 1) + rest, order="F")


def undo_masking_ND2GD_preserve(data_masked, mask) -> np.ndarray:
    """utils/utils.m:125-142."""
    data_masked = np.asarray(data_masked)
    d = dims3(mask)
    rest = data_masked.shape[3:]
    idx = find_mask(mask)
    out = np.zeros((d[0] * d[1] * d[2],) + rest, dtype=data_masked.dtype)
    out[idx, ...] = np.reshape(data_masked, (idx.size,) + rest, order="F")
    return np.reshape(out, d + rest, order="F")


def set_nan_inf_zero(data) -> tuple[np.ndarray, np.ndarray]:
    """utils/utils.m:347-350."""
    data = np.array(data, copy=True)
    bad = ~np.isfinite(data)
    data[bad] = 0
    return data, bad


def remove_img_naninf(data, mask) -> tuple[np.ndarray, np.ndarray]:
    """utils/utils.m:353-378: zero NaN/Inf and drop affected voxels from the mask."""
    data, bad = set_nan_inf_zero(data)
    good = ~bad
    if good.ndim > 3:
        good = good.reshape(good.shape[:3] + (-1,), order="F").all(axis=3)
    mask = np.logical_and(np.asarray(mask) > 0, good)
    return data, mask


def initialise_x0(dims, model_params, start_point) -> dict:
    """utils.initialise_x0: constant starting maps."""
    return {p: np.full(dims, float(v), dtype=np.float32) for p, v in zip(model_params, start_point)}


def prctile(x, p) -> float:
    """MATLAB prctile (Statistics Toolbox default method: midpoint
    interpolation of the sorted sample at positions ((1:n)-0.5)/n)."""
    x = np.sort(np.asarray(x, dtype=np.float64).ravel())
    n = x.size
    if n == 0:
        return float("nan")
    pos = np.array([(k + 0.5) / n * 100.0 for k in range(n)])
    return float(np.interp(p, pos, x, left=x[0], right=x[-1]))
