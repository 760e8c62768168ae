"""Pack ctypes descriptor structs (gc_model, gc_askadam_cfg, gc_mcmc_cfg) into
flat npz entries and back, so that golden fixtures carry the exact
description they were computed with."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from gacelle_amd import _abi

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pack(prefix: str, st) -> dict:
    """Numeric fields only: the network descriptors (gc_model.mlp) carry
    device pointers and are not part of a fixture."""
    out = {}
    for name, ty in st._fields_:
        v = getattr(st, name)
        if isinstance(v, ctypes.Array) and issubclass(v._type_, ctypes.Structure):
            continue
        if isinstance(v, ctypes.Array):
            v = np.array(list(v))
        out[f"{prefix}.{name}"] = np.asarray(v)
    return out


def unpack(prefix: str, cls, data) -> object:
    st = cls()
    for name, ty in cls._fields_:
        key = f"{prefix}.{name}"
        if key not in data:
            continue
        v = data[key]
        cur = getattr(st, name)
        if isinstance(cur, ctypes.Array):
            for i, x in enumerate(np.asarray(v).ravel()):
                cur[i] = x.item()
        else:
            setattr(st, name, np.asarray(v).item())
    return st


def load(name: str) -> dict:
    with np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def model(data) -> _abi.gc_model:
    return unpack("model", _abi.gc_model, data)


def askadam_cfg(data) -> _abi.gc_askadam_cfg:
    return unpack("cfg", _abi.gc_askadam_cfg, data)


def mcmc_cfg(data) -> _abi.gc_mcmc_cfg:
    return unpack("cfg", _abi.gc_mcmc_cfg, data)


def opt(data, key):
    """An optional array entry (None when the fixture has no such input)."""
    return data[key] if key in data else None
