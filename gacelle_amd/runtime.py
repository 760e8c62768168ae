"""Device context and buffer plumbing for the HIP library.

PyTorch is used only to own device memory and the stream; every computation
of the estimation path runs in libgacelle_hip.so.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _abi
from ._abi import GacelleError

_lock = threading.Lock()
_contexts: dict[int, "Context"] = {}


def _torch():
    import torch  # imported lazily: the CPU-side helpers do not need it

    return torch


class Context:
    """One gc_ctx on one HIP device, bound to torch's current stream."""

    def __init__(self, device: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise GacelleError(_abi.GC_ERR_NO_DEVICE, "no HIP device visible (torch.cuda.is_available() is False)")
        self.lib = _abi.load_library()
        self.device = device
        self.torch_device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        rc = self.lib.gc_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise GacelleError(rc, f"gc_ctx_create({device}) failed")
        self.handle = h
        self.bind_stream()

    def bind_stream(self):
        torch = _torch()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
        self.check(self.lib.gc_ctx_set_stream(self.handle, ctypes.c_void_p(stream)))

    def check(self, rc: int):
        if rc != 0:
            msg = self.lib.gc_last_error(self.handle)
            raise GacelleError(rc, msg.decode() if msg else "error")

    def synchronize(self):
        self.check(self.lib.gc_ctx_synchronize(self.handle))

    def set_option(self, option: int, value: int):
        """gc_ctx_set_option: GC_OPT_GRAPHS (0/1) or GC_OPT_MCMC_BLOCK (0 = autotune, 64, 128, 256)."""
        self.check(self.lib.gc_ctx_set_option(self.handle, int(option), int(value)))

    # ---- buffers ----------------------------------------------------------
    def to_device(self, a, dtype=np.float32):
        torch = _torch()
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.torch_device, dtype=_torch_dtype(dtype)).contiguous()
            return t
        arr = np.ascontiguousarray(np.asarray(a, dtype=dtype))
        return torch.from_numpy(arr).to(self.torch_device)

    def empty(self, shape, dtype=np.float32):
        torch = _torch()
        return torch.empty(tuple(int(s) for s in shape), dtype=_torch_dtype(dtype), device=self.torch_device)

    def zeros(self, shape, dtype=np.float32):
        torch = _torch()
        return torch.zeros(tuple(int(s) for s in shape), dtype=_torch_dtype(dtype), device=self.torch_device)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self.lib.gc_ctx_destroy(self.handle)
        except Exception:
            pass


def _torch_dtype(dtype):
    torch = _torch()
    return {
        np.float32: torch.float32,
        np.float64: torch.float64,
        np.int64: torch.int64,
        np.int32: torch.int32,
        np.uint8: torch.uint8,
    }[np.dtype(dtype).type]


def context(device: int = 0) -> Context:
    """Process-wide context for `device` (created on first use)."""
    with _lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = Context(device)
            _contexts[device] = ctx
        else:
            ctx.bind_stream()
        return ctx


def ptr(t) -> ctypes.c_void_p:
    """Device address of a torch tensor (or NULL for None)."""
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def to_host(t) -> np.ndarray:
    return t.detach().cpu().numpy()
