"""Device context and device-memory plumbing for the libgeofem C ABI.

PyTorch is used only to own device memory (tensors on the HIP device) and for
its current stream; all computation runs in libgeofem's HIP kernels on the
context's own stream.  Every library call returns after its work is done, and
torch's current stream is synchronised before a call consumes tensors, so no
cross-stream hazard exists in either direction.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import check, lib


def _torch():
    import torch
    return torch


class Context:
    """One HIP stream on one device (gs_ctx)."""

    def __init__(self, device: int | None = None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device is visible: libgeofem needs a GPU")
        self.device = torch.cuda.current_device() if device is None else int(device)
        h = C.c_void_p()
        st = lib().gs_ctx_create(self.device, C.byref(h))
        if st != _lib.GS_OK:
            raw = lib().gs_last_error(None)
            raise _lib.GeoFemError(st, raw.decode() if raw else "gs_ctx_create")
        self.handle = h
        self._lock = threading.Lock()

    def close(self):
        if getattr(self, "handle", None):
            lib().gs_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers ------------------------------------------------------------
    def torch_device(self):
        return _torch().device("cuda", self.device)

    def sync_torch(self):
        torch = _torch()
        torch.cuda.current_stream(self.device).synchronize()

    def check(self, status, what=""):
        check(status, self.handle, what)

    def empty(self, n, dtype=None):
        torch = _torch()
        return torch.empty(int(n), dtype=dtype or torch.float64,
                           device=self.torch_device())

    def to_device(self, a, dtype=None):
        """numpy / torch -> contiguous float64 tensor on this device."""
        torch = _torch()
        dtype = dtype or torch.float64
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.torch_device(), dtype=dtype)
        else:
            t = torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(
                self.torch_device())
        return t.contiguous()


_default: Context | None = None
_default_lock = threading.Lock()


def default_context() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context()
        return _default


def dptr(t) -> C.c_void_p:
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return C.c_void_p(None)
    return C.c_void_p(t.data_ptr())


__all__ = ["Context", "default_context", "dptr", "_lib"]
