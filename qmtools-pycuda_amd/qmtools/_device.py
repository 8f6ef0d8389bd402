"""Device memory and stream plumbing (PyTorch-ROCm tensors as buffers).

Device buffers are torch tensors on `cuda:<dev>`; the native library sees only
their `data_ptr()` and the raw HIP stream handle of torch's current stream.
torch is imported lazily so that host-only code (basis parsing, grid files)
works on machines without a GPU.
"""
from __future__ import annotations

import numpy

_torch = None
_DEVICE_INDEX = 0


def torch():
    global _torch
    if _torch is None:
        import torch as _t  # noqa: WPS433 (lazy import on purpose)
        _torch = _t
    return _torch


def available() -> bool:
    try:
        return bool(torch().cuda.is_available())
    except Exception:  # pragma: no cover - torch missing entirely
        return False


def require():
    """Return the torch device, or raise if no HIP device is present."""
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("qmtools: no HIP device is available; the grid kernels run on the GPU only")
    return t.device("cuda", _DEVICE_INDEX)


def set_device(index: int):
    global _DEVICE_INDEX
    _DEVICE_INDEX = int(index)


def device_index() -> int:
    return _DEVICE_INDEX


def stream_handle() -> int:
    """Raw hipStream_t of torch's current stream on the active device."""
    t = torch()
    return int(t.cuda.current_stream(require()).cuda_stream)


_NP2T = None


def _dtype(np_dtype):
    t = torch()
    d = numpy.dtype(np_dtype)
    table = {numpy.dtype(numpy.float64): t.float64, numpy.dtype(numpy.float32): t.float32,
             numpy.dtype(numpy.float16): t.float16, numpy.dtype(numpy.int32): t.int32,
             numpy.dtype(numpy.int16): t.int16, numpy.dtype(numpy.int64): t.int64,
             numpy.dtype(numpy.uint8): t.uint8}
    return table[d]


def empty(n, np_dtype):
    return torch().empty(int(n), dtype=_dtype(np_dtype), device=require())


def zeros(n, np_dtype):
    return torch().zeros(int(n), dtype=_dtype(np_dtype), device=require())


def from_host(arr):
    a = numpy.ascontiguousarray(arr)
    return torch().from_numpy(a.copy()).to(require())


def copy_from_host(dst, arr):
    a = numpy.ascontiguousarray(arr)
    src = torch().from_numpy(a).to(dtype=dst.dtype)
    dst.copy_(src.reshape(dst.shape))


def to_host(t):
    return t.detach().cpu().numpy()


def ptr(t) -> int:
    return int(t.data_ptr())


def synchronize():
    torch().cuda.synchronize(require())
