"""Device plumbing: torch-ROCm tensors as HBM buffers and the current stream.

PyTorch is only the allocator and stream provider here; every kernel is
launched through libssmhip's C ABI with raw pointers.
"""

from __future__ import annotations

import numpy as np

from . import _lib

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t  # noqa: PLC0415

        _torch = t
    return _torch


_gpu_ok = False


def require_gpu():
    """Return the torch module after checking (once per process) that a HIP device is usable."""
    global _gpu_ok
    t = torch()
    if not _gpu_ok:
        if not t.cuda.is_available():
            raise RuntimeError("ssmhip needs a ROCm GPU (torch.cuda.is_available() is False); there is no CPU path")
        _lib.load()
        _gpu_ok = True
    return t


def device():
    t = require_gpu()
    return t.device("cuda", t.cuda.current_device())


def stream_ptr(stream=None) -> int:
    t = require_gpu()
    s = stream if stream is not None else t.cuda.current_stream()
    return int(s.cuda_stream)


def to_dev(a: np.ndarray, dtype=None):
    t = require_gpu()
    arr = np.ascontiguousarray(a if dtype is None else a.astype(dtype, copy=False))
    return t.from_numpy(arr).to(device(), non_blocking=False)


def empty(shape, dtype):
    t = require_gpu()
    return t.empty(shape, dtype=dtype, device=device())


def zeros(shape, dtype):
    t = require_gpu()
    return t.zeros(shape, dtype=dtype, device=device())


def ptr(x) -> int | None:
    """Raw device pointer of a tensor (None -> NULL)."""
    if x is None:
        return None
    if not x.is_cuda:
        raise ValueError("expected a device tensor")
    if not x.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return int(x.data_ptr())


# outputs from 1 MiB up to 64 MiB are copied to the host through pinned buffers (torch's
# caching host allocator): smaller ones gain less than the pinned buffer costs, and larger ones
# go through an ordinary copy, so that results a caller keeps hold little page-locked memory
PINNED_COPY_MIN_BYTES = 1 << 20
PINNED_COPY_MAX_BYTES = 64 << 20


def to_host(*tensors):
    """Device tensors -> NumPy arrays, copied on the current stream and waited for once.
    Outputs of 1-64 MiB land in pinned host buffers (a direct DMA, no staging copy); each
    returned array owns its buffer."""
    t = require_gpu()
    out = []
    pinned = []
    for x in tensors:
        if PINNED_COPY_MIN_BYTES <= x.numel() * x.element_size() <= PINNED_COPY_MAX_BYTES:
            h = t.empty(x.shape, dtype=x.dtype, pin_memory=True)
            h.copy_(x, non_blocking=True)
            pinned.append(len(out))
            out.append(h)
        else:
            out.append(None)
    if pinned:
        t.cuda.current_stream().synchronize()
    return [out[i].numpy() if out[i] is not None else x.cpu().numpy() for i, x in enumerate(tensors)]
