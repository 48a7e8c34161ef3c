"""Shared plumbing of the drop-in modules: numpy <-> device tensors."""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("the mwdiem_amd drop-ins need a ROCm GPU")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(a, dtype: torch.dtype, dev: torch.device) -> torch.Tensor:
    """Contiguous device tensor of `a` (numpy array or tensor) in `dtype`."""
    if isinstance(a, torch.Tensor):
        t = a.to(device=dev, dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a))).to(device=dev, dtype=dtype)
    return t.contiguous()


def write_back(dst, src: torch.Tensor):
    """Copy a device result into the caller's array in place (numpy or tensor),
    in the caller's dtype, as pycuda's cuda.InOut did."""
    if isinstance(dst, torch.Tensor):
        dst.copy_(src.reshape(dst.shape).to(dtype=dst.dtype))
    else:
        dst[...] = src.reshape(dst.shape).cpu().numpy().astype(dst.dtype, copy=False)
    return dst


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream or None
