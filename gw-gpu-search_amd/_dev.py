"""Device plumbing for the host modules: torch tensors carry device memory and the current HIP stream; every
computation goes through libgwgs_hip.so (or hipFFT through torch.fft)."""
from __future__ import annotations

import numpy as np
import torch

import _lib

DEV = 'cuda'


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError('gw-gpu-search_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    _lib.load()


def ptr(t):
    """Device address of a tensor (None -> NULL)."""
    return None if t is None else t.data_ptr()


def stream():
    return torch.cuda.current_stream().cuda_stream


def as_dev(x, dtype=None):
    """numpy / torch -> contiguous torch tensor on the GPU."""
    if isinstance(x, torch.Tensor):
        t = x if x.is_cuda else x.to(DEV)
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()
    a = np.ascontiguousarray(x)
    t = torch.from_numpy(a).to(DEV)
    return t if dtype is None else t.to(dtype)


def host(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def i32(x):
    return as_dev(np.asarray(x, np.int32))


def i64(x):
    return as_dev(np.asarray(x, np.int64))
