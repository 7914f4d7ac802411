"""Device plumbing: tensors, streams and the per-thread libslq context.

PyTorch provides device memory and the stream; all arithmetic of the SLQ path
runs in libslq's HIP kernels.
"""

import ctypes

import numpy as np
import torch

from . import _lib


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError(
            "matfree_amd runs on a ROCm GPU and has no CPU fallback "
            "(torch.cuda.is_available() is False)"
        )


def target_device(*candidates):
    """The device of the first CUDA tensor among ``candidates``, else the current one."""
    require_gpu()
    for c in candidates:
        if isinstance(c, torch.Tensor) and c.is_cuda:
            return c.device
    return torch.device("cuda", torch.cuda.current_device())


def as_f32(x, device):
    """A contiguous float32 tensor on ``device`` (copies only when needed)."""
    if isinstance(x, torch.Tensor):
        t = x.detach()
        if t.dtype != torch.float32:
            t = t.to(torch.float32)
        if t.device != device:
            t = t.to(device)
        return t.contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float32), device=device).contiguous()


def scalar(x):
    """A Python float from a number, 0-d array or 0-d tensor."""
    if isinstance(x, torch.Tensor):
        return float(x.detach().cpu())
    return float(np.asarray(x))


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def context(device):
    """libslq context on ``device``, ordered on torch's current stream there."""
    stream = torch.cuda.current_stream(device)
    return _lib.context(device.index, stream.cuda_stream)
