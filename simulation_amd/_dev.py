"""Device-buffer plumbing: torch tensors in, raw pointers and streams out."""
from __future__ import annotations

import ctypes as C

import torch

_f = C.POINTER(C.c_float)


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("a ROCm GPU is required (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def ptr(t: torch.Tensor | None, dtype: torch.dtype = torch.float32) -> int | None:
    """Raw device pointer of a contiguous device tensor (None passes through)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def fptr(t: torch.Tensor | None):
    """float* for struct fields."""
    p = ptr(t)
    return C.cast(p, _f) if p else _f()


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream
