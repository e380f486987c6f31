"""Device-memory plumbing: torch tensors hold the caller's device arrays.

A Julia ``Matrix{Float32}`` of size (Nf, Nt) is column-major; its Python
counterpart here is a torch tensor of shape (Nf, Nt) with strides (1, Nf),
so ``x[i, j]`` (0-based) is the reference's ``x[i+1, j+1]`` and the raw
pointer is exactly the column-major buffer the C-ABI takes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

_hip = None


def _hiprt():
    """The HIP runtime already loaded by torch (same soname, same instance)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpyAsync.restype = C.c_int
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return _hip


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("libzdt_hip needs a GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def is_colmajor(t: torch.Tensor) -> bool:
    return t.dim() == 2 and (t.shape[1] <= 1 or t.stride(1) == t.shape[0]) and (t.shape[0] <= 1 or t.stride(0) == 1)


def colmajor(x, dtype=torch.float32) -> torch.Tensor:
    """Device tensor with column-major storage holding x (copied if needed)."""
    if isinstance(x, np.ndarray):
        a = np.asfortranarray(x, dtype=np.float32 if dtype == torch.float32 else None)
        t = torch.from_numpy(np.ascontiguousarray(a.T)).to(device()).t()
        return t
    if not isinstance(x, torch.Tensor):
        raise TypeError("expected a torch.Tensor or numpy.ndarray")
    if x.dim() != 2:
        raise ValueError("expected a matrix")
    if x.dtype != dtype:
        x = x.to(dtype)
    if not x.is_cuda:
        x = x.to(device())
    if not is_colmajor(x):
        x = x.t().contiguous().t()
    return x


def empty_colmajor(rows: int, cols: int, dtype=torch.float32, dev=None) -> torch.Tensor:
    return torch.empty((cols, rows), dtype=dtype, device=dev or device()).t()


def copy_from_ptr(ptr: int, rows: int, cols: int, complex_: bool = True) -> torch.Tensor:
    """Copy of a column-major device buffer owned by the library."""
    dt = torch.complex64 if complex_ else torch.float32
    out = empty_colmajor(rows, cols, dt)
    nbytes = rows * cols * (8 if complex_ else 4)
    rc = _hiprt().hipMemcpyAsync(C.c_void_p(out.data_ptr()), C.c_void_p(ptr), nbytes, 3, C.c_void_p(stream_ptr()))
    if rc != 0:
        raise RuntimeError(f"hipMemcpyAsync failed ({rc})")
    return out
