"""Batched FP64 slogdet on the MI355X, mirroring ``detnqs.models.slogdet``.

``slogdet(A, use_fast_kernel=True)`` keeps the reference's dispatch rule
(``detnqs/models/slogdet.py:159-188``): real float matrices with 2 <= n <= 64
go to the fused kernel, anything else to the library routine.  The fused path
is differentiable through a custom autograd function, the torch counterpart of
the reference's ``jax.custom_vjp`` (slogdet.py:103-154): the forward keeps the
LU factors and row permutation, the backward returns cot_logabs * A^{-T}
from them (``dq_slogdet_bwd_lu_f64``) instead of refactorising; the sign is
not differentiable.  ``slogdet_bwd_from_a`` is the reference FFI backward
signature (A, cot_sign, cot_logabs) -> grad, which refactorises.

float32 inputs are computed in FP64 and the results rounded to float32: through
``slogdet_fwd`` / ``slogdet_bwd_from_a`` with the FP32 entry points (the
reference's f32 FFI targets), and through ``slogdet`` by promoting to FP64 for
the LU-keeping differentiable path.
"""

from __future__ import annotations

import torch

from . import _lib
from .core import _ptr, device, stream_handle


def _check(A: torch.Tensor):
    if A.dim() < 2 or A.shape[-1] != A.shape[-2]:
        raise ValueError(f"Expected square matrix, got shape {tuple(A.shape)}")
    n = int(A.shape[-1])
    if not 2 <= n <= 64:
        raise ValueError(f"Matrix size {n} outside [2, 64] range")
    if A.device.type != "cuda":
        raise ValueError("fused slogdet needs a device tensor")
    return n


def slogdet_fwd(A: torch.Tensor, out=None):
    """(sign, logabs) of a (..., n, n) FP64 device tensor; ``out`` = (sign, logabs)
    preallocated (contiguous FP64, batch shape) to write into."""
    n = _check(A)
    A = A.contiguous()
    B = A.numel() // (n * n)
    dt = _dtype(A)
    if out is None:
        sign = torch.empty(A.shape[:-2], dtype=dt, device=A.device)
        logabs = torch.empty_like(sign)
    else:
        sign, logabs = out
        for t in (sign, logabs):
            if t.dtype != dt or not t.is_contiguous() or t.numel() != B or t.device != A.device:
                raise ValueError(f"out tensors must be contiguous {dt} with one entry per matrix")
    name = "dq_slogdet_fwd_f64" if dt == torch.float64 else "dq_slogdet_fwd_f32"
    _lib.call(name, _ptr(A), B, n, _ptr(sign), _ptr(logabs), stream_handle())
    return sign, logabs


def _dtype(A: torch.Tensor) -> torch.dtype:
    if A.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"Unsupported dtype {A.dtype}")
    return A.dtype


def slogdet_fwd_lu(A: torch.Tensor):
    """(sign, logabs, lu, perm): L\\U in position order and perm[k] = original row."""
    n = _check(A)
    A = A.contiguous()
    B = A.numel() // (n * n)
    sign = torch.empty(A.shape[:-2], dtype=torch.float64, device=A.device)
    logabs = torch.empty_like(sign)
    lu = torch.empty_like(A)
    perm = torch.empty(A.shape[:-1], dtype=torch.int32, device=A.device)
    _lib.call("dq_slogdet_fwd_lu_f64", _ptr(A), B, n, _ptr(sign), _ptr(logabs), _ptr(lu), _ptr(perm),
              stream_handle())
    return sign, logabs, lu, perm


def slogdet_bwd_from_a(A: torch.Tensor, cot_sign, cot_logabs: torch.Tensor) -> torch.Tensor:
    """Reference FFI backward (logdet_cuda.cu:522-550): grad = cot_logabs * A^{-T}."""
    n = _check(A)
    A = A.contiguous()
    B = A.numel() // (n * n)
    dt = _dtype(A)
    cot = cot_logabs.to(dt).contiguous()
    if cot.numel() != B:
        raise ValueError("cot_logabs batch shape mismatch")
    grad = torch.empty_like(A)
    name = "dq_slogdet_bwd_f64" if dt == torch.float64 else "dq_slogdet_bwd_f32"
    _lib.call(name, _ptr(A), _ptr(None), _ptr(cot), B, n, _ptr(grad), stream_handle())
    return grad


def slogdet_bwd_from_lu(lu: torch.Tensor, perm: torch.Tensor, cot_logabs: torch.Tensor) -> torch.Tensor:
    n = _check(lu)
    B = lu.numel() // (n * n)
    cot = cot_logabs.to(torch.float64).contiguous()
    if cot.numel() != B:
        raise ValueError("cot_logabs batch shape mismatch")
    grad = torch.empty_like(lu)
    _lib.call("dq_slogdet_bwd_lu_f64", _ptr(lu), _ptr(perm), _ptr(cot), B, n, _ptr(grad), stream_handle())
    return grad


class _FusedLogdet(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A):
        sign, logabs, lu, perm = slogdet_fwd_lu(A)
        ctx.save_for_backward(lu, perm)
        ctx.mark_non_differentiable(sign)
        return sign, logabs

    @staticmethod
    def backward(ctx, cot_sign, cot_logabs):
        lu, perm = ctx.saved_tensors
        if cot_logabs is None:
            return torch.zeros_like(lu)
        return slogdet_bwd_from_lu(lu, perm, cot_logabs)


def slogdet(A: torch.Tensor, *, use_fast_kernel: bool = True):
    """Sign and log|det| of (..., n, n); same dispatch as the reference."""
    if not use_fast_kernel:
        return torch.linalg.slogdet(A)
    is_real = A.dtype in (torch.float32, torch.float64)
    n = A.shape[-1] if (A.dim() >= 2 and A.shape[-1] == A.shape[-2]) else 0
    if not (is_real and 2 <= n <= 64):
        return torch.linalg.slogdet(A)
    if A.device.type != "cuda":
        A = A.to(device())
    if A.dtype == torch.float32:
        s, la = _FusedLogdet.apply(A.to(torch.float64))
        return s.to(torch.float32), la.to(torch.float32)
    return _FusedLogdet.apply(A)


#: flag of a tile code: the kernel that reads the bounds at run time, also at
#: n equal to its bucket's column count (12, 16, 24, 32)
TILE_RUNTIME_N = 0x100


def tile_choices(n: int) -> list:
    """Tile codes the 8 < n <= 64 kernels are built for at this n: lanes per
    matrix, and at n = 12, 16, 24 and 32 each also | TILE_RUNTIME_N (the same
    arithmetic with the bounds read at run time rather than fixed)."""
    import ctypes
    buf = (ctypes.c_int * 8)()
    cnt = ctypes.c_int()
    _lib.call("dq_slogdet_tile_choices", int(n), buf, 8, ctypes.byref(cnt))
    return list(buf[: cnt.value])


#: kernel kinds of the tile table
KINDS = {"forward": 0, "forward_lu": 1, "backward": 2, "backward_lu": 3}


def _kind(kind) -> int:
    if isinstance(kind, str):
        if kind not in KINDS:
            raise ValueError(f"unknown slogdet kernel kind {kind!r}; one of {sorted(KINDS)}")
        return KINDS[kind]
    return int(kind)


def get_tile(kind, n: int) -> int:
    """Tile code (lanes per matrix, maybe | TILE_RUNTIME_N) used for launches of this kernel kind ("forward",
    "forward_lu", "backward", "backward_lu" or 0-3) at this n (8 < n <= 64)."""
    import ctypes
    v = ctypes.c_int()
    _lib.call("dq_slogdet_tile_get", _kind(kind), int(n), ctypes.byref(v))
    return v.value


def set_tile(kind, n: int, lanes: int) -> None:
    """Pick the tile code for every later launch of this kernel kind at
    this n in the process (0 restores the default); ValueError for an unbuilt
    choice."""
    _lib.call("dq_slogdet_tile_set", _kind(kind), int(n), int(lanes))


__all__ = ["slogdet", "slogdet_fwd", "slogdet_fwd_lu", "slogdet_bwd_from_a", "slogdet_bwd_from_lu",
           "tile_choices", "get_tile", "set_tile", "KINDS", "TILE_RUNTIME_N"]
