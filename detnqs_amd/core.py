"""Drop-in mirror of ``detnqs.core`` for the local-energy path, backed by
libdetnqs_hip.so on an MI355X.

Signatures, argument meaning, defaults, return layouts and exception types
follow the reference nanobind bridge (``csrc/bridge/bridge.cpp:235-772``,
re-exported by ``detnqs/core.py:19-88``):

* ``IntCtx(fcidump_path, num_orb)``, ``.get_e_nuc()``, ``.hb_prepare(threshold)``,
  ``.hb_clear()``                                  (bridge.cpp:242-248)
* ``get_ham_diag(dets, int_ctx)``                  (bridge.cpp:450-457)
* ``get_local_connections(dets, int_ctx, n_orb, use_heatbath, eps1)``
                                                  (bridge.cpp:635-673)
* ``get_ham_vv``, ``get_ham_block``, ``get_ham_conn``, ``get_ham_conn_amp``,
  ``get_ham_eff``, ``compute_variational_energy`` (bridge.cpp:459-714)

Inputs are host NumPy arrays as in the reference (C-contiguous uint64 (N, 2)
determinants, float64 vectors); they are moved to HBM, the kernels run there,
and results come back as NumPy arrays.  ``*_device`` variants take and return
device tensors (torch is used only for device memory and streams).

Errors: invalid arguments raise ``ValueError``, out-of-range indices
``IndexError``, HIP failures ``RuntimeError`` (the nanobind translations of
std::invalid_argument / out_of_range / runtime_error).
"""

from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib

HEATBATH_THRESH = 1e-15  # ham_utils.hpp:42, default of IntCtx.hb_prepare


# ----------------------------------------------------------------------------- device plumbing

def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("detnqs_amd needs a HIP device (MI355X / gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream: Optional[torch.cuda.Stream] = None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t: Optional[torch.Tensor]) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def dets_to_device(dets) -> torch.Tensor:
    """uint64 (N, 2) host dets -> int64 (N, 2) device tensor (same bits)."""
    if isinstance(dets, torch.Tensor):
        t = dets
        if t.dtype not in (torch.int64, torch.uint64) or t.dim() != 2 or t.shape[1] != 2:
            raise ValueError("dets must be a (N, 2) 64-bit tensor")
        if t.dtype == torch.uint64:
            t = t.view(torch.int64)
        return t.to(device()).contiguous()
    a = np.asarray(dets)
    if a.dtype != np.uint64:
        a = a.astype(np.uint64)
    a = np.ascontiguousarray(a)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError(f"dets must have shape (N, 2), got {a.shape}")
    return torch.from_numpy(a.view(np.int64)).to(device())


def dets_to_host(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64).reshape(-1, 2)


def _f64_to_device(x, n: int, name: str) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        t = x.to(device=device(), dtype=torch.float64).contiguous()
    else:
        t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device())
    if t.dim() != 1 or t.shape[0] != n:
        raise ValueError(f"{name} size mismatch: expected {n}, got {tuple(t.shape)}")
    return t


# ----------------------------------------------------------------------------- IntCtx

class IntCtx:
    """Molecular integral context (MO/SO + Heat-Bath table), HBM resident.

    Mirrors ``IntCtx`` of the reference bridge (bridge.cpp:108-128)."""

    def __init__(self, fcidump_path: str, num_orb: int):
        device()
        h = ctypes.c_void_p()
        _lib.call("dq_ctx_create_fcidump", str(fcidump_path).encode(), int(num_orb), ctypes.byref(h))
        self._h = h
        self._read_info()

    @classmethod
    def from_integrals(cls, n_orb: int, n_elec: int, h1e: np.ndarray, h2e: np.ndarray, e_nuc: float) -> "IntCtx":
        """Context from packed integral arrays (integral_mo.hpp:30-49 key order)."""
        device()
        n1 = n_orb * (n_orb + 1) // 2
        h1 = np.ascontiguousarray(h1e, dtype=np.float64)
        h2 = np.ascontiguousarray(h2e, dtype=np.float64)
        if h1.size != n1 or h2.size != n1 * (n1 + 1) // 2:
            raise ValueError("packed integral sizes do not match n_orb")
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        _lib.call("dq_ctx_create", int(n_orb), int(n_elec), h1.ctypes.data_as(ctypes.c_void_p),
                  h2.ctypes.data_as(ctypes.c_void_p), float(e_nuc), ctypes.byref(h))
        self._h = h
        self._read_info()
        return self

    def _read_info(self):
        n, ne, en = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.call("dq_ctx_info", self._h, ctypes.byref(n), ctypes.byref(ne), ctypes.byref(en))
        self.n_orb, self.n_elec, self._e_nuc = n.value, ne.value, en.value
        self._has_hb = False

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def get_e_nuc(self) -> float:
        return self._e_nuc

    def hb_prepare(self, threshold: float = HEATBATH_THRESH) -> None:
        _lib.call("dq_hb_prepare", self._h, float(threshold))
        self._has_hb = True

    def hb_clear(self) -> None:
        _lib.call("dq_hb_clear", self._h)
        self._has_hb = False

    @property
    def has_hb(self) -> bool:
        return self._has_hb

    def hb_export(self) -> dict:
        """Heat-bath table as host arrays (row_offsets, a, b, w)."""
        n = ctypes.c_int64()
        _lib.call("dq_hb_export", self._h, ctypes.byref(n), None, None, None, None)
        nso = 2 * self.n_orb
        off = np.empty(nso * (nso - 1) // 2 + 1, dtype=np.uint64)
        a = np.empty(n.value, dtype=np.int32)
        b = np.empty(n.value, dtype=np.int32)
        w = np.empty(n.value, dtype=np.float64)
        _lib.call("dq_hb_export", self._h, ctypes.byref(n), off.ctypes.data_as(ctypes.c_void_p),
                  a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                  w.ctypes.data_as(ctypes.c_void_p))
        return {"row_offsets": off, "a": a, "b": b, "w": w}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().dq_ctx_destroy(h)
            except Exception:
                pass
            self._h = None


# ----------------------------------------------------------------------------- diagonal

def get_ham_diag_device(dets: torch.Tensor, int_ctx: IntCtx) -> torch.Tensor:
    d = dets_to_device(dets)
    out = torch.empty(d.shape[0], dtype=torch.float64, device=d.device)
    _lib.call("dq_ham_diag", int_ctx.handle, _ptr(d), d.shape[0], _ptr(out), stream_handle())
    return out


def get_ham_diag(dets, int_ctx: IntCtx) -> np.ndarray:
    """Diagonal elements <x|H|x> (bridge.cpp:450-457)."""
    return get_ham_diag_device(dets_to_device(dets), int_ctx).cpu().numpy()


# ----------------------------------------------------------------------------- local connections

def local_conn_count(dets: torch.Tensor, int_ctx: IntCtx, n_orb: int, use_heatbath: bool = False,
                     eps1: float = 1e-6, offsets: Optional[torch.Tensor] = None, want_total: bool = True):
    """Count pass: int64 offsets[n+1] (exclusive scan of kept row sizes)."""
    n = dets.shape[0]
    if offsets is None:
        offsets = torch.empty(n + 1, dtype=torch.int64, device=dets.device)
    tot = ctypes.c_int64(-1)
    _lib.call("dq_local_conn_count", int_ctx.handle, _ptr(dets), n, int(n_orb), int(bool(use_heatbath)),
              float(eps1), _ptr(offsets), ctypes.byref(tot) if want_total else None, stream_handle())
    return offsets, (tot.value if want_total else None)


def local_conn_fill(dets: torch.Tensor, int_ctx: IntCtx, n_orb: int, use_heatbath: bool, eps1: float,
                    offsets: torch.Tensor, out_dets: torch.Tensor, out_vals: torch.Tensor) -> None:
    _lib.call("dq_local_conn_fill", int_ctx.handle, _ptr(dets), dets.shape[0], int(n_orb), int(bool(use_heatbath)),
              float(eps1), _ptr(offsets), _ptr(out_dets), _ptr(out_vals), stream_handle())


def get_local_connections_device(dets, int_ctx: IntCtx, n_orb: int, use_heatbath: bool = False,
                                 eps1: float = 1e-6):
    """Device form: (offsets int64[n+1], dets int64[M,2], values f64[M]) tensors."""
    d = dets_to_device(dets)
    off, total = local_conn_count(d, int_ctx, n_orb, use_heatbath, eps1)
    od = torch.empty((total, 2), dtype=torch.int64, device=d.device)
    ov = torch.empty(total, dtype=torch.float64, device=d.device)
    if total > 0:
        local_conn_fill(d, int_ctx, n_orb, use_heatbath, eps1, off, od, ov)
    return off, od, ov


def get_local_connections(dets, int_ctx: IntCtx, n_orb: int, use_heatbath: bool = False,
                          eps1: float = 1e-6) -> dict:
    """H|x> for a batch as {offsets, dets, values} (bridge.cpp:635-673).

    ``offsets`` is int32 as in the reference when the total fits, int64 when it
    does not (the reference's int32 offsets overflow there, local_conn.cpp:129-132)."""
    off, od, ov = get_local_connections_device(dets, int_ctx, n_orb, use_heatbath, eps1)
    offsets = off.cpu().numpy()
    if offsets[-1] < 2 ** 31:
        offsets = offsets.astype(np.int32)
    return {"offsets": offsets, "dets": dets_to_host(od), "values": ov.cpu().numpy()}


__all__ = [
    "IntCtx",
    "HEATBATH_THRESH",
    "get_ham_diag",
    "get_ham_diag_device",
    "get_local_connections",
    "get_local_connections_device",
    "local_conn_count",
    "local_conn_fill",
    "dets_to_device",
    "dets_to_host",
    "device",
    "stream_handle",
]
