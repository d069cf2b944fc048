"""Serving form: a device-resident continuous-batching loop for local-energy
requests (``dq_server_*``).

Concurrent VMC runs (e.g. the geometries of a potential-energy scan, each with
its own :class:`~detnqs_amd.core.IntCtx`) submit ``compute_variational_energy``
requests of uneven size from any thread; :meth:`Server.submit` returns a ticket
and :meth:`Server.wait` the energy with its timestamps.  The loop admits
requests in arrival order while their device memory fits the HBM budget, keeps
integral contexts and keyed V sets resident, and runs every admitted request's
rows in shared launches.  A V set kept under a ``v_key`` also keeps its
lane-ordered H_VV (:class:`~detnqs_amd.batch.VVCache` layout), so repeated
requests on it (VMC inner steps: same V, new amplitudes) are one pass over
stored elements instead of the string walk.  Each result is bit-identical to
``core.compute_variational_energy`` on the same inputs.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from .batch import Tuning
from .core import IntCtx


@dataclass
class Result:
    energy: float
    t_submit: float
    t_admit: float
    t_first_launch: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


@dataclass
class Stats:
    submitted: int
    completed: int
    launches: int
    rows: int
    v_hits: int
    v_misses: int
    v_evictions: int
    v_cached: int
    arena_bytes: int
    arena_used: int
    arena_peak: int
    queued: int
    running: int
    vv_builds: int
    vv_rows: int


class Server:
    """Continuous-batching local-energy server on the current device."""

    def __init__(self, hbm_budget_bytes: int = 8 << 30, rows_per_launch: int = 32768, max_inflight: int = 2,
                 tuning: Optional[Tuning] = None, vv_layout: bool = True):
        t = (tuning or Tuning(4, True, 0)).c()
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), int(rows_per_launch), int(max_inflight), t,
                                int(bool(vv_layout)))
        h = ctypes.c_void_p()
        _lib.call("dq_server_create", ctypes.byref(cfg), ctypes.byref(h))
        self._h = h
        self._keep = {}  # ticket -> inputs that must stay alive until the wait

    def submit(self, ctx: IntCtx, dets, coeffs, n_orb: int, use_hb: bool = True, eps1: float = 1e-6,
               v_key: int = 0, method: str = "strings") -> int:
        """Queue one request.  ``dets`` (N, 2) uint64 and ``coeffs`` (N,) float64
        may be host arrays (copied at submit) or device tensors (used in place
        until the result is waited for; dets are copied to the server)."""
        if use_hb and not ctx.has_hb:
            raise ValueError("compute_variational_energy: Heat-Bath mode requires non-null hb_table")
        if method not in _lib.METHODS:
            raise ValueError(f"unknown energy method {method!r}")
        keep = []
        if isinstance(dets, torch.Tensor):
            d = dets.contiguous()
            if d.dtype == torch.uint64:
                d = d.view(torch.int64)
            if not d.is_cuda or d.dim() != 2 or d.shape[1] != 2 or d.dtype != torch.int64:
                raise ValueError("device dets must be a (N, 2) 64-bit CUDA tensor")
            dptr, d_on, n = d.data_ptr(), 1, d.shape[0]
            keep.append(d)
        else:
            a = np.ascontiguousarray(np.asarray(dets), dtype=np.uint64)
            if a.ndim != 2 or a.shape[1] != 2:
                raise ValueError(f"dets must have shape (N, 2), got {a.shape}")
            dptr, d_on, n = a.ctypes.data, 0, a.shape[0]
            keep.append(a)
        if isinstance(coeffs, torch.Tensor) and coeffs.is_cuda:
            c = coeffs.to(torch.float64).contiguous()
            cptr, c_on = c.data_ptr(), 1
        else:
            c = np.ascontiguousarray(np.asarray(coeffs.cpu() if isinstance(coeffs, torch.Tensor) else coeffs),
                                     dtype=np.float64)
            cptr, c_on = c.ctypes.data, 0
        keep.append(c)
        if c.shape[0] != n or len(c.shape) != 1:
            raise ValueError("coefficient size mismatch")
        stream = torch.cuda.current_stream().cuda_stream if (d_on or c_on) else 0
        req = _lib.ServerRequest(ctx.handle.value, dptr if n else None, d_on, cptr if n else None, c_on, int(n),
                                 int(n_orb), int(bool(use_hb)), float(eps1), int(v_key), stream,
                                 _lib.METHODS[method])
        t = ctypes.c_int64()
        _lib.call("dq_server_submit", self._h, ctypes.byref(req), ctypes.byref(t))
        if d_on or c_on:
            self._keep[t.value] = keep  # device inputs are read by the loop later
        return t.value

    def wait(self, ticket: int, timeout: Optional[float] = None) -> Result:
        r = _lib.ServerResult()
        try:
            _lib.call("dq_server_wait", self._h, int(ticket), -1.0 if timeout is None else float(timeout),
                      ctypes.byref(r))
        except TimeoutError:
            raise  # still running: its device inputs stay referenced
        except Exception:
            self._keep.pop(ticket, None)
            raise
        self._keep.pop(ticket, None)
        return Result(r.energy, r.t_submit, r.t_admit, r.t_first_launch, r.t_done)

    @staticmethod
    def request_bytes(n_rows: int, coeffs_on_device: bool = False) -> int:
        """Arena bytes a request of n_rows holds while it runs."""
        b = ctypes.c_int64()
        _lib.call("dq_server_request_bytes", int(n_rows), int(bool(coeffs_on_device)), ctypes.byref(b))
        return b.value

    def stats(self) -> Stats:
        s = _lib.ServerStats()
        _lib.call("dq_server_stats_get", self._h, ctypes.byref(s))
        return Stats(*[getattr(s, k) for k, _ in _lib.ServerStats._fields_])

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.call("dq_server_destroy", h)
        self._h = None
        self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server", "Result", "Stats"]
