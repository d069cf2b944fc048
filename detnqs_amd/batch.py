"""Multi-instance fused local energy (batched and served forms).

An *instance* is one evaluation of ``compute_variational_energy``
(bridge.cpp:680-714, local_conn.cpp:148-229): an integral context, a selected
space V (device determinants and their DetMap), amplitudes psi_V, screening
mode.  :class:`ElocPlan` runs row ranges of many instances in one launch of
``dq_eloc_plan_*``; the per-row arithmetic is that of the single-instance
kernel, so every sigma_i, and every energy, is bit-identical to
``compute_variational_energy`` on that instance alone.

:func:`eloc_batched` is the batched form: all instances in one persistent
launch, optionally captured into a HIP graph and replayed.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import torch

from . import _lib
from .core import IntCtx, _f64_to_device, _ptr, dets_to_device, stream_handle


class DetMap:
    """Device DetMap of V (det_space.cpp:29-65): ``strict`` rejects duplicates
    (``from_ordered``), otherwise the first occurrence wins."""

    def __init__(self, dets_dev: torch.Tensor, strict: bool = False):
        self.n = int(dets_dev.shape[0])
        h = ctypes.c_void_p()
        _lib.call("dq_detmap_build", _ptr(dets_dev), self.n, int(bool(strict)), stream_handle(), ctypes.byref(h))
        self._h = h

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.call("dq_detmap_free", h)
            except Exception:
                pass
            self._h = None


class VVCache:
    """Lane-ordered H_VV of one instance's V (``dq_vvcache_*``): the elements
    the string kernel would generate, in the order it adds them, so energies
    from it are bit-identical to ``compute_variational_energy`` on any
    amplitudes.  Built once per V (a count and a fill pass of the string
    walk); each later energy is one pass over the layout."""

    def __init__(self, inst: "Instance"):
        if inst.method != "strings":
            raise ValueError("a cached H_VV serves the string method only")
        h = ctypes.c_void_p()
        _lib.call("dq_vvcache_build", inst.ctx.handle, inst.vmap.handle, _ptr(inst.dets), inst.n_rows, inst.n_orb,
                  int(inst.use_hb), float(inst.eps1), stream_handle(), ctypes.byref(h))
        self._h = h
        self.n_rows = inst.n_rows
        n, r, z = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.call("dq_vvcache_info", h, ctypes.byref(n), ctypes.byref(r), ctypes.byref(z))
        self.rounds, self.nnz = int(r.value), int(z.value)

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def nbytes(self) -> int:
        return 12 * 64 * self.rounds + 8 * (self.n_rows + 1)

    def export(self):
        """(off int64[n+1], col uint32 as int32[64 rounds], val float64[64 rounds]) device tensors."""
        dev = torch.device("cuda", torch.cuda.current_device())
        off = torch.empty(self.n_rows + 1, dtype=torch.int64, device=dev)
        col = torch.empty(max(64 * self.rounds, 1), dtype=torch.int32, device=dev)
        val = torch.empty(max(64 * self.rounds, 1), dtype=torch.float64, device=dev)
        _lib.call("dq_vvcache_export", self._h, _ptr(off), _ptr(col), _ptr(val), stream_handle())
        return off, col[: 64 * self.rounds], val[: 64 * self.rounds]

    def energy(self, coeffs: torch.Tensor, sigma: Optional[torch.Tensor] = None) -> float:
        e = ctypes.c_double()
        _lib.call("dq_vvcache_energy", self._h, _ptr(coeffs), _ptr(sigma), ctypes.byref(e), stream_handle())
        return e.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.call("dq_vvcache_free", h)
            except Exception:
                pass
            self._h = None


@dataclass
class Instance:
    """One local-energy evaluation: sigma = H_VV psi and E = psi^T sigma."""

    ctx: IntCtx
    dets: torch.Tensor  # (n, 2) int64 on device
    coeffs: torch.Tensor  # (n,) float64 on device
    n_orb: int
    use_hb: bool = True
    eps1: float = 1e-6
    sigma: Optional[torch.Tensor] = None
    vmap: Optional[DetMap] = field(default=None, repr=False)
    method: str = "strings"
    vv: Optional[VVCache] = field(default=None, repr=False)

    @classmethod
    def create(cls, ctx: IntCtx, dets, coeffs, n_orb: int, use_hb: bool = True, eps1: float = 1e-6,
               method: str = "strings") -> "Instance":
        if method not in _lib.METHODS:
            raise ValueError(f"unknown energy method {method!r}")
        d = dets_to_device(dets)
        c = _f64_to_device(coeffs, d.shape[0], "coefficient")
        if use_hb and not ctx.has_hb:
            raise ValueError("compute_variational_energy: Heat-Bath mode requires non-null hb_table")
        inst = cls(ctx, d, c, int(n_orb), bool(use_hb), float(eps1))
        inst.sigma = torch.empty(d.shape[0], dtype=torch.float64, device=d.device)
        inst.vmap = DetMap(d)
        inst.method = method
        return inst

    @property
    def n_rows(self) -> int:
        return int(self.dets.shape[0])

    def cache_vv(self) -> VVCache:
        """Build (once) the lane-ordered H_VV; later launches use it."""
        if self.vv is None:
            self.vv = VVCache(self)
        return self.vv

    def segment(self, row_begin: int = 0, row_end: Optional[int] = None, energy_slot: int = -1) -> _lib.ElocSegment:
        return _lib.ElocSegment(
            ctx=self.ctx.handle.value, vmap=self.vmap.handle.value, d_dets=self.dets.data_ptr(),
            d_coeffs=self.coeffs.data_ptr(), d_sigma=self.sigma.data_ptr(), n_rows=self.n_rows,
            row_begin=int(row_begin), row_end=self.n_rows if row_end is None else int(row_end),
            n_orb=self.n_orb, use_hb=int(self.use_hb), eps1=self.eps1, energy_slot=int(energy_slot),
            method=_lib.METHODS[self.method], vv=self.vv.handle.value if self.vv is not None else None)


@dataclass
class Tuning:
    """Launch shape (dq_eloc_tuning)."""

    waves_per_block: int = 4
    persistent: bool = True
    grid_blocks: int = 0

    def c(self) -> _lib.ElocTuning:
        return _lib.ElocTuning(int(self.waves_per_block), int(bool(self.persistent)), int(self.grid_blocks))


class ElocPlan:
    """One launch over row ranges of many instances.

    ``ranges`` is a list of ``(instance, row_begin, row_end, energy_slot)``;
    ``energy_slot >= 0`` asks for ``energy[slot] = psi^T sigma`` of that
    instance after the launch (all its rows must be computed by then)."""

    def __init__(self, ranges: Sequence, tuning: Optional[Tuning] = None):
        self.tuning = tuning or Tuning()
        self._keep = list(ranges)  # instances must outlive the plan
        segs = (_lib.ElocSegment * max(1, len(ranges)))()
        for k, (inst, b, e, slot) in enumerate(ranges):
            segs[k] = inst.segment(b, e, slot)
        h = ctypes.c_void_p()
        t = self.tuning.c()
        _lib.call("dq_eloc_plan_create", segs, len(ranges), ctypes.byref(t), ctypes.byref(h))
        self._h = h
        self.n_energy = sum(1 for r in ranges if r[3] >= 0)
        self.total_rows = sum(int(r[2]) - int(r[1]) for r in ranges)
        info = _lib.ElocTuning()
        _lib.call("dq_eloc_plan_info", h, None, None, None, ctypes.byref(info))
        self.grid_blocks = info.grid_blocks
        self._graph_energy = None

    def launch(self, energy: Optional[torch.Tensor] = None, stream=None) -> None:
        _lib.call("dq_eloc_plan_launch", self._h, _ptr(energy), stream_handle(stream))

    def capture(self, energy: Optional[torch.Tensor] = None) -> None:
        """Record the launch into a HIP graph (kernels only)."""
        self._graph_energy = energy
        _lib.call("dq_eloc_plan_capture", self._h, _ptr(energy))

    def replay(self, stream=None) -> None:
        _lib.call("dq_eloc_plan_graph_launch", self._h, stream_handle(stream))

    def check(self, stream=None) -> None:
        _lib.call("dq_eloc_plan_check", self._h, stream_handle(stream))

    def free(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.call("dq_eloc_plan_free", h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def eloc_batched(instances: List[Instance], tuning: Optional[Tuning] = None, graph: bool = False) -> torch.Tensor:
    """Energies psi^T H psi of every instance from one launch (or graph replay);
    returns a device float64 tensor, and leaves each instance's sigma filled."""
    if not instances:
        return torch.zeros(0, dtype=torch.float64)
    energy = torch.empty(len(instances), dtype=torch.float64, device=instances[0].dets.device)
    plan = ElocPlan([(x, 0, x.n_rows, k) for k, x in enumerate(instances)], tuning)
    if graph:
        plan.capture(energy)
        plan.replay()
    else:
        plan.launch(energy)
    plan.check()
    plan.free()
    return energy


__all__ = ["DetMap", "VVCache", "Instance", "Tuning", "ElocPlan", "eloc_batched"]
