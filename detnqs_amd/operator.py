"""Device Hamiltonian operator and energy step (``dq_operator_*``).

Mirrors ``detnqs/operator/kernel.py:48-145`` (``build_vv_operator``,
``build_proxy_operator``: SciPy CSR/CSC products in a host callback) and the
real-wavefunction energy steps of ``detnqs/operator/functional.py:31-268``
(variational, proxy, asymmetric), with the products and reductions on the
device.  Each product entry is the sequential column-order sum SciPy forms, so
it matches SciPy bit for bit; energies and cotangents use fixed-order dots.
"""

from __future__ import annotations

import ctypes
from typing import NamedTuple, Optional

import numpy as np
import torch

from . import _lib
from .core import _ptr, device, stream_handle

MODES = {"variational": 0, "proxy": 1, "asymmetric": 2}


class VVContraction(NamedTuple):
    V: torch.Tensor


class ProxyContraction(NamedTuple):
    V: torch.Tensor
    P: torch.Tensor


def _u32(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device()).to(torch.int32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x).astype(np.uint32)).view(np.int32)).to(device())


def _f64(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device())


class DeviceOperator:
    """H_VV, and for the proxy space H_VP, H_PV = H_VP^T and diag(H_PP), on the device.

    ``vv`` and ``vp`` are COO dicts {row, col, val} as ``core.get_ham_*`` return
    them (host or device arrays)."""

    def __init__(self, n_v: int, vv: dict, vp: Optional[dict] = None, n_p: int = 0, h_pp=None):
        self.n_v, self.n_p = int(n_v), int(n_p)
        r, c, v = _u32(vv["row"]), _u32(vv["col"]), _f64(vv["val"])
        if vp is not None and self.n_p > 0:
            pr, pc, pv = _u32(vp["row"]), _u32(vp["col"]), _f64(vp["val"])
            hp = _f64(h_pp)
            if hp.numel() != self.n_p:
                raise ValueError("diag(H_PP) size must equal n_p")
        else:
            pr = pc = torch.zeros(0, dtype=torch.int32, device=device())
            pv = torch.zeros(0, dtype=torch.float64, device=device())
            hp = None
            self.n_p = 0
        h = ctypes.c_void_p()
        _lib.call("dq_operator_create", self.n_v, self.n_p, _ptr(r), _ptr(c), _ptr(v), r.numel(), _ptr(pr), _ptr(pc),
                  _ptr(pv), pr.numel(), _ptr(hp), stream_handle(), ctypes.byref(h))
        self._h = h

    def apply(self, psi_v: torch.Tensor, psi_p: Optional[torch.Tensor] = None):
        """(y_V, y_P): y_V = H_VV psi_V + H_VP psi_P, y_P = H_PV psi_V + diag(H_PP) psi_P."""
        xv = _f64(psi_v)
        xp = _f64(psi_p) if (psi_p is not None and self.n_p) else None
        if xv.numel() != self.n_v or (xp is not None and xp.numel() != self.n_p):
            raise ValueError("vector size mismatch")
        if self.n_p and xp is None:
            raise ValueError("psi_P is required for a proxy operator")
        yv = torch.empty(self.n_v, dtype=torch.float64, device=xv.device)
        yp = torch.empty(self.n_p, dtype=torch.float64, device=xv.device) if self.n_p else None
        _lib.call("dq_operator_apply", self._h, _ptr(xv), _ptr(xp), _ptr(yv), _ptr(yp), stream_handle())
        return yv, yp

    def energy_step(self, sign, logabs, mode: str = "variational", eps: float = 1e-12):
        """Real-wavefunction energy step: returns (E, cot_logabs, psi, H psi) with
        psi = sign exp(logabs - max), the contraction of ``mode`` and
        cot_logabs = 2 psi (H psi - E psi) / max(den, eps)."""
        if mode not in MODES:
            raise ValueError(f"Unsupported mode: {mode!r}")
        m = MODES[mode]
        n = self.n_v if m == 0 else self.n_v + self.n_p
        s, la = _f64(sign), _f64(logabs)
        if s.numel() != n or la.numel() != n:
            raise ValueError(f"sign/logabs must have {n} entries for mode {mode!r}")
        psi = torch.empty(n, dtype=torch.float64, device=s.device)
        hpsi = torch.empty(n, dtype=torch.float64, device=s.device)
        cot = torch.empty(n, dtype=torch.float64, device=s.device)
        e = ctypes.c_double()
        _lib.call("dq_energy_step", self._h, m, _ptr(s), _ptr(la), float(eps), _ptr(psi), _ptr(hpsi), _ptr(cot),
                  ctypes.byref(e), stream_handle())
        return e.value, cot, psi, hpsi

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.call("dq_operator_free", h)
            except Exception:
                pass
            self._h = None


def build_vv_operator(ham_vv: dict, n_v: Optional[int] = None):
    """psi_V -> VVContraction(V = H_VV psi_V) (kernel.py:48-83)."""
    nv = int(ham_vv.get("shape", (n_v, n_v))[0] if n_v is None else n_v)
    op = DeviceOperator(nv, ham_vv)

    def operator(psi_v):
        return VVContraction(V=op.apply(psi_v)[0])

    operator.device_operator = op
    return operator


def build_proxy_operator(ham_vv: dict, ham_vp: dict, h_diag_p):
    """(psi_V, psi_P) -> ProxyContraction(V, P) (kernel.py:86-140)."""
    nv = int(ham_vv["shape"][0])
    np_ = int(ham_vp["shape"][1])
    op = DeviceOperator(nv, ham_vv, ham_vp, np_, h_diag_p)

    def operator(psi_v, psi_p):
        yv, yp = op.apply(psi_v, psi_p)
        return ProxyContraction(V=yv, P=yp)

    operator.device_operator = op
    return operator


__all__ = ["DeviceOperator", "VVContraction", "ProxyContraction", "build_vv_operator", "build_proxy_operator",
           "MODES"]
