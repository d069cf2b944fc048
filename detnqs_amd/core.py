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


# ----------------------------------------------------------------------------- projected blocks

class Blocks:
    """Device-resident H_VV / H_VP / P result (dq_blocks)."""

    def __init__(self, h: ctypes.c_void_p):
        self._h = h
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.call("dq_blocks_sizes", h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        self.nnz_vv, self.nnz_vp, self.n_p = a.value, b.value, c.value

    def to_device(self):
        """(vv_row, vv_col, vv_val, vp_row, vp_col, vp_val, det_P) as device tensors."""
        dev = device()
        i32 = lambda m: torch.empty(m, dtype=torch.int32, device=dev)  # noqa: E731
        f64 = lambda m: torch.empty(m, dtype=torch.float64, device=dev)  # noqa: E731
        out = (i32(self.nnz_vv), i32(self.nnz_vv), f64(self.nnz_vv), i32(self.nnz_vp), i32(self.nnz_vp),
               f64(self.nnz_vp), torch.empty((self.n_p, 2), dtype=torch.int64, device=dev))
        _lib.call("dq_blocks_copy", self._h, *[_ptr(t) for t in out], 1, stream_handle())
        return out

    def to_host(self):
        u32 = lambda m: np.empty(m, dtype=np.uint32)  # noqa: E731
        out = (u32(self.nnz_vv), u32(self.nnz_vv), np.empty(self.nnz_vv), u32(self.nnz_vp), u32(self.nnz_vp),
               np.empty(self.nnz_vp), np.empty((self.n_p, 2), dtype=np.uint64))
        _lib.call("dq_blocks_copy", self._h, *[a.ctypes.data_as(ctypes.c_void_p) for a in out], 0,
                  stream_handle())
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().dq_blocks_free(h)
            except Exception:
                pass
            self._h = None


def _coo(row, col, val, shape) -> dict:
    return {"row": row, "col": col, "val": val, "shape": (int(shape[0]), int(shape[1]))}


def ham_blocks_device(dets, int_ctx: IntCtx, n_orb: int, mode: int, use_heatbath: bool = False, eps1: float = 1e-6,
                      psi=None, ext=None) -> Blocks:
    d = dets_to_device(dets)
    e = dets_to_device(ext) if ext is not None else None
    p = _f64_to_device(psi, d.shape[0], "psi_v") if psi is not None else None
    h = ctypes.c_void_p()
    _lib.call("dq_ham_blocks", int_ctx.handle, _ptr(d), d.shape[0], _ptr(e), 0 if e is None else e.shape[0],
              int(n_orb), int(mode), int(bool(use_heatbath)), float(eps1), _ptr(p), stream_handle(), ctypes.byref(h))
    return Blocks(h)


def _blocks_dict(B: Blocks, n_v: int, n_p_shape: int, with_p: bool) -> dict:
    vr, vc, vv, pr, pc, pv, dp = B.to_host()
    out = {"H_VV": _coo(vr, vc, vv, (n_v, n_v)), "H_VP": _coo(pr, pc, pv, (n_v if with_p else 0, n_p_shape))}
    if with_p:
        out["det_P"] = dp
        out["size_P"] = int(B.n_p)
    return out


def get_ham_vv(dets_V, int_ctx: IntCtx, n_orb: int) -> dict:
    """H_VV in canonical COO (bridge.cpp:459-467; build_ham.cpp:402-422)."""
    n = np.asarray(dets_V).shape[0]
    B = ham_blocks_device(dets_V, int_ctx, n_orb, 0)
    vr, vc, vv, *_ = B.to_host()
    return _coo(vr, vc, vv, (n, n) if n else (0, 0))


def get_ham_block(bra_dets, ket_dets=None, int_ctx: IntCtx = None, n_orb: int = 0) -> dict:
    """H_VV and H_VP against an explicit P (bridge.cpp:469-507; build_ham.cpp:424-460)."""
    n = np.asarray(bra_dets).shape[0]
    if n == 0:
        return {"H_VV": _coo(*(np.zeros(0, np.uint32),) * 2, np.zeros(0), (0, 0)),
                "H_VP": _coo(*(np.zeros(0, np.uint32),) * 2, np.zeros(0), (0, 0))}
    has_p = ket_dets is not None and np.asarray(ket_dets).shape[0] > 0
    B = ham_blocks_device(bra_dets, int_ctx, n_orb, 0, ext=ket_dets if has_p else None)
    vr, vc, vv, pr, pc, pv, dp = B.to_host()
    out = {"H_VV": _coo(vr, vc, vv, (n, n)),
           "H_VP": _coo(pr, pc, pv, (n, dp.shape[0]) if has_p else (0, 0))}
    if has_p:
        out["det_P"] = dp
        out["size_P"] = int(dp.shape[0])
    return out


def get_ham_conn(dets_V, int_ctx: IntCtx, n_orb: int, use_heatbath: bool = False, eps1: float = 1e-6) -> dict:
    """H_VV, H_VP and the discovered P with static screening (bridge.cpp:509-541)."""
    n = np.asarray(dets_V).shape[0]
    if use_heatbath and not int_ctx.has_hb:
        raise ValueError("Heat-Bath table not initialized")
    if n == 0:
        z = np.zeros(0, np.uint32)
        return {"H_VV": _coo(z, z, np.zeros(0), (0, 0)), "H_VP": _coo(z, z, np.zeros(0), (0, 0)),
                "det_P": np.zeros((0, 2), np.uint64), "size_P": 0}
    B = ham_blocks_device(dets_V, int_ctx, n_orb, 1, use_heatbath, eps1)
    return _blocks_dict(B, n, B.n_p, True)


def get_ham_conn_amp(dets_V, psi_v, int_ctx: IntCtx, n_orb: int, eps1: float = 1e-6) -> dict:
    """Dynamic amplitude-weighted screening (bridge.cpp:547-586; build_ham.cpp:497-535)."""
    if not int_ctx.has_hb:
        raise ValueError("Heat-Bath table not initialized")
    n = np.asarray(dets_V).shape[0]
    if np.asarray(psi_v).shape[0] != n:
        raise ValueError("size mismatch: dets_V and psi_v")
    if n == 0:
        z = np.zeros(0, np.uint32)
        return {"H_VV": _coo(z, z, np.zeros(0), (0, 0)), "H_VP": _coo(z, z, np.zeros(0), (0, 0)),
                "det_P": np.zeros((0, 2), np.uint64), "size_P": 0}
    B = ham_blocks_device(dets_V, int_ctx, n_orb, 2, True, eps1, psi=psi_v)
    return _blocks_dict(B, n, B.n_p, True)


def _coo_arrays(d: dict):
    r = np.ascontiguousarray(d["row"]).astype(np.uint32)
    c = np.ascontiguousarray(d["col"]).astype(np.uint32)
    v = np.ascontiguousarray(d["val"], dtype=np.float64)
    shape = tuple(int(x) for x in d.get("shape", (0, 0)))
    return r, c, v, shape


def get_ham_eff(H_VV: dict, H_VP: dict, h_pp_diag, e_ref: float, reg_type: str = "sigma", epsilon: float = 1e-12,
                upper_only: bool = True) -> dict:
    """H_eff = H_VV + H_VP D^-1 H_VP^T (bridge.cpp:593-629; ham_eff.cpp:64-233)."""
    if reg_type not in ("sigma", "linear_shift"):
        raise ValueError(f"invalid reg_type: {reg_type}")
    vr, vc, vv, vshape = _coo_arrays(H_VV)
    pr, pc, pv, pshape = _coo_arrays(H_VP)
    hpp = np.ascontiguousarray(h_pp_diag, dtype=np.float64)
    n_p = hpp.shape[0]
    if n_p > 0:
        if pshape[1] != n_p:
            raise ValueError("H_VP columns != h_pp_diag size")
        n_v = vshape[0] if vshape[0] else pshape[0]
        if pshape[0] != n_v:
            raise ValueError("H_VP rows != H_VV rows")
    else:
        n_v = vshape[0]
    dev = device()
    t = lambda a: torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to(dev)  # noqa: E731
    dvr, dvc, dvv, dpr, dpc, dpv, dh = (t(vr), t(vc), t(vv), t(pr), t(pc), t(pv), t(hpp))
    h = ctypes.c_void_p()
    _lib.call("dq_ham_eff", _ptr(dvr), _ptr(dvc), _ptr(dvv), vv.size, _ptr(dpr), _ptr(dpc), _ptr(dpv), pv.size,
              n_v, _ptr(dh), n_p, float(e_ref), 1 if reg_type == "sigma" else 0, float(epsilon), stream_handle(),
              ctypes.byref(h))
    B = Blocks(h)
    er, ec, ev, *_ = B.to_host()
    return _coo(er, ec, ev, (n_v, n_v))


def compute_variational_energy_device(dets, coeffs, int_ctx: IntCtx, n_orb: int, use_heatbath: bool = False,
                                      eps1: float = 1e-6, want_sigma: bool = False, method: str = "strings"):
    """``method``: "strings" (default) visits only the determinants of V that can
    connect to a row, through V's alpha/beta string groups; "walk" generates every
    connected determinant and looks it up in V (local_conn.cpp:148-229).  Both
    find the same (y, H_xy) pairs; the sums differ only in rounding order."""
    if method not in _lib.METHODS:
        raise ValueError(f"unknown energy method {method!r}")
    d = dets_to_device(dets)
    c = _f64_to_device(coeffs, d.shape[0], "coefficient")
    sig = torch.empty(d.shape[0], dtype=torch.float64, device=d.device) if want_sigma else None
    e = ctypes.c_double()
    _lib.call("dq_variational_energy_method", int_ctx.handle, _ptr(d), _ptr(c), d.shape[0], int(n_orb),
              int(bool(use_heatbath)), float(eps1), _lib.METHODS[method], _ptr(sig), ctypes.byref(e),
              stream_handle())
    return (e.value, sig) if want_sigma else e.value


def compute_variational_energy(dets, coeffs, int_ctx: IntCtx, n_orb: int, use_heatbath: bool = False,
                               eps1: float = 1e-6, method: str = "strings") -> float:
    """<Psi|H|Psi> numerator on V (bridge.cpp:680-714; local_conn.cpp:148-229)."""
    n = np.asarray(dets).shape[0]
    if np.asarray(coeffs).shape[0] != n:
        raise ValueError("coefficient size mismatch")
    if use_heatbath and not int_ctx.has_hb:
        raise ValueError("compute_variational_energy: Heat-Bath mode requires non-null hb_table")
    if n == 0:
        return 0.0
    return compute_variational_energy_device(dets, coeffs, int_ctx, n_orb, use_heatbath, eps1, method=method)


def trim_memory() -> None:
    """Returns the library's cached device blocks (large temporaries of the
    set / PT2 paths) to the driver."""
    _lib.call("dq_trim_memory")


# ----------------------------------------------------------------------------- determinant batches

def prepare_det_batch_device(dets, ref, n_orb: int, n_alpha: int, n_beta: int, kmax: int = 0,
                             need_k: bool = False, need_phase: bool = False, need_hp: bool = False,
                             need_hp_pos: bool = False) -> dict:
    """Device form of :func:`prepare_det_batch`: a dict of torch tensors in HBM
    (hp_mask as bool)."""
    d = dets_to_device(dets)
    r = np.asarray(ref, dtype=np.uint64).reshape(-1)
    if r.shape != (2,):
        raise ValueError(f"ref must be (2,), got {r.shape}")
    B = d.shape[0]
    need_k = bool(need_k or need_hp)  # bridge.cpp:369
    flags = ((_lib.DQ_PREP_K if need_k else 0) | (_lib.DQ_PREP_PHASE if need_phase else 0)
             | (_lib.DQ_PREP_HP if need_hp else 0) | (_lib.DQ_PREP_HP_POS if need_hp_pos else 0))
    if n_alpha < 0 or n_beta < 0:
        raise ValueError("invalid electron counts")
    K = max(int(kmax), 0)
    dev = d.device
    i32 = dict(dtype=torch.int32, device=dev)
    out = {"occ": torch.empty((B, n_alpha + n_beta), **i32)}
    if need_k:
        out["k"] = torch.empty(B, dtype=torch.int8, device=dev)
    if need_phase:
        out["phase"] = torch.empty(B, dtype=torch.int8, device=dev)
    if need_hp:
        out["holes"] = torch.empty((B, K), **i32)
        out["parts"] = torch.empty((B, K), **i32)
        out["hp_mask"] = torch.empty((B, K), dtype=torch.bool, device=dev)
        if need_hp_pos:
            out["holes_pos"] = torch.empty((B, K), **i32)
            out["parts_pos"] = torch.empty((B, K), **i32)
    g = out.get
    _lib.call("dq_prepare_det_batch", _ptr(d), B, int(r[0]), int(r[1]), int(n_orb), int(n_alpha), int(n_beta),
              int(kmax), flags, _ptr(out["occ"]), _ptr(g("k")), _ptr(g("phase")), _ptr(g("holes")),
              _ptr(g("parts")), _ptr(g("hp_mask")), _ptr(g("holes_pos")), _ptr(g("parts_pos")), stream_handle())
    return out


def prepare_det_batch(dets, ref, n_orb: int, n_alpha: int, n_beta: int, kmax: int = 0, need_k: bool = False,
                      need_phase: bool = False, need_hp: bool = False, need_hp_pos: bool = False) -> dict:
    """Batch features for the neural ansatz (bridge.cpp:347-444; det_batch.cpp:189-319):
    ``occ`` int32 (B, n_alpha + n_beta) and, as requested, ``k`` int8, ``phase``
    int8, ``holes`` / ``parts`` int32 (B, kmax), ``hp_mask`` bool (B, kmax),
    ``holes_pos`` / ``parts_pos`` int32 (B, kmax).  Spin orbitals in block order
    (beta orbital mo is n_orb + mo)."""
    out = prepare_det_batch_device(dets, ref, n_orb, n_alpha, n_beta, kmax, need_k, need_phase, need_hp,
                                   need_hp_pos)
    return {k: v.cpu().numpy() for k, v in out.items()}


def gen_fci_dets_device(n_orb: int, n_alpha: int, n_beta: int) -> torch.Tensor:
    n = ctypes.c_int64()
    _lib.call("dq_fci_size", int(n_orb), int(n_alpha), int(n_beta), ctypes.byref(n))
    out = torch.empty((n.value, 2), dtype=torch.int64, device=device())
    _lib.call("dq_gen_fci_dets", int(n_orb), int(n_alpha), int(n_beta), _ptr(out), n.value, stream_handle())
    return out


def gen_fci_dets(n_orb: int, n_alpha: int, n_beta: int) -> np.ndarray:
    """Full CI space (bridge.cpp:253-260; det_enum.cpp:87-107): beta words outer,
    each spin's words ascending."""
    return dets_to_host(gen_fci_dets_device(n_orb, n_alpha, n_beta))


# ----------------------------------------------------------------------------- connected / perturbative spaces

class DetSet:
    """Device-resident sorted determinant set (dq_detset), optionally with values."""

    def __init__(self, h: ctypes.c_void_p):
        self._h = h
        n, hv = ctypes.c_int64(), ctypes.c_int()
        _lib.call("dq_detset_size", h, ctypes.byref(n), ctypes.byref(hv))
        self.n, self.has_values = n.value, bool(hv.value)

    def to_device(self):
        d = torch.empty((self.n, 2), dtype=torch.int64, device=device())
        v = torch.empty(self.n, dtype=torch.float64, device=device()) if self.has_values else None
        _lib.call("dq_detset_copy", self._h, _ptr(d), _ptr(v), 1, stream_handle())
        return d, v

    def to_host(self):
        d = np.empty((self.n, 2), dtype=np.uint64)
        v = np.empty(self.n, dtype=np.float64) if self.has_values else None
        _lib.call("dq_detset_copy", self._h, d.ctypes.data_as(ctypes.c_void_p),
                  None if v is None else v.ctypes.data_as(ctypes.c_void_p), 0, stream_handle())
        return d, v

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().dq_detset_free(h)
            except Exception:
                pass
            self._h = None


def gen_connected_dets_device(ref_dets, n_orb: int, chunk_items: int = 0) -> DetSet:
    d = dets_to_device(ref_dets)
    h = ctypes.c_void_p()
    _lib.call("dq_gen_connected", _ptr(d), d.shape[0], int(n_orb), int(chunk_items), stream_handle(),
              ctypes.byref(h))
    return DetSet(h)


def gen_connected_dets(ref_dets, n_orb: int, chunk_items: int = 0) -> np.ndarray:
    """Every single and double excitation of ``ref_dets``, sorted unique
    (bridge.cpp:262-271).  ``chunk_items`` bounds the candidates per device pass."""
    return gen_connected_dets_device(ref_dets, n_orb, chunk_items).to_host()[0]


def gen_perturbative_dets(ref_dets, n_orb: int, int_ctx: Optional[IntCtx] = None, psi_v=None, mode: str = "none",
                          eps1: float = 1e-6, chunk_items: int = 0) -> np.ndarray:
    """P = C \\ V, sorted (bridge.cpp:278-338): mode 'none' (all singles and
    doubles), 'static' (heat-bath eps1), 'dynamic' (amplitude-weighted)."""
    return gen_perturbative_dets_device(ref_dets, n_orb, int_ctx, psi_v, mode, eps1, chunk_items).to_host()[0]


def gen_perturbative_dets_device(ref_dets, n_orb: int, int_ctx: Optional[IntCtx] = None, psi_v=None,
                                 mode: str = "none", eps1: float = 1e-6, chunk_items: int = 0) -> DetSet:
    """Device form of :func:`gen_perturbative_dets`: the set stays in HBM."""
    d = dets_to_device(ref_dets)
    n = d.shape[0]
    if mode not in _lib.SCREEN_MODES:
        # the reference checks V's duplicates first (bridge.cpp:285-290), then the mode
        _check_unique(d)
        raise ValueError(f"invalid screen mode: {mode}")
    m = _lib.SCREEN_MODES[mode]
    psi = None
    if m != 0:
        if int_ctx is None:
            raise ValueError("gen_perturbative_dets: int_ctx is required for screened modes")
        if not int_ctx.has_hb:
            _check_unique(d)
            raise ValueError("Heat-Bath table not initialized")
        if m == 2:
            if psi_v is None:
                raise ValueError("psi_v size mismatch")
            psi = _f64_to_device(psi_v, n, "psi_v")
    h = ctypes.c_void_p()
    _lib.call("dq_gen_perturbative", int_ctx.handle if int_ctx is not None else None, _ptr(d), n, int(n_orb),
              _ptr(psi), m, float(eps1), int(chunk_items), stream_handle(), ctypes.byref(h))
    return DetSet(h)


def _check_unique(d: torch.Tensor) -> None:
    if d.shape[0] and torch.unique(d, dim=0).shape[0] != d.shape[0]:
        raise ValueError("DetMap::from_ordered: duplicate determinant found")


def compute_pt2_device(dets_V, coeffs_V, int_ctx: IntCtx, n_orb: int, e_ref: float, use_heatbath: bool = False,
                       eps1: float = 1e-6, chunk_items: int = 0, want_external: bool = False):
    """EN-PT2 on the device; with ``want_external`` also the external set as
    (dets, v_a) device tensors, v_a = sum_i H_ai c_i."""
    d = dets_to_device(dets_V)
    c = _f64_to_device(coeffs_V, d.shape[0], "coefficient")
    ei, ee, ne = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    h = ctypes.c_void_p()
    _lib.call("dq_compute_pt2", int_ctx.handle, _ptr(d), _ptr(c), d.shape[0], int(n_orb), float(e_ref),
              int(bool(use_heatbath)), float(eps1), int(chunk_items), ctypes.byref(ei), ctypes.byref(ee),
              ctypes.byref(ne), stream_handle(), ctypes.byref(h) if want_external else None)
    out = {"e_pt2_internal": ei.value, "e_pt2_external": ee.value, "n_ext": ne.value}
    if want_external:
        out["external"] = DetSet(h).to_device()
    return out


def compute_pt2(dets_V, coeffs_V, int_ctx: IntCtx, n_orb: int, e_ref: float, use_heatbath: bool = False,
                eps1: float = 1e-6) -> dict:
    """Decomposed EN-PT2 (bridge.cpp:721-771; local_conn.cpp:247-418):
    {'e_pt2_internal', 'e_pt2_external', 'n_ext'}."""
    n = np.asarray(dets_V).shape[0] if not isinstance(dets_V, torch.Tensor) else dets_V.shape[0]
    if (coeffs_V.shape[0] if isinstance(coeffs_V, torch.Tensor) else np.asarray(coeffs_V).shape[0]) != n:
        raise ValueError("coefficient size mismatch")
    if use_heatbath and not int_ctx.has_hb:
        raise ValueError("Heat-Bath table not initialized")
    return compute_pt2_device(dets_V, coeffs_V, int_ctx, n_orb, e_ref, use_heatbath, eps1)


__all__ = [
    "DetSet",
    "trim_memory",
    "prepare_det_batch",
    "prepare_det_batch_device",
    "gen_fci_dets",
    "gen_fci_dets_device",
    "gen_connected_dets",
    "gen_connected_dets_device",
    "gen_perturbative_dets",
    "gen_perturbative_dets_device",
    "compute_pt2",
    "compute_pt2_device",
    "Blocks",
    "ham_blocks_device",
    "get_ham_vv",
    "get_ham_block",
    "get_ham_conn",
    "get_ham_conn_amp",
    "get_ham_eff",
    "compute_variational_energy",
    "compute_variational_energy_device",
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
