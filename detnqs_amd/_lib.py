"""Loader for libdetnqs_hip.so (the C-ABI declared in include/detnqs_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C detnqs_amd/csrc``).  There is no fallback: if it is missing, or no
HIP device is visible, calls raise.
"""

from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdetnqs_hip.so")

DQ_OK, DQ_EINVAL, DQ_ERANGE, DQ_ERUNTIME, DQ_ETIMEDOUT = 0, -1, -2, -3, -4
DQ_ELOC_STRINGS, DQ_ELOC_WALK = 0, 1
METHODS = {"strings": DQ_ELOC_STRINGS, "walk": DQ_ELOC_WALK}

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double



class ElocSegment(ctypes.Structure):
    """dq_eloc_segment: one row range of one instance."""
    _fields_ = [
        ("ctx", _vp), ("vmap", _vp), ("d_dets", _vp), ("d_coeffs", _vp), ("d_sigma", _vp),
        ("n_rows", _i64), ("row_begin", _i64), ("row_end", _i64),
        ("n_orb", _int), ("use_hb", _int), ("eps1", _dbl), ("energy_slot", _int), ("method", _int),
        ("vv", _vp),
    ]


class ElocTuning(ctypes.Structure):
    """dq_eloc_tuning: launch shape of a multi-instance energy plan."""
    _fields_ = [("waves_per_block", _int), ("persistent", _int), ("grid_blocks", _int)]


class ServerConfig(ctypes.Structure):
    """dq_server_config."""
    _fields_ = [("hbm_budget_bytes", _i64), ("rows_per_launch", _i64), ("max_inflight", _int),
                ("tune", ElocTuning), ("vv_layout", _int)]


class ServerRequest(ctypes.Structure):
    """dq_server_request."""
    _fields_ = [("ctx", _vp), ("dets", _vp), ("dets_on_device", _int), ("coeffs", _vp),
                ("coeffs_on_device", _int), ("n_rows", _i64), ("n_orb", _int), ("use_hb", _int),
                ("eps1", _dbl), ("v_key", ctypes.c_uint64), ("stream", _vp), ("method", _int)]


class ServerResult(ctypes.Structure):
    """dq_server_result."""
    _fields_ = [("energy", _dbl), ("status", _int), ("t_submit", _dbl), ("t_admit", _dbl),
                ("t_first_launch", _dbl), ("t_done", _dbl)]


class ServerStats(ctypes.Structure):
    """dq_server_stats."""
    _fields_ = [(k, _i64) for k in ("submitted", "completed", "launches", "rows", "v_hits", "v_misses",
                                     "v_evictions", "v_cached", "arena_bytes", "arena_used", "arena_peak",
                                     "queued", "running", "vv_builds", "vv_rows")]


# name -> argtypes (all return int status)
_SIGS = {
    "dq_version": [],
    "dq_trim_memory": [],
    "dq_ctx_create_fcidump": [ctypes.c_char_p, _int, ctypes.POINTER(_vp)],
    "dq_ctx_create": [_int, _int, _vp, _vp, _dbl, ctypes.POINTER(_vp)],
    "dq_ctx_destroy": [_vp],
    "dq_ctx_info": [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_dbl)],
    "dq_hb_prepare": [_vp, _dbl],
    "dq_hb_clear": [_vp],
    "dq_hb_export": [_vp, ctypes.POINTER(_i64), _vp, _vp, _vp, _vp],
    "dq_ham_diag": [_vp, _vp, _i64, _vp, _vp],
    "dq_local_conn_count": [_vp, _vp, _i64, _int, _int, _dbl, _vp, ctypes.POINTER(_i64), _vp],
    "dq_local_conn_fill": [_vp, _vp, _i64, _int, _int, _dbl, _vp, _vp, _vp, _vp],
    "dq_ham_blocks": [_vp, _vp, _i64, _vp, _i64, _int, _int, _int, _dbl, _vp, _vp, ctypes.POINTER(_vp)],
    "dq_blocks_sizes": [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "dq_blocks_copy": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp],
    "dq_blocks_free": [_vp],
    "dq_ham_eff": [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _i64, _vp, _i64, _dbl, _int, _dbl, _vp,
                   ctypes.POINTER(_vp)],
    "dq_variational_energy": [_vp, _vp, _vp, _i64, _int, _int, _dbl, _vp, ctypes.POINTER(_dbl), _vp],
    "dq_variational_energy_method": [_vp, _vp, _vp, _i64, _int, _int, _dbl, _int, _vp, ctypes.POINTER(_dbl), _vp],
    "dq_variational_energy_noalloc": [_vp, _vp, _vp, _i64, _int, _int, _dbl, _vp, _vp, _vp],
    "dq_noalloc_limits": [ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "dq_detmap_build": [_vp, _i64, _int, _vp, ctypes.POINTER(_vp)],
    "dq_vvcache_build": [_vp, _vp, _vp, _i64, _int, _int, _dbl, _vp, ctypes.POINTER(_vp)],
    "dq_vvcache_info": [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "dq_vvcache_export": [_vp, _vp, _vp, _vp, _vp],
    "dq_vvcache_energy": [_vp, _vp, _vp, ctypes.POINTER(_dbl), _vp],
    "dq_vvcache_free": [_vp],
    "dq_detmap_free": [_vp],
    "dq_eloc_plan_create": [ctypes.POINTER(ElocSegment), _int, ctypes.POINTER(ElocTuning), ctypes.POINTER(_vp)],
    "dq_eloc_plan_info": [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_int), ctypes.POINTER(_int),
                          ctypes.POINTER(ElocTuning)],
    "dq_eloc_plan_launch": [_vp, _vp, _vp],
    "dq_eloc_plan_capture": [_vp, _vp],
    "dq_eloc_plan_graph_launch": [_vp, _vp],
    "dq_eloc_plan_check": [_vp, _vp],
    "dq_eloc_plan_free": [_vp],
    "dq_server_create": [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)],
    "dq_server_request_bytes": [_i64, _int, ctypes.POINTER(_i64)],
    "dq_server_submit": [_vp, ctypes.POINTER(ServerRequest), ctypes.POINTER(_i64)],
    "dq_server_wait": [_vp, _i64, _dbl, ctypes.POINTER(ServerResult)],
    "dq_server_stats_get": [_vp, ctypes.POINTER(ServerStats)],
    "dq_server_destroy": [_vp],
    "dq_operator_create": [_i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, ctypes.POINTER(_vp)],
    "dq_operator_apply": [_vp, _vp, _vp, _vp, _vp, _vp],
    "dq_energy_step": [_vp, _int, _vp, _vp, _dbl, _vp, _vp, _vp, ctypes.POINTER(_dbl), _vp],
    "dq_operator_free": [_vp],
    "dq_slogdet_fwd_f64": [_vp, _i64, _int, _vp, _vp, _vp],
    "dq_slogdet_bwd_f64": [_vp, _vp, _vp, _i64, _int, _vp, _vp],
    "dq_slogdet_fwd_lu_f64": [_vp, _i64, _int, _vp, _vp, _vp, _vp, _vp],
    "dq_slogdet_fwd_f32": [_vp, _i64, _int, _vp, _vp, _vp],
    "dq_slogdet_bwd_f32": [_vp, _vp, _vp, _i64, _int, _vp, _vp],
    "dq_slogdet_bwd_lu_f64": [_vp, _vp, _vp, _i64, _int, _vp, _vp],
    "dq_slogdet_tile_choices": [_int, _vp, _int, _vp],
    "dq_slogdet_tile_set": [_int, _int, _int],
    "dq_slogdet_tile_get": [_int, _int, _vp],
    "dq_prepare_det_batch": [_vp, _i64, ctypes.c_uint64, ctypes.c_uint64, _int, _int, _int, _int, _int,
                             _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "dq_fci_size": [_int, _int, _int, ctypes.POINTER(_i64)],
    "dq_gen_fci_dets": [_int, _int, _int, _vp, _i64, _vp],
    "dq_gen_connected": [_vp, _i64, _int, _i64, _vp, ctypes.POINTER(_vp)],
    "dq_gen_perturbative": [_vp, _vp, _i64, _int, _vp, _int, _dbl, _i64, _vp, ctypes.POINTER(_vp)],
    "dq_compute_pt2": [_vp, _vp, _vp, _i64, _int, _dbl, _int, _dbl, _i64, ctypes.POINTER(_dbl),
                       ctypes.POINTER(_dbl), ctypes.POINTER(_i64), _vp, ctypes.POINTER(_vp)],
    "dq_detset_size": [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_int)],
    "dq_detset_copy": [_vp, _vp, _vp, _int, _vp],
    "dq_detset_free": [_vp],
}

DQ_PREP_K, DQ_PREP_PHASE, DQ_PREP_HP, DQ_PREP_HP_POS = 1, 2, 4, 8
SCREEN_MODES = {"none": 0, "static": 1, "dynamic": 2}

_lock = threading.Lock()
_lib = None


def exported_symbols():
    return list(_SIGS.keys()) + ["dq_last_error"]


def load() -> ctypes.CDLL:
    """Load the library once (raises ImportError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libdetnqs_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.dq_last_error.restype = ctypes.c_char_p
        L.dq_last_error.argtypes = []
        for name, args in _SIGS.items():
            if not hasattr(L, name):
                continue  # reported by call() when used
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
        return L


class DQError(RuntimeError):
    pass


def check(rc: int) -> None:
    """Map a C-ABI status to the exception the reference bridge raises."""
    if rc == DQ_OK:
        return
    msg = load().dq_last_error().decode(errors="replace")
    if rc == DQ_EINVAL:
        raise ValueError(msg)
    if rc == DQ_ERANGE:
        raise IndexError(msg)
    if rc == DQ_ETIMEDOUT:
        raise TimeoutError(msg)
    raise DQError(msg)


def call(name: str, *args) -> None:
    L = load()
    if not hasattr(L, name):
        raise DQError(f"{name} is not exported by {LIB_PATH}")
    check(getattr(L, name)(*args))
