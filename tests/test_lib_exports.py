"""CPU: libdetnqs_hip.so loads (no GPU needed) and exports every symbol that
include/detnqs_hip.h declares."""

import ctypes
import os
import re

from helpers import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "detnqs_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dq_\w+)\s*\(", src, flags=re.M)))


def test_library_exports_header():
    from detnqs_amd import _lib
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.dq_version() >= 1
    assert set(names) <= set(_lib.exported_symbols())


def test_error_text_without_gpu():
    from detnqs_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.dq_ctx_create(0, 0, None, None, ctypes.c_double(0.0), ctypes.byref(h))
    assert rc == _lib.DQ_EINVAL
    assert b"null" in L.dq_last_error() or b"orbitals" in L.dq_last_error()


def test_slogdet_tile_table_without_gpu():
    """The tile table is host-only: choices per size bucket, set/get, and the
    errors for sizes without tiles and unbuilt choices."""
    from detnqs_amd import slogdet as S
    rt = S.TILE_RUNTIME_N
    assert S.tile_choices(12) == [4, 8, 16, 4 | rt, 8 | rt, 16 | rt]
    assert S.tile_choices(11) == [4, 8, 16] and S.tile_choices(13) == [4, 8, 16]
    assert S.tile_choices(16) == [4, 8, 16, 4 | rt, 8 | rt, 16 | rt]
    assert S.tile_choices(24) == [8, 16, 32, 8 | rt, 16 | rt, 32 | rt]
    assert S.tile_choices(32) == [16, 32, 16 | rt, 32 | rt]
    assert S.tile_choices(33) == [64] and S.tile_choices(64) == [64]
    for n in (9, 12, 13, 16, 20, 24, 31, 32, 40):
        assert all(S.get_tile(k, n) in S.tile_choices(n) for k in range(4)), n
    d = S.get_tile("forward_lu", 16)
    assert d in S.tile_choices(16)
    S.set_tile("forward_lu", 16, 16)
    assert S.get_tile("forward_lu", 16) == 16 and S.get_tile("backward_lu", 16) in S.tile_choices(16)
    S.set_tile(1, 16, 0)
    assert S.get_tile(1, 16) == d
    import pytest
    for n, lanes in ((16, 32), (24, 4), (40, 16), (13, 8 | rt), (64, 64 | rt)):
        with pytest.raises(ValueError):
            S.set_tile("forward", n, lanes)
    for n in (8, 65):
        with pytest.raises(ValueError):
            S.get_tile(0, n)
    with pytest.raises(ValueError):
        S.set_tile(4, 16, 8)
    with pytest.raises(ValueError):
        S.get_tile("sideways", 16)
