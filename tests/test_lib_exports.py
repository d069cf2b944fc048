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
