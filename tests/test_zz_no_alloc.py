"""GPU: the allocation-free energy entry point (dq_variational_energy_noalloc)
against the allocating one (dq_variational_energy_method, string method),
called through the C-ABI: sigma and energy bit-identical, on two calls in a
row (the second smaller, so the shared scratch array holds the first call's
leftovers), with and without a caller sigma, plus the size limit."""

import ctypes

import numpy as np
import pytest

from helpers import write_tmp_fcidump

pytestmark = pytest.mark.gpu


def _alloc_path(L, ctx, d, c, n_orb, use_hb, eps):
    import torch
    n = d.shape[0]
    sig = torch.empty(n, dtype=torch.float64, device="cuda")
    e = ctypes.c_double()
    rc = L.dq_variational_energy_method(ctx.handle, d.data_ptr(), c.data_ptr(), n, n_orb, int(use_hb), eps, 0,
                                        sig.data_ptr(), ctypes.byref(e), None)
    assert rc == 0, L.dq_last_error()
    return e.value, sig.cpu().numpy()


def _noalloc_path(L, ctx, d, c, n_orb, use_hb, eps, want_sigma):
    import torch
    n = d.shape[0]
    sig = torch.empty(n, dtype=torch.float64, device="cuda") if want_sigma else None
    e = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    rc = L.dq_variational_energy_noalloc(ctx.handle, d.data_ptr(), c.data_ptr(), n, n_orb, int(use_hb), eps,
                                         sig.data_ptr() if sig is not None else None, e.data_ptr(), None)
    assert rc == 0, L.dq_last_error()
    torch.cuda.synchronize()
    return float(e.cpu().numpy()[0]), (sig.cpu().numpy() if sig is not None else None)


def test_noalloc_matches_allocating_entry_point(tmp_path):
    import torch
    from detnqs_amd import _lib, core, synth
    L = _lib.load()
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    dets = synth.config1_dets(60000, seed=1)
    rng = np.random.default_rng(12)
    for n, use_hb, eps, want_sigma in ((60000, True, 1e-7, True), (9000, True, 1e-5, False),
                                       (9000, False, 0.0, True), (1, True, 1e-7, True)):
        d = core.dets_to_device(dets[:n])
        c = torch.from_numpy(rng.standard_normal(n)).cuda()
        e_ref, s_ref = _alloc_path(L, ctx, d, c, 28, use_hb, eps)
        e_got, s_got = _noalloc_path(L, ctx, d, c, 28, use_hb, eps, want_sigma)
        print(f"noalloc n={n} hb={use_hb} eps={eps}: {e_got!r} vs {e_ref!r}", flush=True)
        assert np.float64(e_got).tobytes() == np.float64(e_ref).tobytes()
        if want_sigma:
            assert s_got.tobytes() == s_ref.tobytes()


def test_noalloc_limits_and_empty():
    import torch
    from detnqs_amd import _lib, core, synth
    L = _lib.load()
    mx, by = ctypes.c_int64(), ctypes.c_int64()
    assert L.dq_noalloc_limits(ctypes.byref(mx), ctypes.byref(by)) == 0
    assert mx.value == 262144 and by.value == 64 << 20
    ints = synth.make_integrals(13, 10, 40, seed=3)
    ctx = core.IntCtx.from_integrals(13, 10, ints.h1e, ints.h2e, ints.e_nuc)
    d = torch.zeros((mx.value + 1, 2), dtype=torch.int64, device="cuda")
    c = torch.zeros(mx.value + 1, dtype=torch.float64, device="cuda")
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = L.dq_variational_energy_noalloc(ctx.handle, d.data_ptr(), c.data_ptr(), mx.value + 1, 13, 0, 0.0, None,
                                         e.data_ptr(), None)
    assert rc == _lib.DQ_ERANGE and b"allocation-free maximum" in L.dq_last_error()
    e.fill_(5.0)
    rc = L.dq_variational_energy_noalloc(ctx.handle, d.data_ptr(), c.data_ptr(), 0, 13, 0, 0.0, None,
                                         e.data_ptr(), None)
    assert rc == 0
    torch.cuda.synchronize()
    assert float(e.cpu()[0]) == 0.0
