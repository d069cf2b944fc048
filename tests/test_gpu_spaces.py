"""GPU parity for the rows next to the local-energy path: prepare_det_batch,
the FCI space, the connected / perturbative spaces and EN-PT2, against the
oracle on the same seeded inputs.

Bars: features, sets and counts bit-exact; EN-PT2 amplitudes v_a bit-exact
when the candidates fit one pass (the device then adds each v_a in row
order, as the oracle and the reference's one-thread run do), and within
1e-13 relative over several passes; the internal and external energies
(reductions in another order) within 1e-12 relative."""

import tempfile

import numpy as np
import pytest

from helpers import H2O_STO3G, fci_dets, write_tmp_fcidump

pytestmark = pytest.mark.gpu


def _random_batch(rng, n_orb, na, nb, B):
    a = np.argsort(rng.random((B, n_orb)), axis=1)
    out = np.zeros((B, 2), dtype=np.uint64)
    one = np.uint64(1)
    for j in range(na):
        out[:, 0] |= one << a[:, j].astype(np.uint64)
    b = np.argsort(rng.random((B, n_orb)), axis=1)
    for j in range(nb):
        out[:, 1] |= one << b[:, j].astype(np.uint64)
    return out


def _ctx_pair(ints, n):
    from detnqs_amd import core
    from oracle.oracle import Lib
    with tempfile.TemporaryDirectory() as td:
        p = write_tmp_fcidump(td, ints)
        ctx = core.IntCtx(p, n)
        octx = Lib("oracle").ctx(p, n)
    ctx.hb_prepare(1e-15)
    octx.hb_prepare(1e-15)
    return ctx, octx


# ----------------------------------------------------------------------------- prepare_det_batch

@pytest.mark.parametrize("n,na,nb,kmax", [(13, 5, 5, 4), (28, 7, 7, 6), (64, 20, 17, 9), (40, 0, 3, 2),
                                          (60, 12, 12, 30), (28, 7, 7, 50), (64, 32, 32, 40), (5, 0, 0, 3)])
def test_prepare_det_batch_matches_oracle(n, na, nb, kmax):
    from detnqs_amd import core
    from oracle.oracle import Lib
    O = Lib("oracle")
    rng = np.random.default_rng(n)
    D = _random_batch(rng, n, na, nb, 3001)
    ref = (sum(1 << i for i in range(na)), sum(1 << i for i in range(nb)))
    for flags in ((False, False, False, False), (True, True, False, False), (False, True, True, False),
                  (True, True, True, True)):
        got = core.prepare_det_batch(D, ref, n, na, nb, kmax, *flags)
        exp = O.prepare_det_batch(D, ref, n, na, nb, kmax, *flags)
        assert got.keys() == exp.keys()
        for k in exp:
            assert got[k].dtype == exp[k].dtype, k
            assert np.array_equal(got[k], exp[k]), (flags, k)


def test_prepare_det_batch_large_and_empty():
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    D = synth.config1_dets(300_000, seed=1)
    ref = (0x7F, 0x7F)
    got = core.prepare_det_batch(D, ref, 28, 7, 7, 8, False, True, True, True)
    exp = Lib("oracle").prepare_det_batch(D, ref, 28, 7, 7, 8, False, True, True, True)
    for k in exp:
        assert np.array_equal(got[k], exp[k]), k
    empty = core.prepare_det_batch(np.zeros((0, 2), dtype=np.uint64), ref, 28, 7, 7, 3, True, True, True, True)
    assert empty["occ"].shape == (0, 14) and empty["holes_pos"].shape == (0, 3)


def test_prepare_det_batch_unaligned_outputs():
    """Output arrays at offsets that rule out 16-byte stores give the same rows."""
    import ctypes
    import torch
    from detnqs_amd import _lib, core
    rng = np.random.default_rng(3)
    B, n, na, nb, K = 1001, 28, 7, 7, 6
    D = core.dets_to_device(_random_batch(rng, n, na, nb, B))
    ref = (0x7F, 0x7F)
    exp = core.prepare_det_batch_device(D, ref, n, na, nb, K, True, True, True, True)
    dev = D.device
    raw32 = [torch.full((B * w + 3,), -7, dtype=torch.int32, device=dev) for w in (na + nb, K, K, K, K)]
    raw8 = [torch.zeros(B + 3, dtype=torch.int8, device=dev) for _ in range(2)] + \
        [torch.zeros(B * K + 3, dtype=torch.uint8, device=dev)]
    off = lambda t, k: ctypes.c_void_p(t.data_ptr() + k * t.element_size())  # noqa: E731
    _lib.call("dq_prepare_det_batch", ctypes.c_void_p(D.data_ptr()), B, ref[0], ref[1], n, na, nb, K, 15,
              off(raw32[0], 1), off(raw8[0], 1), off(raw8[1], 1), off(raw32[1], 1), off(raw32[2], 1),
              off(raw8[2], 1), off(raw32[3], 1), off(raw32[4], 1), core.stream_handle())
    names = ("occ", "holes", "parts", "holes_pos", "parts_pos")
    for t, k in zip(raw32, names):
        assert torch.equal(t[1:1 + exp[k].numel()].view(exp[k].shape), exp[k]), k
        assert int(t[0]) == -7 and int(t[1 + exp[k].numel()]) == -7, k
    assert torch.equal(raw8[0][1:1 + B], exp["k"]) and torch.equal(raw8[1][1:1 + B], exp["phase"])
    assert torch.equal(raw8[2][1:1 + B * K].view(B, K).bool(), exp["hp_mask"])


def test_prepare_det_batch_errors_match_oracle():
    from detnqs_amd import core
    from oracle.oracle import Lib
    O = Lib("oracle")
    D = np.array([[0b11, 0b1], [0b111, 0b1], [0b11 | (1 << 9), 0b1]] * 100, dtype=np.uint64)
    cases = [((0b11, 0b1), 8, 2, 1, 0, False, False, False, False),
             ((0b11, 0b1), 8, 3, 1, 0, False, False, False, False),
             ((0b11, 0b10), 8, 2, 1, 2, False, True, True, True),
             ((0b111, 0b1), 8, 2, 1, 2, False, False, True, True),
             ((0b111, 0b1), 8, 2, 1, 0, False, True, False, False),
             ((0b11, 0b1), 8, 2, 1, 2, False, False, False, True)]
    for args in cases:
        with pytest.raises(ValueError) as eg:
            core.prepare_det_batch(D, *args)
        with pytest.raises(ValueError) as eo:
            O.prepare_det_batch(D, *args)
        assert str(eg.value) == str(eo.value), args


# ----------------------------------------------------------------------------- FCI space

@pytest.mark.parametrize("n,na,nb", [(1, 1, 0), (7, 5, 5), (13, 5, 4), (14, 7, 7), (64, 64, 0), (64, 1, 2)])
def test_gen_fci_matches_oracle(n, na, nb):
    from detnqs_amd import core
    from oracle.oracle import Lib
    got = core.gen_fci_dets(n, na, nb)
    exp = Lib("oracle").gen_fci(n, na, nb)
    assert np.array_equal(got, exp)


def test_gen_fci_errors():
    from detnqs_amd import core
    with pytest.raises(ValueError, match="out of range"):
        core.gen_fci_dets(0, 0, 0)
    with pytest.raises(ValueError, match="invalid electrons"):
        core.gen_fci_dets(6, 7, 1)


# ----------------------------------------------------------------------------- connected / perturbative spaces

def test_gen_connected_matches_oracle():
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    O = Lib("oracle")
    _, dets = synth.config0()
    for V in (dets[:1], dets[:500], dets[::7]):
        exp = O.gen_connected(V, 13)
        assert np.array_equal(core.gen_connected_dets(V, 13), exp)
        assert np.array_equal(core.gen_connected_dets(V, 13, chunk_items=5000), exp)
    assert core.gen_connected_dets(np.zeros((0, 2), dtype=np.uint64), 13).shape == (0, 2)


def test_gen_connected_wide_words():
    """n_orb > 32: the 128-bit sort path."""
    from detnqs_amd import core
    from oracle.oracle import Lib
    rng = np.random.default_rng(7)
    for n, na, nb in ((36, 4, 3), (64, 3, 2), (33, 2, 2)):
        V = np.unique(_random_batch(rng, n, na, nb, 40), axis=0)
        exp = Lib("oracle").gen_connected(V, n)
        assert np.array_equal(core.gen_connected_dets(V, n), exp)
        assert np.array_equal(core.gen_connected_dets(V, n, chunk_items=20000), exp)


def test_gen_perturbative_matches_oracle():
    from detnqs_amd import core, synth
    ints, dets = synth.config0()
    ctx, octx = _ctx_pair(ints, 13)
    rng = np.random.default_rng(1)
    V = dets[rng.permutation(dets.shape[0])[:700]]
    psi = rng.standard_normal(700) * 0.05
    for mode, eps in (("none", 0.0), ("static", 1e-6), ("static", 1e-2), ("static", 0.0), ("dynamic", 1e-5),
                      ("dynamic", 1e-3)):
        exp = octx.gen_perturbative(V, 13, psi=psi, mode=mode, eps1=eps)
        got = core.gen_perturbative_dets(V, 13, ctx, psi, mode, eps)
        assert np.array_equal(got, exp), (mode, eps)
        got2 = core.gen_perturbative_dets(V, 13, ctx, psi, mode, eps, chunk_items=30000)
        assert np.array_equal(got2, exp), (mode, eps, "chunked")


def test_candidate_keys_one_and_many_sectors():
    """Candidate sets of V in one (n_alpha, n_beta) sector sort colex-rank keys
    (fewer radix passes); V over several sectors sorts the packed strings.
    Both give the oracle's sets and PT2 terms, also chunked.  The electron
    counts also pick the external-term kernel: at most 8 per spin (13 and 24
    orbitals), at most 16 (20 orbitals, 10a 9b), more (32 orbitals, 17a 3b)."""
    from detnqs_amd import core, synth
    ints, dets = synth.config0()
    ctx, octx = _ctx_pair(ints, 13)
    rng = np.random.default_rng(21)
    mixed = np.concatenate([dets[:300], _random_batch(rng, 13, 4, 6, 150), _random_batch(rng, 13, 6, 5, 150)])
    mixed = mixed[np.sort(np.unique(mixed, axis=0, return_index=True)[1])]
    ints24 = synth.make_integrals(24, 12, 30, seed=24)
    ctx24, octx24 = _ctx_pair(ints24, 24)
    one24 = np.unique(_random_batch(rng, 24, 6, 6, 120), axis=0)
    ints20 = synth.make_integrals(20, 19, 30, seed=20)
    ctx20, octx20 = _ctx_pair(ints20, 20)
    one20 = np.unique(_random_batch(rng, 20, 10, 9, 40), axis=0)
    ints32 = synth.make_integrals(32, 20, 30, seed=32)
    ctx32, octx32 = _ctx_pair(ints32, 32)
    one32 = np.unique(_random_batch(rng, 32, 17, 3, 20), axis=0)
    for V, c, o, n, chunk in ((mixed, ctx, octx, 13, 5000), (one24, ctx24, octx24, 24, 300_000),
                              (one20, ctx20, octx20, 20, 100_000), (one32, ctx32, octx32, 32, 200_000)):
        psi = rng.standard_normal(V.shape[0]) * 0.05
        for mode, eps in (("static", 1e-5), ("dynamic", 1e-4)):
            exp = o.gen_perturbative(V, n, psi=psi, mode=mode, eps1=eps)
            assert np.array_equal(core.gen_perturbative_dets(V, n, c, psi, mode, eps), exp), (n, mode)
            assert np.array_equal(core.gen_perturbative_dets(V, n, c, psi, mode, eps, chunk_items=chunk), exp), \
                (n, mode, "chunked")
        coef = rng.standard_normal(V.shape[0])
        coef /= np.linalg.norm(coef)
        E = o.variational_energy(V, coef, n) - 0.1
        exp = o.compute_pt2(V, coef, n, E, True, 1e-5)
        _pt2_check(core.compute_pt2_device(V, coef, c, n, E, True, 1e-5, want_external=True), exp, exact_vals=True)
        _pt2_check(core.compute_pt2_device(V, coef, c, n, E, True, 1e-5, chunk_items=chunk, want_external=True),
                   exp, exact_vals=False)


def test_gen_perturbative_wide_and_errors():
    from detnqs_amd import core, synth
    ints = synth.make_integrals(36, 8, 10, seed=77)
    ctx, octx = _ctx_pair(ints, 36)
    rng = np.random.default_rng(3)
    V = np.unique(_random_batch(rng, 36, 4, 4, 30), axis=0)
    psi = rng.standard_normal(V.shape[0])
    for mode, eps in (("none", 0.0), ("static", 1e-4), ("dynamic", 1e-4)):
        assert np.array_equal(core.gen_perturbative_dets(V, 36, ctx, psi, mode, eps),
                              octx.gen_perturbative(V, 36, psi=psi, mode=mode, eps1=eps)), mode
    with pytest.raises(ValueError, match="duplicate"):
        core.gen_perturbative_dets(np.concatenate([V, V[:1]]), 36, ctx, None, "static", 1e-4)
    with pytest.raises(ValueError, match="invalid screen mode"):
        core.gen_perturbative_dets(V, 36, ctx, None, "fancy", 1e-4)
    ctx.hb_clear()
    with pytest.raises(ValueError, match="Heat-Bath table not initialized"):
        core.gen_perturbative_dets(V, 36, ctx, None, "static", 1e-4)
    with pytest.raises(ValueError, match="orbitals at or above"):
        core.gen_perturbative_dets(np.array([[1 << 40, 1]], dtype=np.uint64), 36, None, None, "none")


# ----------------------------------------------------------------------------- EN-PT2

def _pt2_check(got, exp, exact_vals):
    assert got["n_ext"] == exp["n_ext"]
    gd, gv = got["external"]
    gd = gd.cpu().numpy().view(np.uint64)
    gv = gv.cpu().numpy()
    assert np.array_equal(gd, exp["ext_dets"])
    if exact_vals:
        assert gv.tobytes() == exp["ext_vals"].tobytes()
    else:
        assert np.allclose(gv, exp["ext_vals"], rtol=1e-13, atol=1e-16)
    for k in ("e_pt2_internal", "e_pt2_external"):
        assert abs(got[k] - exp[k]) <= 1e-12 * abs(exp[k]) + 1e-300, k


def test_pt2_matches_oracle_config0():
    from detnqs_amd import core, synth
    ints, dets = synth.config0()
    ctx, octx = _ctx_pair(ints, 13)
    rng = np.random.default_rng(11)
    V = dets[:2000]
    coef = rng.standard_normal(2000) * np.exp(-np.arange(2000) / 300.0)
    coef[5::97] = 1e-9
    coef /= np.linalg.norm(coef)
    E = octx.variational_energy(V, coef, 13) - 0.05
    for hbm, eps in ((True, 1e-6), (True, 1e-4), (False, 1e-5), (False, 0.0)):
        exp = octx.compute_pt2(V, coef, 13, E, hbm, eps)
        got = core.compute_pt2_device(V, coef, ctx, 13, E, hbm, eps, want_external=True)
        _pt2_check(got, exp, exact_vals=True)
        got = core.compute_pt2_device(V, coef, ctx, 13, E, hbm, eps, chunk_items=100_000, want_external=True)
        _pt2_check(got, exp, exact_vals=False)
    plain = core.compute_pt2(V, coef, ctx, 13, E, True, 1e-6)
    assert set(plain) == {"e_pt2_internal", "e_pt2_external", "n_ext"}


def test_pt2_eigenvector_and_duplicates():
    from detnqs_amd import core
    from oracle.oracle import Lib
    ctx = core.IntCtx(H2O_STO3G, 7)
    octx = Lib("oracle").ctx(H2O_STO3G, 7)
    F = fci_dets(7, 5, 5)
    vv = octx.ham_vv(F, 7)
    H = np.zeros((F.shape[0],) * 2)
    H[vv["row"], vv["col"]] = vv["val"]
    w, U = np.linalg.eigh(H)
    got = core.compute_pt2(F, U[:, 0], ctx, 7, float(w[0]), False, 0.0)
    assert abs(got["e_pt2_internal"]) < 1e-20 and got["n_ext"] == 0 and got["e_pt2_external"] == 0.0
    # duplicates in V: first occurrence wins, as in the reference's tolerant DetMap
    V = np.concatenate([F[:30], F[5:10]])
    c = np.linspace(1.0, 0.1, V.shape[0])
    c /= np.linalg.norm(c)
    exp = octx.compute_pt2(V, c, 7, -84.0, False, 1e-8)
    got = core.compute_pt2_device(V, c, ctx, 7, -84.0, False, 1e-8, want_external=True)
    _pt2_check(got, exp, exact_vals=True)
    # duplicates before other rows: coefficients are read at the rank of the
    # first occurrence (local_conn.cpp:272, 322-333)
    V = np.concatenate([F[:30], F[5:10], F[30:80]])
    c = np.linspace(1.0, 0.1, V.shape[0])
    c /= np.linalg.norm(c)
    exp = octx.compute_pt2(V, c, 7, -84.0, False, 1e-8)
    got = core.compute_pt2_device(V, c, ctx, 7, -84.0, False, 1e-8, want_external=True)
    _pt2_check(got, exp, exact_vals=True)


def test_pt2_wide_words():
    from detnqs_amd import core, synth
    ints = synth.make_integrals(34, 8, 10, seed=5)
    ctx, octx = _ctx_pair(ints, 34)
    rng = np.random.default_rng(8)
    V = np.unique(_random_batch(rng, 34, 4, 4, 60), axis=0)
    coef = rng.standard_normal(V.shape[0])
    coef /= np.linalg.norm(coef)
    E = octx.variational_energy(V, coef, 34) - 0.2
    for hbm, eps in ((True, 1e-5), (False, 1e-6)):
        exp = octx.compute_pt2(V, coef, 34, E, hbm, eps)
        got = core.compute_pt2_device(V, coef, ctx, 34, E, hbm, eps, want_external=True)
        _pt2_check(got, exp, exact_vals=True)


def test_pt2_errors():
    from detnqs_amd import core
    ctx = core.IntCtx(H2O_STO3G, 7)
    F = fci_dets(7, 5, 5)[:10]
    with pytest.raises(ValueError, match="coefficient size mismatch"):
        core.compute_pt2(F, np.ones(9), ctx, 7, -80.0)
    with pytest.raises(ValueError, match="Heat-Bath table not initialized"):
        core.compute_pt2(F, np.ones(10), ctx, 7, -80.0, True, 1e-6)
    empty = core.compute_pt2(np.zeros((0, 2), dtype=np.uint64), np.zeros(0), ctx, 7, -80.0)
    assert empty == {"e_pt2_internal": 0.0, "e_pt2_external": 0.0, "n_ext": 0}


def test_block_cache_reuse_across_streams_and_trim():
    """Results stay identical when the cached blocks of one call are reused by
    the next on another stream, and after the cache is returned to the driver."""
    import torch
    from detnqs_amd import core, synth
    ints, dets = synth.config0()
    ctx, _ = _ctx_pair(ints, 13)
    V = dets[:1500]
    first = core.gen_perturbative_dets(V, 13, ctx, None, "static", 1e-6, chunk_items=200_000)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        again = core.gen_perturbative_dets(V, 13, ctx, None, "static", 1e-6, chunk_items=150_000)
    side.synchronize()
    assert np.array_equal(first, again)
    core.trim_memory()
    assert np.array_equal(core.gen_perturbative_dets(V, 13, ctx, None, "static", 1e-6), first)


def test_pt2_wide_context_global_tables():
    """62 orbitals: the diagonal tables exceed the LDS budget of the external
    term kernel, which then reads them from HBM; no heat-bath table."""
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    ints = synth.make_integrals(62, 6, 8, seed=9)
    ctx = core.IntCtx.from_integrals(62, 6, ints.h1e, ints.h2e, ints.e_nuc)
    octx = Lib("oracle").ctx(ints=ints)
    rng = np.random.default_rng(12)
    V = np.unique(_random_batch(rng, 62, 3, 3, 25), axis=0)
    coef = rng.standard_normal(V.shape[0])
    coef /= np.linalg.norm(coef)
    E = octx.variational_energy(V, coef, 62) - 0.3
    exp = octx.compute_pt2(V, coef, 62, E, False, 1e-7)
    got = core.compute_pt2_device(V, coef, ctx, 62, E, False, 1e-7, want_external=True)
    _pt2_check(got, exp, exact_vals=True)
