"""GPU parity of get_local_connections away from the small fixtures: the
28-orbital benchmark system, unequal spin counts at 48 orbitals, the 64-orbital
limit (words using bit 63), and 32 electrons (496 occupied pairs per row), each
bit-exact against the oracle (structure, order, phases, values)."""

import numpy as np
import pytest

from helpers import assert_conn_equal

pytestmark = pytest.mark.gpu


def _random_dets(rng, n_orb, na, nb, B):
    out = np.zeros((B, 2), dtype=np.uint64)
    one = np.uint64(1)
    for i in range(B):
        for w, k in ((0, na), (1, nb)):
            for o in rng.choice(n_orb, k, replace=False):
                out[i, w] |= one << np.uint64(o)
    return np.unique(out, axis=0)


def _pair(ints, n, hb=True):
    from detnqs_amd import core
    from oracle.oracle import Lib
    ctx = core.IntCtx.from_integrals(n, ints.n_elec, ints.h1e, ints.h2e, ints.e_nuc)
    octx = Lib("oracle").ctx(ints=ints)
    if hb:
        ctx.hb_prepare(1e-15)
        octx.hb_prepare(1e-15)
    return ctx, octx


def _check(ctx, octx, dets, n, use_hb, eps):
    from detnqs_amd import core
    got = core.get_local_connections(dets, ctx, n, use_heatbath=use_hb, eps1=eps)
    exp = octx.local_connections(dets, n, use_hb, eps)
    assert_conn_equal(got, exp)
    return int(exp["offsets"][-1])


def test_benchmark_system_rows():
    from detnqs_amd import synth
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    dets = synth.config1_dets(3000, seed=1)[::10]
    ctx, octx = _pair(ints, 28)
    for use_hb, eps in ((True, 1e-7), (True, 1e-3), (True, 0.0), (False, 1e-7)):
        assert _check(ctx, octx, dets, 28, use_hb, eps) > 0


def test_unequal_spins_48_orbitals():
    from detnqs_amd import synth
    ints = synth.make_integrals(48, 11, 12, seed=48)
    rng = np.random.default_rng(48)
    dets = _random_dets(rng, 48, 6, 5, 60)
    ctx, octx = _pair(ints, 48)
    for use_hb, eps in ((True, 1e-5), (True, 0.05), (False, 1e-6)):
        _check(ctx, octx, dets, 48, use_hb, eps)


def test_64_orbital_words():
    from detnqs_amd import synth
    ints = synth.make_integrals(64, 6, 8, seed=64)
    rng = np.random.default_rng(64)
    dets = _random_dets(rng, 64, 3, 3, 40)
    dets[0] = (np.uint64(1) << np.uint64(63)) | np.uint64(3), (np.uint64(1) << np.uint64(63)) | np.uint64(5)
    ctx, octx = _pair(ints, 64, hb=False)
    _check(ctx, octx, dets, 64, False, 1e-6)


def test_32_orbital_words():
    """n_orb = 32: words using bit 31, heat-bath on."""
    from detnqs_amd import synth
    ints = synth.make_integrals(32, 8, 8, seed=32)
    rng = np.random.default_rng(32)
    dets = _random_dets(rng, 32, 4, 4, 40)
    top = np.uint64(1) << np.uint64(31)
    dets[0] = top | np.uint64(7), top | np.uint64(7)
    dets[1] = top | np.uint64(0b1011), np.uint64(0b1111)
    dets = np.unique(dets, axis=0)
    ctx, octx = _pair(ints, 32)
    for use_hb, eps in ((True, 1e-6), (True, 0.0), (True, 1e-3)):
        assert _check(ctx, octx, dets, 32, use_hb, eps) > 0


def test_32_electrons_many_pairs():
    from detnqs_amd import synth
    ints = synth.make_integrals(40, 32, 10, seed=40)
    rng = np.random.default_rng(40)
    dets = _random_dets(rng, 40, 16, 16, 12)
    ctx, octx = _pair(ints, 40)
    for use_hb, eps in ((True, 1e-4), (True, 1e-2)):
        _check(ctx, octx, dets, 40, use_hb, eps)


def test_projected_blocks_wide_words():
    """get_ham_conn / get_ham_conn_amp at 36 orbitals: P ranked through the
    128-bit path (the packed-key ranking covers n_orb <= 32)."""
    from detnqs_amd import core, synth
    from helpers import assert_coo_equal
    ints = synth.make_integrals(36, 8, 10, seed=36)
    rng = np.random.default_rng(36)
    dets = _random_dets(rng, 36, 4, 4, 40)
    ctx, octx = _pair(ints, 36)
    for got, exp in ((core.get_ham_conn(dets, ctx, 36, True, 1e-4), octx.ham_conn(dets, 36, True, 1e-4)),
                     (core.get_ham_conn(dets, ctx, 36, False, 1e-6), octx.ham_conn(dets, 36, False, 1e-6))):
        assert_coo_equal(got["H_VV"], exp["H_VV"])
        assert_coo_equal(got["H_VP"], exp["H_VP"])
        assert np.array_equal(got["det_P"], exp["det_P"]) and got["size_P"] == exp["size_P"]
    psi = rng.standard_normal(dets.shape[0]) * 0.1
    got, exp = core.get_ham_conn_amp(dets, psi, ctx, 36, 1e-4), octx.ham_conn_amp(dets, psi, 36, 1e-4)
    assert_coo_equal(got["H_VP"], exp["H_VP"])
    assert np.array_equal(got["det_P"], exp["det_P"])
