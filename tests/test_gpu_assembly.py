"""GPU parity: Slater–Condon row assembly (get_local_connections) and the
diagonal against the oracle, bit-exact (structure, order, phases, values)."""

import numpy as np
import pytest

from helpers import H2O_STO3G, assert_conn_equal, fci_dets, same_bits, write_tmp_fcidump

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg0(tmp_path_factory):
    from detnqs_amd import core, synth
    from oracle.oracle import Lib

    ints, dets = synth.config0()
    path = write_tmp_fcidump(tmp_path_factory.mktemp("c0"), ints)
    ctx = core.IntCtx(path, 13)
    ctx.hb_prepare(1e-15)
    octx = Lib("oracle").ctx(path, 13)
    octx.hb_prepare(1e-15)
    return ints, dets, ctx, octx


@pytest.fixture(scope="module")
def h2o():
    from detnqs_amd import core
    from oracle.oracle import Lib

    ctx = core.IntCtx(H2O_STO3G, 7)
    ctx.hb_prepare(1e-15)
    octx = Lib("oracle").ctx(H2O_STO3G, 7)
    octx.hb_prepare(1e-15)
    return ctx, octx, fci_dets(7, 5, 5)


def test_hb_table_matches_oracle(cfg0):
    _, _, ctx, octx = cfg0
    g, e = ctx.hb_export(), octx.hb_export()
    for k in ("row_offsets", "a", "b"):
        assert np.array_equal(g[k].astype(np.int64), e[k].astype(np.int64)), k
    assert same_bits(g["w"], e["w"])


def test_diag_bit_exact(cfg0):
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    assert same_bits(core.get_ham_diag(dets, ctx), octx.ham_diag(dets))


@pytest.mark.parametrize("use_hb,eps", [(False, 1e-6), (True, 1e-6), (True, 0.05), (True, 0.0)])
def test_local_connections_config0(cfg0, use_hb, eps):
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    sub = dets[:1500]
    got = core.get_local_connections(sub, ctx, 13, use_heatbath=use_hb, eps1=eps)
    exp = octx.local_connections(sub, 13, use_hb, eps)
    assert_conn_equal(got, exp)
    assert got["offsets"].dtype == np.int32


def test_local_connections_h2o_fci(h2o):
    from detnqs_amd import core
    ctx, octx, dets = h2o
    for use_hb in (False, True):
        got = core.get_local_connections(dets, ctx, 7, use_heatbath=use_hb, eps1=1e-4)
        exp = octx.local_connections(dets, 7, use_hb, 1e-4)
        assert_conn_equal(got, exp)


def test_local_connections_full_config0_stats(cfg0):
    """All 7,241 rows: per-row counts and per-row value sums (same order) match."""
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    off, od, ov = core.get_local_connections_device(dets, ctx, 13, True, 1e-6)
    counts = np.diff(off.cpu().numpy())
    ec, _ = octx.local_conn_stats(dets, 13, True, 1e-6)
    assert np.array_equal(counts, ec)


def test_chunked_fill_matches_single_call(cfg0):
    """Streaming row chunks through one buffer gives the same bytes."""
    import torch
    from detnqs_amd import core
    _, dets, ctx, _ = cfg0
    d = core.dets_to_device(dets[:900])
    off, total = core.local_conn_count(d, ctx, 13, True, 1e-6)
    full_d = torch.empty((total, 2), dtype=torch.int64, device=d.device)
    full_v = torch.empty(total, dtype=torch.float64, device=d.device)
    core.local_conn_fill(d, ctx, 13, True, 1e-6, off, full_d, full_v)
    offs = off.cpu().numpy()
    for start, stop in [(0, 300), (300, 301), (301, 900)]:
        m = int(offs[stop] - offs[start])
        cd = torch.empty((max(m, 1), 2), dtype=torch.int64, device=d.device)
        cv = torch.empty(max(m, 1), dtype=torch.float64, device=d.device)
        core.local_conn_fill(d[start:stop], ctx, 13, True, 1e-6, off[start:stop + 1], cd, cv)
        assert torch.equal(cd[:m], full_d[offs[start]:offs[stop]])
        assert torch.equal(cv[:m], full_v[offs[start]:offs[stop]])


def test_empty_and_errors(cfg0):
    from detnqs_amd import core
    _, dets, ctx, _ = cfg0
    got = core.get_local_connections(np.zeros((0, 2), dtype=np.uint64), ctx, 13, True, 1e-6)
    assert got["offsets"].tolist() == [0] and got["values"].size == 0
    ints = cfg0[0]
    c = core.IntCtx.from_integrals(ints.n_orb, ints.n_elec, ints.h1e, ints.h2e, ints.e_nuc)
    with pytest.raises(ValueError):
        core.get_local_connections(dets[:4], c, 13, use_heatbath=True)  # no HB table
    with pytest.raises(ValueError):
        core.get_local_connections(dets[:4], c, 14)  # n_orb beyond the integrals
    with pytest.raises(ValueError):
        core.get_local_connections(dets[:4, :1], c, 13)
    assert c.get_e_nuc() == ints.e_nuc
    bad = dets[:8].copy()
    bad[3, 0] |= np.uint64(1 << 13)  # orbital 13 with n_orb = 13
    with pytest.raises(ValueError):
        core.get_local_connections(bad, ctx, 13, True, 1e-6)
    # the flag is cleared: the next call succeeds
    assert core.get_local_connections(dets[:8], ctx, 13, True, 1e-6)["values"].size > 0
    d = core.get_ham_diag(bad, c)
    assert np.isnan(d[3]) and np.all(np.isfinite(np.delete(d, 3)))
