"""GPU parity: projection onto V (get_ham_vv / block / conn / conn_amp),
H_eff SpGEMM and the fused variational energy against the oracle.
COO structure, P sets and values are bit-exact; the energy (a reduction whose
order the reference leaves to OpenMP) is within 1e-12 relative."""

import numpy as np
import pytest

from helpers import H2O_STO3G, assert_coo_equal, explicit_p_sets, fci_dets, write_tmp_fcidump

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


def _blocks_equal(got, exp):
    assert_coo_equal(got["H_VV"], exp["H_VV"])
    assert_coo_equal(got["H_VP"], exp["H_VP"])
    assert np.array_equal(got["det_P"], exp["det_P"])
    assert got["size_P"] == exp["size_P"]


def test_ham_vv_bit_exact(cfg0, h2o):
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    assert_coo_equal(core.get_ham_vv(dets[:3000], ctx, 13), octx.ham_vv(dets[:3000], 13))
    hctx, hoctx, F = h2o
    assert_coo_equal(core.get_ham_vv(F, hctx, 7), hoctx.ham_vv(F, 7))


def test_ham_block_explicit_p(h2o):
    from detnqs_amd import core
    ctx, octx, F = h2o
    got = core.get_ham_block(F[:50], F[50:80], ctx, 7)
    exp = octx.ham_block(F[:50], F[50:80], 7)
    assert_coo_equal(got["H_VV"], exp["H_VV"])
    assert_coo_equal(got["H_VP"], exp["H_VP"])
    assert np.array_equal(got["det_P"], exp["det_P"])


def test_ham_block_explicit_p_overlap(cfg0, h2o):
    """P overlapping V: kets that lie in V go to H_VV, never to H_VP; P members
    connected to nothing give empty columns (build_ham.cpp:97-105)."""
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    V = dets[:300]
    Pc = octx.ham_conn(V, 13, True, 1e-3)["det_P"]
    for P in explicit_p_sets(V, Pc, dets[5000:5200], seed=5):
        got, exp = core.get_ham_block(V, P, ctx, 13), octx.ham_block(V, P, 13)
        assert_coo_equal(got["H_VV"], exp["H_VV"])
        assert_coo_equal(got["H_VP"], exp["H_VP"])
        assert np.array_equal(got["det_P"], exp["det_P"])
    hctx, hoctx, F = h2o
    perm = np.random.default_rng(41).permutation(F.shape[0])
    got = core.get_ham_block(F[perm[:60]], F[perm[40:200]], hctx, 7)
    exp = hoctx.ham_block(F[perm[:60]], F[perm[40:200]], 7)
    assert_coo_equal(got["H_VV"], exp["H_VV"])
    assert_coo_equal(got["H_VP"], exp["H_VP"])
    with pytest.raises(ValueError, match="duplicate"):
        core.get_ham_block(V, np.concatenate([Pc[:10], Pc[3:4]]), ctx, 13)


@pytest.mark.parametrize("use_hb,eps", [(True, 1e-3), (True, 1e-6), (False, 1e-6)])
def test_ham_conn_static(cfg0, use_hb, eps):
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    V = dets[:300]
    _blocks_equal(core.get_ham_conn(V, ctx, 13, use_hb, eps), octx.ham_conn(V, 13, use_hb, eps))


def test_ham_conn_amp(cfg0):
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    V = dets[:300]
    psi = np.random.default_rng(77).standard_normal(300) * 0.1
    psi[5] = 0.0  # tau capped by max(|psi|, 1e-12)
    _blocks_equal(core.get_ham_conn_amp(V, psi, ctx, 13, 1e-3), octx.ham_conn_amp(V, psi, 13, 1e-3))


def test_ham_conn_mixed_sectors(cfg0):
    """Rows of different electron numbers in one V: P mixes sectors, so its
    ket keys are the packed (alpha << n_orb | beta) words instead of the
    per-sector colex ranks; each row is ordered within its own sector."""
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    V = dets[:120].copy()
    less_a = dets[120:200].copy()
    top = np.array([int(a).bit_length() - 1 for a in less_a[:, 0]], dtype=np.uint64)
    less_a[:, 0] ^= np.uint64(1) << top                               # one alpha electron fewer
    more_b = dets[200:260].copy()
    free = [next(p for p in range(13) if not (int(b) >> p) & 1) for b in more_b[:, 1]]
    more_b[:, 1] |= np.array([1 << p for p in free], dtype=np.uint64)  # one beta electron more
    W = np.concatenate([V, less_a, more_b])
    _, first = np.unique(W, axis=0, return_index=True)
    W = W[np.sort(first)]
    assert len({(bin(int(a)).count("1"), bin(int(b)).count("1")) for a, b in W}) == 3
    _blocks_equal(core.get_ham_conn(W, ctx, 13, True, 1e-4), octx.ham_conn(W, 13, True, 1e-4))
    psi = np.random.default_rng(78).standard_normal(W.shape[0]) * 0.1
    _blocks_equal(core.get_ham_conn_amp(W, psi, ctx, 13, 1e-3), octx.ham_conn_amp(W, psi, 13, 1e-3))


def test_ham_conn_empty_p(h2o):
    """V = the whole FCI space: every connection lies in V, H_VP and P are empty."""
    from detnqs_amd import core
    ctx, octx, F = h2o
    got, exp = core.get_ham_conn(F, ctx, 7, True, 1e-6), octx.ham_conn(F, 7, True, 1e-6)
    _blocks_equal(got, exp)
    assert got["size_P"] == 0 and got["H_VP"]["val"].size == 0


@pytest.mark.parametrize("n_a,n_b", [(16, 16), (3, 29)])
def test_ham_conn_32_orbitals(n_a, n_b):
    """The widest strings the ket keys take (32 orbitals): half filling has the
    most connections per row (94,849), the lopsided case the widest colex
    range on one side."""
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    ints = synth.make_integrals(32, n_a + n_b, 12, seed=3200 + n_a)
    rng = np.random.default_rng(32 + n_a)
    rows = set()
    while len(rows) < 12:
        a = sum(1 << int(p) for p in rng.choice(32, n_a, replace=False))
        b = sum(1 << int(p) for p in rng.choice(32, n_b, replace=False))
        rows.add((a, b))
    V = np.array(sorted(rows), dtype=np.uint64)
    ctx = core.IntCtx.from_integrals(32, n_a + n_b, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    octx = Lib("oracle").ctx(ints=ints)
    octx.hb_prepare(1e-15)
    _blocks_equal(core.get_ham_conn(V, ctx, 32, True, 1e-2), octx.ham_conn(V, 32, True, 1e-2))


def test_ham_eff_bit_exact(cfg0, h2o):
    from detnqs_amd import core
    from oracle.oracle import Lib
    O = Lib("oracle")
    _, dets, ctx, octx = cfg0
    V = dets[:300]
    psi = np.random.default_rng(77).standard_normal(300) * 0.1
    blk = octx.ham_conn_amp(V, psi, 13, 1e-3)
    hpp = octx.ham_diag(blk["det_P"])
    e_ref = float(hpp.min() - 1.0)
    for reg in ("sigma", "linear_shift"):
        got = core.get_ham_eff(blk["H_VV"], blk["H_VP"], hpp, e_ref, reg, 1e-6)
        exp = O.ham_eff(blk["H_VV"], blk["H_VP"], hpp, e_ref, reg, 1e-6)
        assert_coo_equal(got, exp)
    # skip rule active (tiny dinv * ||col||^2) and the empty-P path
    hctx, hoctx, F = h2o
    hf = np.array([[0b11111, 0b11111]], dtype=np.uint64)
    b2 = hoctx.ham_conn(hf, 7, True, 1e-5)
    hpp2 = hoctx.ham_diag(b2["det_P"])
    e2 = float(hoctx.ham_diag(hf)[0])
    assert_coo_equal(core.get_ham_eff(b2["H_VV"], b2["H_VP"], hpp2, e2, "sigma", 1e-10),
                     O.ham_eff(b2["H_VV"], b2["H_VP"], hpp2, e2, "sigma", 1e-10))
    assert_coo_equal(core.get_ham_eff(b2["H_VV"], b2["H_VP"], hpp2, e2 - 1e6, "sigma", 1e-10),
                     O.ham_eff(b2["H_VV"], b2["H_VP"], hpp2, e2 - 1e6, "sigma", 1e-10))
    empty_vp = {"row": np.zeros(0, np.uint32), "col": np.zeros(0, np.uint32), "val": np.zeros(0), "shape": (1, 0)}
    assert_coo_equal(core.get_ham_eff(b2["H_VV"], empty_vp, np.zeros(0), e2),
                     O.ham_eff(b2["H_VV"], empty_vp, np.zeros(0), e2))


@pytest.mark.parametrize("lds_cols", ["128", "64"])
def test_ham_eff_windows_and_sort_path(cfg0, monkeypatch, lds_cols):
    """Narrow accumulator windows: 128 columns gives three LDS passes over
    V = 300; 64 gives five, beyond the LDS limit, so the triple sort path
    runs.  Columns longer than one wavefront (up to 98 rows) are present."""
    from detnqs_amd import core
    from oracle.oracle import Lib
    O = Lib("oracle")
    _, dets, ctx, octx = cfg0
    V = dets[:300]
    psi = np.random.default_rng(77).standard_normal(300) * 0.1
    blk = octx.ham_conn_amp(V, psi, 13, 1e-3)
    assert np.bincount(blk["H_VP"]["col"]).max() > 64
    hpp = octx.ham_diag(blk["det_P"])
    e_ref = float(hpp.min() - 1.0)
    monkeypatch.setenv("DQ_HEFF_LDS_COLS", lds_cols)
    for reg in ("sigma", "linear_shift"):
        got = core.get_ham_eff(blk["H_VV"], blk["H_VP"], hpp, e_ref, reg, 1e-6)
        exp = O.ham_eff(blk["H_VV"], blk["H_VP"], hpp, e_ref, reg, 1e-6)
        assert_coo_equal(got, exp)


def test_ham_eff_unsorted_input(cfg0):
    """H_VP given out of order and with an explicit zero: the canonicalising
    sort path (coo_to_csr) instead of the already-canonical one."""
    from detnqs_amd import core
    from oracle.oracle import Lib
    O = Lib("oracle")
    _, dets, ctx, octx = cfg0
    V = dets[:200]
    blk = octx.ham_conn(V, 13, True, 1e-3)
    hpp = octx.ham_diag(blk["det_P"])
    e_ref = float(hpp.min() - 1.0)
    vp = dict(blk["H_VP"])
    perm = np.random.default_rng(3).permutation(vp["row"].size)
    vp = {"row": vp["row"][perm], "col": vp["col"][perm], "val": vp["val"][perm].copy(), "shape": vp["shape"]}
    vp["val"][5] = 0.0
    got = core.get_ham_eff(blk["H_VV"], vp, hpp, e_ref, "sigma", 1e-6)
    exp = O.ham_eff(blk["H_VV"], vp, hpp, e_ref, "sigma", 1e-6)
    assert_coo_equal(got, exp)


def _synthetic_blocks(rng, n_v, n_p, per_row):
    """Random H_VV / H_VP in COO, out of order, with some empty H_VP rows,
    columns shared by many rows, and duplicate entries whose parts are dyadic
    (their sum is exact in any order)."""
    vv_r, vv_c = np.meshgrid(np.arange(n_v), np.arange(n_v), indexing="ij")
    keep = rng.random(vv_r.shape) < 0.3
    vv = {"row": vv_r[keep].astype(np.uint32), "col": vv_c[keep].astype(np.uint32),
          "val": rng.standard_normal(int(keep.sum())), "shape": (n_v, n_v)}
    rows, cols = [], []
    hot = rng.choice(n_p, size=max(1, n_p // 10), replace=False)  # columns most rows touch
    for i in range(n_v):
        if i % 5 == 3:
            continue  # empty row
        c = np.unique(np.concatenate([rng.choice(n_p, per_row, replace=False), hot[rng.random(hot.size) < 0.8]]))
        rows.append(np.full(c.size, i)), cols.append(c)
    r = np.concatenate(rows).astype(np.uint32) if rows else np.zeros(0, np.uint32)
    c = np.concatenate(cols).astype(np.uint32) if cols else np.zeros(0, np.uint32)
    v = rng.standard_normal(r.size) * 0.1
    dup = rng.random(r.size) < 0.05
    v[dup] = 0.375
    r = np.concatenate([r, r[dup]])
    c = np.concatenate([c, c[dup]])
    v = np.concatenate([v, np.full(int(dup.sum()), 0.125)])
    perm = rng.permutation(r.size)
    vp = {"row": r[perm], "col": c[perm], "val": v[perm], "shape": (n_v, n_p)}
    hpp = rng.standard_normal(n_p) + 2.0
    return vv, vp, hpp


@pytest.mark.parametrize("n_v,n_p,lds_cols", [(1, 40, None), (3, 200, None), (70, 300, "64"), (130, 700, None),
                                              (200, 900, "128")])
def test_ham_eff_synthetic_edges(monkeypatch, n_v, n_p, lds_cols):
    """Tiny V, a partial last window (70 rows in windows of 64), empty H_VP
    rows, duplicates and out-of-order input, against the oracle."""
    from detnqs_amd import core
    from oracle.oracle import Lib
    O = Lib("oracle")
    rng = np.random.default_rng(n_v * 1000 + n_p)
    vv, vp, hpp = _synthetic_blocks(rng, n_v, n_p, per_row=min(25, n_p // 2))
    if lds_cols:
        monkeypatch.setenv("DQ_HEFF_LDS_COLS", lds_cols)
    for reg, e_ref in (("sigma", -1.0), ("linear_shift", 0.5)):
        got = core.get_ham_eff(vv, vp, hpp, e_ref, reg, 1e-8)
        exp = O.ham_eff(vv, vp, hpp, e_ref, reg, 1e-8)
        assert_coo_equal(got, exp)


@pytest.mark.parametrize("use_hb", [False, True])
def test_variational_energy(cfg0, use_hb):
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    V = dets[:3000]
    c = np.random.default_rng(5).standard_normal(3000)
    c /= np.linalg.norm(c)
    c[10] = 1e-7  # skipped row: c^2 < 1e-12
    e_gpu, sig = core.compute_variational_energy_device(V, c, ctx, 13, use_hb, 1e-4, want_sigma=True)
    e_orc, osig = octx.variational_energy(V, c, 13, use_hb, 1e-4, want_sigma=True)
    assert abs(e_gpu - e_orc) <= 1e-12 * abs(e_orc)
    sig = sig.cpu().numpy()
    assert np.all(np.abs(sig - osig) <= 1e-12 * np.maximum(np.abs(osig), 1.0))
    assert sig[10] == 0.0


def test_energy_zero_variance_h2o(h2o):
    """FCI eigenvector: E_loc(x) = E_FCI - e_nuc for every x."""
    from detnqs_amd import core
    ctx, octx, F = h2o
    vv = octx.ham_vv(F, 7)
    H = np.zeros((F.shape[0],) * 2)
    H[vv["row"], vv["col"]] = vv["val"]
    w, U = np.linalg.eigh(H)
    psi = U[:, 0]
    e, sig = core.compute_variational_energy_device(F, psi, ctx, 7, False, 0.0, want_sigma=True)
    sig = sig.cpu().numpy()
    big = np.abs(psi) > 1e-6
    assert np.allclose(sig[big] / psi[big], w[0], atol=1e-9)
    assert abs(e - w[0]) < 1e-10


def test_energy_same_alone_and_in_batch(cfg0):
    """Per-row sigma does not depend on the batch it is computed in."""
    from detnqs_amd import core
    _, dets, ctx, _ = cfg0
    V = dets[:2000]
    c = np.random.default_rng(9).standard_normal(2000)
    _, s_all = core.compute_variational_energy_device(V, c, ctx, 13, True, 1e-5, want_sigma=True)
    _, s_again = core.compute_variational_energy_device(V, c, ctx, 13, True, 1e-5, want_sigma=True)
    assert s_all.cpu().numpy().tobytes() == s_again.cpu().numpy().tobytes()


def test_duplicates_and_errors(cfg0):
    from detnqs_amd import core
    _, dets, ctx, octx = cfg0
    V = np.concatenate([dets[:10], dets[3:4]])
    with pytest.raises(ValueError, match="duplicate"):
        core.get_ham_vv(V, ctx, 13)
    with pytest.raises(ValueError):
        core.get_ham_conn_amp(dets[:10], np.ones(9), ctx, 13)
    # energy tolerates duplicates (first occurrence wins): equals the deduplicated sum
    c = np.linspace(0.1, 1.0, 11)
    e = core.compute_variational_energy(V, c, ctx, 13)
    eo = octx.variational_energy(V, c, 13, False, 1e-6)
    assert abs(e - eo) <= 1e-12 * abs(eo)
