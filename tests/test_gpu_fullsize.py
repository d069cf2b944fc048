"""Full-size parity.

get_local_connections at the benchmark's full size: 28 orbitals 7a+7b,
1,000,000 config-1 determinants, heat-bath eps1 = 1e-7 (3.07e10 elements),
streamed through a 4e8-element buffer as a caller would.

Every element is checked for size-independent properties on the device
(orbital range, 7 alpha and 7 beta electrons, |H| > 1e-15 and finite, each
row led by its own determinant), and 64 rows spread over the space (the first
8, the last 8 and 48 between) are compared bit for bit with the oracle: the
same rows computed inside the full-size call, chunk boundaries included.

get_ham_conn at the projected-assembly benchmark's size (20,000 rows of V,
static heat-bath eps1 = 1e-6) is checked the same way: properties on every
element, bit-exact rows against the oracle; gen_perturbative_dets at the
same size through the properties of P = C(V) minus V; compute_pt2 at the
same size through its exact scaling with the coefficients."""

import time

import numpy as np
import pytest
import torch

from helpers import assert_conn_equal

pytestmark = pytest.mark.gpu

N_ROWS = 1_000_000
EPS1 = 1e-7
CAP = 400_000_000


def _popcount(x: torch.Tensor) -> torch.Tensor:
    """Bit count of non-negative int64 values (SWAR)."""
    x = x - ((x >> 1) & 0x5555555555555555)
    x = (x & 0x3333333333333333) + ((x >> 2) & 0x3333333333333333)
    x = (x + (x >> 4)) & 0x0F0F0F0F0F0F0F0F
    return (x * 0x0101010101010101) >> 56


def _chunks(offs: np.ndarray, cap: int):
    out, start, n = [], 0, offs.size - 1
    while start < n:
        stop = int(np.searchsorted(offs, offs[start] + cap, side="right")) - 1
        stop = min(max(stop, start + 1), n)
        assert offs[stop] - offs[start] <= cap
        out.append((start, stop))
        start = stop
    return out


def test_config1_full_size():
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    dets = synth.config1_dets(N_ROWS, seed=1)
    n = dets.shape[0]
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    d = core.dets_to_device(dets)
    print(f"{n} determinants on the device", flush=True)
    off, total = core.local_conn_count(d, ctx, 28, True, EPS1)
    offs = off.cpu().numpy()
    print(f"count pass: {total} elements", flush=True)
    assert offs[0] == 0 and offs[-1] == total
    assert np.all(np.diff(offs) >= 1)  # every row keeps at least its diagonal
    assert total > 3e10

    sample = sorted(set(np.linspace(0, n - 1, 48).astype(np.int64).tolist()) | set(range(8)) |
                    set(range(n - 8, n)))
    got = {}
    buf_d = torch.empty((CAP, 2), dtype=torch.int64, device=d.device)
    buf_v = torch.empty(CAP, dtype=torch.float64, device=d.device)
    chunks = _chunks(offs, CAP)
    t0 = time.perf_counter()
    print(f"{len(chunks)} chunks", flush=True)
    for c, (a, b) in enumerate(chunks):
        core.local_conn_fill(d[a:b], ctx, 28, True, EPS1, off[a:b + 1], buf_d, buf_v)
        m = int(offs[b] - offs[a])
        od, ov = buf_d[:m], buf_v[:m]
        assert bool(((od >= 0) & (od < (1 << 28))).all())
        assert bool((_popcount(od) == 7).all())
        assert bool(torch.isfinite(ov).all()) and bool((ov.abs() > 1e-15).all())
        starts = off[a:b] - off[a]
        assert bool((od[starts] == d[a:b]).all())  # every row starts with its own determinant
        print(f"chunk {c + 1}/{len(chunks)}: rows {a}-{b}, {m} elements checked, {time.perf_counter() - t0:.1f} s",
              flush=True)
        for r in sample:
            if a <= r < b:
                s0, s1 = int(offs[r] - offs[a]), int(offs[r + 1] - offs[a])
                got[r] = (od[s0:s1].cpu().numpy().view(np.uint64), ov[s0:s1].cpu().numpy())
    assert sorted(got) == sample

    exp = Lib("oracle").ctx(ints=ints)
    exp.hb_prepare(1e-15)
    want = exp.local_connections(dets[sample], 28, True, EPS1)
    sizes = np.array([got[r][1].size for r in sample])
    have = {"offsets": np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64),
            "dets": np.concatenate([got[r][0] for r in sample]),
            "values": np.concatenate([got[r][1] for r in sample])}
    assert_conn_equal(have, want)


PROJ_ROWS = 20_000
PROJ_EPS1 = 1e-6


def _row_range(rows: np.ndarray, r: int):
    return int(np.searchsorted(rows, r, side="left")), int(np.searchsorted(rows, r, side="right"))


def _row_entries(blocks: dict, V: np.ndarray, r: int):
    """Row r of H_VV and H_VP as (determinant key, value) pairs sorted by key."""
    keys, vals = [], []
    for name, table in (("H_VV", V), ("H_VP", blocks["det_P"])):
        coo = blocks[name]
        a, b = _row_range(coo["row"], r)
        c = coo["col"][a:b].astype(np.int64)
        keys.append((table[c, 0] << np.uint64(28)) | table[c, 1])
        vals.append(coo["val"][a:b])
    k, v = np.concatenate(keys), np.concatenate(vals)
    o = np.argsort(k, kind="stable")
    return k[o], v[o]


def test_ham_conn_full_size():
    """get_ham_conn at the projected-assembly benchmark's size: V = the first
    20,000 config-1 determinants, static heat-bath eps1 = 1e-6 (6.1e8 H_VP
    elements, |P| = 3.1e8).

    Every element and every P determinant is checked for properties that hold
    at any size (canonical (row, column) order with no repeats, P strictly
    ascending, 7 alpha and 7 beta electrons, no P determinant in V, every P
    determinant used), and 24 rows (the first 8, the last 8 and 8 between)
    are compared bit for bit with the oracle's projection of those rows onto
    themselves: a row's connections, as (determinant, value) pairs, do not
    depend on which of them fall in V."""
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    V = synth.config1_dets(PROJ_ROWS, seed=1)
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    t0 = time.perf_counter()
    got = core.get_ham_conn(V, ctx, 28, True, PROJ_EPS1)
    vp, dp = got["H_VP"], got["det_P"]
    n_p = got["size_P"]
    print(f"get_ham_conn: nnz(H_VV) {got['H_VV']['val'].size}, nnz(H_VP) {vp['val'].size}, |P| {n_p}, "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    assert vp["val"].size > 6e8 and n_p > 3e8 and dp.shape == (n_p, 2)

    for name, ncol in (("H_VV", PROJ_ROWS), ("H_VP", n_p)):
        coo = got[name]
        assert coo["row"].max() < PROJ_ROWS and coo["col"].max() < ncol
        key = (coo["row"].astype(np.uint64) << np.uint64(32)) | coo["col"].astype(np.uint64)
        assert np.all(key[1:] > key[:-1]), name  # canonical order, no repeats
        assert np.all(np.isfinite(coo["val"])) and np.all(np.abs(coo["val"]) > 1e-15)
        del key
        print(f"{name}: order and values checked, {time.perf_counter() - t0:.1f} s", flush=True)
    assert np.all(np.bincount(vp["col"], minlength=n_p) > 0)  # every P determinant is used

    kp = (dp[:, 0] << np.uint64(28)) | dp[:, 1]
    assert np.all(kp[1:] > kp[:-1])
    assert np.all(np.bitwise_count(dp[:, 0]) == 7) and np.all(np.bitwise_count(dp[:, 1]) == 7)
    assert np.all(dp < np.uint64(1 << 28))
    kv = np.sort((V[:, 0] << np.uint64(28)) | V[:, 1])
    hit = np.minimum(np.searchsorted(kv, kp), kv.size - 1)
    assert not np.any(kv[hit] == kp)  # P and V are disjoint
    del kp, hit
    print(f"P checked, {time.perf_counter() - t0:.1f} s", flush=True)

    sample = sorted(set(range(8)) | set(range(PROJ_ROWS - 8, PROJ_ROWS)) |
                    set(np.linspace(8, PROJ_ROWS - 9, 8).astype(np.int64).tolist()))
    octx = Lib("oracle").ctx(ints=ints)
    octx.hb_prepare(1e-15)
    sub = V[sample]
    want = octx.ham_conn(sub, 28, True, PROJ_EPS1)
    for j, r in enumerate(sample):
        gk, gv = _row_entries(got, V, r)
        wk, wv = _row_entries(want, sub, j)
        assert np.array_equal(gk, wk), r
        assert np.array_equal(gv.view(np.uint64), wv.view(np.uint64)), r
    print(f"{len(sample)} rows equal to the oracle's, {time.perf_counter() - t0:.1f} s", flush=True)


def test_perturbative_space_full_size():
    """gen_perturbative_dets (static heat-bath eps1 = 1e-6) at the P-discovery
    benchmark's size, V = the first 20,000 config-1 determinants (|P| = 3.1e8).

    P = C(V) minus V, so: P is strictly ascending, holds 7 alpha and 7 beta
    electrons, misses V; the oracle's P of 24 sampled rows, less V, lies
    inside it (every connection of those rows was found); and 4,096 P
    determinants spread over the set are each at most a double excitation
    from some determinant of V."""
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    V = synth.config1_dets(PROJ_ROWS, seed=1)
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    t0 = time.perf_counter()
    P = core.gen_perturbative_dets(V, 28, ctx, None, "static", PROJ_EPS1)
    print(f"gen_perturbative_dets: |P| {P.shape[0]}, {time.perf_counter() - t0:.1f} s", flush=True)
    assert P.shape[0] > 3e8 and P.shape[1] == 2

    kp = (P[:, 0] << np.uint64(28)) | P[:, 1]
    assert np.all(kp[1:] > kp[:-1])
    assert np.all(np.bitwise_count(P[:, 0]) == 7) and np.all(np.bitwise_count(P[:, 1]) == 7)
    assert np.all(P < np.uint64(1 << 28))
    kv = np.sort((V[:, 0] << np.uint64(28)) | V[:, 1])
    hit = np.minimum(np.searchsorted(kv, kp), kv.size - 1)
    assert not np.any(kv[hit] == kp)
    del hit
    print(f"P order, electrons and disjointness checked, {time.perf_counter() - t0:.1f} s", flush=True)

    sample = sorted(set(range(8)) | set(range(PROJ_ROWS - 8, PROJ_ROWS)) |
                    set(np.linspace(8, PROJ_ROWS - 9, 8).astype(np.int64).tolist()))
    octx = Lib("oracle").ctx(ints=ints)
    octx.hb_prepare(1e-15)
    ws = octx.gen_perturbative(V[sample], 28, None, "static", PROJ_EPS1)
    kw = (ws[:, 0] << np.uint64(28)) | ws[:, 1]
    hw = np.minimum(np.searchsorted(kv, kw), kv.size - 1)
    kw = kw[kv[hw] != kw]  # the sampled rows' connections outside V
    at = np.minimum(np.searchsorted(kp, kw), kp.size - 1)
    assert kw.size > 0 and np.all(kp[at] == kw)
    print(f"{kw.size} connections of {len(sample)} sampled rows found in P, {time.perf_counter() - t0:.1f} s",
          flush=True)

    pick = P[np.linspace(0, P.shape[0] - 1, 4096).astype(np.int64)]
    diff = (np.bitwise_count(pick[:, None, 0] ^ V[None, :, 0]).astype(np.int32) +
            np.bitwise_count(pick[:, None, 1] ^ V[None, :, 1]).astype(np.int32))
    assert np.all(diff.min(axis=1) <= 4)  # at most two electrons moved


def test_pt2_full_size_scaling():
    """compute_pt2 at the PT2 benchmark's size (20,000 rows of V), unscreened
    (all doubles, eps1 = 0, so the external space does not depend on the
    coefficients).  Doubling every coefficient doubles sigma_i, r_i and every
    v_a exactly (a power-of-two scale is exact in FP64), so with the same
    candidates and the same summation order both terms must come out exactly
    four times larger, and n_ext must not change."""
    from detnqs_amd import core, synth
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    V = synth.config1_dets(PROJ_ROWS, seed=1)
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    rng = np.random.default_rng(5)
    c = rng.uniform(0.05, 1.0, PROJ_ROWS) * rng.choice([-1.0, 1.0], PROJ_ROWS)
    t0 = time.perf_counter()
    a = core.compute_pt2(V, c, ctx, 28, -1000.0, False, 0.0)
    b = core.compute_pt2(V, 2.0 * c, ctx, 28, -1000.0, False, 0.0)
    print(f"compute_pt2 x2: {a}, {b}, {time.perf_counter() - t0:.1f} s", flush=True)
    assert a["n_ext"] > 3e8 and b["n_ext"] == a["n_ext"]
    assert a["e_pt2_internal"] < 0.0 and a["e_pt2_external"] < 0.0  # every denominator is negative
    assert b["e_pt2_internal"] == 4.0 * a["e_pt2_internal"]
    assert b["e_pt2_external"] == 4.0 * a["e_pt2_external"]
