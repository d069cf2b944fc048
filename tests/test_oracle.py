"""Pins the oracle (CPU): reference golden digests, the reference's own KATs,
brute-force second quantisation, closed forms, and (when the reference build
exists in this container) direct equality with it."""

import hashlib
import os
from itertools import combinations

import numpy as np
import pytest

from helpers import (E_FCI_H2O_STO3G, GOLDEN, H2_631G, H2O_STO3G, explicit_p_sets, fci_dets,
                     write_tmp_fcidump)

from oracle.oracle import Lib, ref_available

O = Lib("oracle")


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype.str).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def conn_digest(d):
    return digest(np.asarray(d["offsets"], dtype=np.int64), d["dets"].astype(np.uint64), d["values"].astype(np.float64))


def coo_digest(c):
    return digest(c["row"].astype(np.uint32), c["col"].astype(np.uint32), c["val"].astype(np.float64),
                  np.asarray(c["shape"], dtype=np.uint32))


@pytest.fixture(scope="module")
def c0(tmp_path_factory):
    from detnqs_amd import synth
    ints, dets = synth.config0()
    p = write_tmp_fcidump(tmp_path_factory.mktemp("c0"), ints)
    c = O.ctx(p, 13)
    c.hb_prepare(1e-15)
    return ints, dets, c, p, np.load(os.path.join(GOLDEN, "golden_config0.npz"))


@pytest.fixture(scope="module")
def h2o():
    c = O.ctx(H2O_STO3G, 7)
    c.hb_prepare(1e-15)
    return c, fci_dets(7, 5, 5), np.load(os.path.join(GOLDEN, "golden_h2o.npz"))


# ----------------------------------------------------------------------------- golden (reference build)

def test_golden_config0_tables_and_rows(c0):
    ints, dets, c, _, g = c0
    hb = c.hb_export()
    assert digest(hb["row_offsets"], hb["a"], hb["b"], hb["w"]) == str(g["c0_hb"])
    assert digest(c.ham_diag(dets)) == str(g["c0_diag"])
    for tag, (hbm, eps) in {"full": (False, 1e-6), "hb6": (True, 1e-6), "hb2": (True, 5e-2)}.items():
        d = c.local_connections(dets[:1500], 13, hbm, eps)
        assert d["values"].size == int(g[f"c0_conn_{tag}_nnz"])
        assert conn_digest(d) == str(g[f"c0_conn_{tag}"]), tag
        for r in (0, 7, 1499):
            a, b = d["offsets"][r], d["offsets"][r + 1]
            assert np.array_equal(d["dets"][a:b], g[f"c0_conn_{tag}_row{r}_dets"])
            assert d["values"][a:b].tobytes() == g[f"c0_conn_{tag}_row{r}_vals"].tobytes()


def test_golden_config0_blocks_eff_energy(c0):
    ints, dets, c, _, g = c0
    V = dets[:3000]
    assert coo_digest(c.ham_vv(V, 13)) == str(g["c0_vv"])
    blk = c.ham_conn(V[:300], 13, True, 1e-3)
    assert coo_digest(blk["H_VV"]) == str(g["c0_conn_hb_vv"])
    assert coo_digest(blk["H_VP"]) == str(g["c0_conn_hb_vp"])
    assert digest(blk["det_P"]) == str(g["c0_conn_hb_P"])
    rng = np.random.Generator(np.random.PCG64(int(g["c0_amp_psi_seed"])))
    psi = rng.standard_normal(300) * 0.1
    amp = c.ham_conn_amp(V[:300], psi, 13, 1e-3)
    assert coo_digest(amp["H_VV"]) == str(g["c0_amp_vv"])
    assert coo_digest(amp["H_VP"]) == str(g["c0_amp_vp"])
    assert digest(amp["det_P"]) == str(g["c0_amp_P"])
    hpp = c.ham_diag(amp["det_P"])
    e_ref = float(hpp.min() - 1.0)
    for reg in ("sigma", "linear_shift"):
        assert coo_digest(O.ham_eff(amp["H_VV"], amp["H_VP"], hpp, e_ref, reg, 1e-6)) == str(g[f"c0_eff_{reg}"])
    coef = rng.standard_normal(V.shape[0])
    coef /= np.linalg.norm(coef)
    # the reference reduces rows in OpenMP order: equal within 1e-12 relative
    for tag, hbm in (("full", False), ("hb", True)):
        e = c.variational_energy(V, coef, 13, hbm, 1e-4)
        assert abs(e - float(g[f"c0_energy_{tag}"])) <= 1e-12 * abs(float(g[f"c0_energy_{tag}"]))


def test_golden_config0_explicit_p_block(c0):
    """get_ham_block with an explicit external set (build_ham.cpp:402-460,
    bridge.cpp:469-507): P disjoint from V, and P overlapping V with members
    that connect to nothing; digests written from the reference build."""
    ints, dets, c, _, g = c0
    V = dets[:300]
    blk = c.ham_conn(V, 13, True, 1e-3)
    assert digest(blk["det_P"]) == str(g["c0_conn_hb_P"])
    Pd, Po = explicit_p_sets(V, blk["det_P"], dets[5000:5200], seed=5)
    assert np.intersect1d(Pd.view("V16"), V.view("V16")).size == 0
    assert np.intersect1d(Po.view("V16"), V.view("V16")).size > 0
    for tag, P in (("disjoint", Pd), ("overlap", Po)):
        b = c.ham_block(V, P, 13)
        assert b["H_VP"]["val"].size == int(g[f"c0_block_{tag}_vp_nnz"]), tag
        assert coo_digest(b["H_VV"]) == str(g[f"c0_block_{tag}_vv"]), tag
        assert coo_digest(b["H_VP"]) == str(g[f"c0_block_{tag}_vp"]), tag
        assert digest(b["det_P"]) == str(g[f"c0_block_{tag}_P"]), tag
    # duplicates in the external set are refused, as DetMap::from_ordered(.., true) does
    with pytest.raises(RuntimeError, match="duplicate determinant"):
        c.ham_block(V, np.concatenate([Pd[:10], Pd[3:4]]), 13)


def test_explicit_p_block_dense_closed_form(h2o):
    """H_VV and H_VP of get_ham_block equal the dense FCI Hamiltonian restricted
    to V x V and V x (P minus V), entries with |H| <= 1e-15 dropped; a ket of P
    that lies in V contributes to H_VV only (build_ham.cpp:97-105).  The dense H
    is the reference-pinned ham_vv over the whole FCI space."""
    c, F, g = h2o
    vv = c.ham_vv(F, 7)
    H = np.zeros((F.shape[0],) * 2)
    H[vv["row"], vv["col"]] = vv["val"]
    rng = np.random.default_rng(41)
    perm = rng.permutation(F.shape[0])
    iv, ip = perm[:60], perm[40:200]
    b = c.ham_block(F[iv], F[ip], 7)
    assert coo_digest(b["H_VV"]) == str(g["h2o_block_vv"])
    assert coo_digest(b["H_VP"]) == str(g["h2o_block_vp"])
    assert np.array_equal(b["det_P"], F[ip])
    Dvv = H[np.ix_(iv, iv)]
    Dvp = H[np.ix_(iv, ip)]
    Dvp[:, np.isin(ip, iv)] = 0.0
    for D, coo, shape in ((Dvv, b["H_VV"], (60, 60)), (Dvp, b["H_VP"], (60, 160))):
        r, cc = np.nonzero(np.abs(D) > 1e-15)
        order = np.lexsort((cc, r))
        assert tuple(coo["shape"]) == shape
        assert np.array_equal(coo["row"].astype(np.int64), r[order])
        assert np.array_equal(coo["col"].astype(np.int64), cc[order])
        assert coo["val"].tobytes() == D[r[order], cc[order]].tobytes()
    # and on a tiny system against second quantisation from first principles
    from detnqs_amd import synth
    ints = synth.make_integrals(5, 4, 6, seed=23)
    ct = O.ctx(ints=ints)
    T = fci_dets(5, 2, 2)
    Hsq = _fock_hamiltonian(ints, 5, T)
    iv, ip = np.arange(0, 100, 3), np.arange(1, 100, 2)
    b = ct.ham_block(T[iv], T[ip], 5)
    Dvp = np.zeros((iv.size, ip.size))
    Dvp[b["H_VP"]["row"], b["H_VP"]["col"]] = b["H_VP"]["val"]
    exp = Hsq[np.ix_(iv, ip)].copy()
    exp[:, np.isin(ip, iv)] = 0.0
    assert np.allclose(Dvp, exp, atol=1e-12, rtol=0)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
def test_explicit_p_block_equal_to_reference_build(c0, h2o):
    R = Lib("ref")
    ints, dets, c, path, _ = c0
    cr = R.ctx(path, 13)
    V = dets[:300]
    blk = c.ham_conn(V, 13, True, 1e-3)
    Pd, Po = explicit_p_sets(V, blk["det_P"], dets[5000:5200], seed=5)
    for P in (Pd, Po, None, np.zeros((0, 2), dtype=np.uint64), dets[:5]):
        a, b = c.ham_block(V, P, 13), cr.ham_block(V, P, 13)
        for k in ("H_VV", "H_VP"):
            assert coo_digest(a[k]) == coo_digest(b[k]), k
        assert np.array_equal(a["det_P"], b["det_P"])
    ch, F, _ = h2o
    crh = R.ctx(H2O_STO3G, 7)
    rng = np.random.default_rng(9)
    for _ in range(3):
        perm = rng.permutation(F.shape[0])
        V, P = F[perm[:80]], F[perm[50:300]]
        a, b = ch.ham_block(V, P, 7), crh.ham_block(V, P, 7)
        for k in ("H_VV", "H_VP"):
            assert coo_digest(a[k]) == coo_digest(b[k]), k


def test_golden_h2o(h2o):
    c, F, g = h2o
    assert conn_digest(c.local_connections(F, 7, False, 1e-4)) == str(g["h2o_conn_full"])
    assert conn_digest(c.local_connections(F, 7, True, 1e-4)) == str(g["h2o_conn_hb"])
    assert coo_digest(c.ham_vv(F, 7)) == str(g["h2o_vv"])


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
def test_equal_to_reference_build_random(c0):
    from detnqs_amd import synth
    R = Lib("ref")
    for seed in (3, 4):
        ints = synth.make_integrals(9, 6, 12, seed=seed)
        rng = np.random.default_rng(seed)
        dets = np.unique(synth._random_excitations(rng, 9, 3, 3, rng.integers(1, 6, size=400)), axis=0)
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            p = write_tmp_fcidump(td, ints)
            co, cr = O.ctx(p, 9), R.ctx(p, 9)
        co.hb_prepare(1e-15)
        cr.hb_prepare(1e-15)
        for hbm, eps in ((False, 1e-6), (True, 0.02)):
            a, b = co.local_connections(dets, 9, hbm, eps), cr.local_connections(dets, 9, hbm, eps)
            assert conn_digest(a) == conn_digest(b)
        assert coo_digest(co.ham_vv(dets, 9)) == coo_digest(cr.ham_vv(dets, 9))


# ----------------------------------------------------------------------------- reference KATs

def test_phase_kats():
    """tests/cpp/test_determinant.cpp:78-149 (ket a={0,1}, b={0,1})."""
    ket = [0b0011, 0b0011]
    assert O.analyze(ket, ket) == (0, 0, 0, 1.0)
    assert O.analyze([0b0110, 0b0011], ket) == (1, 1, 0, -1.0)
    assert O.analyze([0b0011, 0b0110], ket) == (1, 0, 1, -1.0)
    assert O.analyze([0b1100, 0b0011], ket) == (2, 2, 0, 1.0)
    assert O.analyze([0b0110, 0b1001], ket)[0] == 2 and O.analyze([0b0110, 0b1001], ket)[3] == -1.0
    d, _, _, ph = O.analyze([0b1100, 0b0101], ket)
    assert d > 2 and ph == 0.0


def test_collector_counts_closed_form():
    """tests/cpp/test_determinant.cpp:155-189: 8 singles, 18 doubles, 26 total for n=4, 2a2b;
    checked through get_local_connections with a Hamiltonian of nonzero couplings."""
    from detnqs_amd import synth
    ints = synth.make_integrals(4, 4, 6, seed=5)
    c = O.ctx(ints=ints)
    d = c.local_connections(np.array([[0b0011, 0b0011]], dtype=np.uint64), 4, False)
    conn = d["dets"]
    assert conn.shape[0] == 1 + 26
    assert np.unique(conn, axis=0).shape[0] == 27
    # general closed form S = na*va + nb*vb, D = C(na,2)C(va,2) + C(nb,2)C(vb,2) + na*va*nb*vb
    from math import comb
    ints = synth.make_integrals(8, 6, 10, seed=6)
    c = O.ctx(ints=ints)
    x = np.array([[0b00010111, 0b00001011]], dtype=np.uint64)  # na=4, nb=3
    na, nb, va, vb = 4, 3, 4, 5
    d = c.local_connections(x, 8, False)
    S = na * va + nb * vb
    D = comb(na, 2) * comb(va, 2) + comb(nb, 2) * comb(vb, 2) + na * va * nb * vb
    assert d["dets"].shape[0] == 1 + S + D


def test_brute_force_classification():
    """Every connected det of a row has excitation degree 1 or 2 and every FCI det
    of degree <= 2 appears exactly once (n=6, 2a2b)."""
    from detnqs_amd import synth
    ints = synth.make_integrals(6, 4, 8, seed=8)
    c = O.ctx(ints=ints)
    F = fci_dets(6, 2, 2)
    for x in F[::7]:
        d = c.local_connections(x[None, :], 6, False)
        got = {(int(a), int(b)) for a, b in d["dets"]}
        deg = lambda y: (bin(int(y[0]) ^ int(x[0])).count("1") + bin(int(y[1]) ^ int(x[1])).count("1")) // 2
        exp = {(int(a), int(b)) for a, b in F if deg((a, b)) <= 2}
        assert got == exp and len(got) == d["dets"].shape[0]


def _sq_phase(bra, ket):
    """Maximum-coincidence sign, by definition: bra's ascending orbital list per
    spin with each removed orbital replaced by the added one (sorted pairing),
    sign of the permutation sorting it; alpha and beta strings multiply."""
    par = 0
    for s in (0, 1):
        b, k = int(bra[s]), int(ket[s])
        holes = sorted(i for i in range(64) if (k >> i) & 1 and not (b >> i) & 1)
        parts = sorted(i for i in range(64) if (b >> i) & 1 and not (k >> i) & 1)
        L = sorted(i for i in range(64) if (b >> i) & 1)
        for h, p in zip(holes, parts):
            L[L.index(p)] = h
        par += sum(1 for i in range(len(L)) for j in range(i + 1, len(L)) if L[i] > L[j])
    return -1.0 if par & 1 else 1.0


def test_phase_brute_force():
    rng = np.random.default_rng(11)
    F = fci_dets(7, 3, 3)
    for _ in range(600):
        bra, ket = F[rng.integers(len(F))], F[rng.integers(len(F))]
        deg, _, _, ph = O.analyze(bra, ket)
        if deg > 2:
            assert ph == 0.0
        else:
            assert ph == _sq_phase(bra, ket)


def _fock_hamiltonian(ints, n, dets):
    """Dense H from second quantisation: H = sum h_pq a+_p a_q + 1/2 sum <pq|rs> a+_p a+_q a_s a_r,
    spin orbitals ordered alpha block then beta block (Jordan–Wigner signs)."""
    h1 = ints.h1_dense()
    g = ints.h2_dense()  # (pq|rs)
    idx = {(int(a), int(b)): i for i, (a, b) in enumerate(dets)}
    nso = 2 * n

    def occ_vec(d):
        a, b = int(d[0]), int(d[1])
        return [(a >> i) & 1 for i in range(n)] + [(b >> i) & 1 for i in range(n)]

    def annihilate(v, p):
        if not v[p]:
            return None, 0
        s = sum(v[:p]) & 1
        w = list(v)
        w[p] = 0
        return w, s

    def create(v, p):
        if v[p]:
            return None, 0
        s = sum(v[:p]) & 1
        w = list(v)
        w[p] = 1
        return w, s

    def key(v):
        a = sum(1 << i for i in range(n) if v[i])
        b = sum(1 << i for i in range(n) if v[n + i])
        return (a, b)

    def spin(p):
        return 0 if p < n else 1

    def mo(p):
        return p % n

    H = np.zeros((len(dets), len(dets)))
    for j, d in enumerate(dets):
        v = occ_vec(d)
        for p in range(nso):
            for q in range(nso):
                if spin(p) != spin(q):
                    continue
                w, s1 = annihilate(v, q)
                if w is None:
                    continue
                w, s2 = create(w, p)
                if w is None:
                    continue
                H[idx[key(w)], j] += (-1) ** (s1 + s2) * h1[mo(p), mo(q)]
        for p in range(nso):
            for q in range(nso):
                for r in range(nso):
                    for s in range(nso):
                        if spin(p) != spin(r) or spin(q) != spin(s):
                            continue
                        w, t1 = annihilate(v, r)
                        if w is None:
                            continue
                        w, t2 = annihilate(w, s)
                        if w is None:
                            continue
                        w, t3 = create(w, q)
                        if w is None:
                            continue
                        w, t4 = create(w, p)
                        if w is None:
                            continue
                        # <pq|rs> = (pr|qs)
                        H[idx[key(w)], j] += 0.5 * (-1) ** (t1 + t2 + t3 + t4) * g[mo(p), mo(r), mo(q), mo(s)]
    return H


def test_slater_condon_vs_second_quantisation():
    from detnqs_amd import synth
    ints = synth.make_integrals(4, 4, 6, seed=21)
    c = O.ctx(ints=ints)
    F = fci_dets(4, 2, 2)
    Hsq = _fock_hamiltonian(ints, 4, F)
    vv = c.ham_vv(F, 4)
    H = np.zeros_like(Hsq)
    H[vv["row"], vv["col"]] = vv["val"]
    assert np.allclose(H, Hsq, atol=1e-12, rtol=0)


def test_h2o_fci_energy_and_hermiticity(h2o):
    """E_FCI(H2O/STO-3G) = -75.01240139 within 1e-7 relative (test_hamiltonian.cpp:111,282-295)."""
    c, F, _ = h2o
    vv = c.ham_vv(F, 7)
    H = np.zeros((F.shape[0],) * 2)
    H[vv["row"], vv["col"]] = vv["val"]
    assert np.abs(H - H.T).max() < 1e-12
    w, U = np.linalg.eigh(H)
    e = w[0] + c.e_nuc
    assert abs(e - E_FCI_H2O_STO3G) <= 1e-7 * abs(E_FCI_H2O_STO3G)
    # zero-variance: E_loc(x) = E_FCI - e_nuc for every x of the FCI eigenvector
    psi = U[:, 0]
    _, sigma = c.variational_energy(F, psi, 7, False, 0.0, want_sigma=True)
    big = np.abs(psi) > 1e-6
    assert np.allclose(sigma[big] / psi[big], w[0], atol=1e-9)
    # HB discovery from HF: 20 <= |P| <= 140 and HF not in P (test_hamiltonian.cpp:201-225)
    hf = np.array([[0b11111, 0b11111]], dtype=np.uint64)
    blk = c.ham_conn(hf, 7, True, 1e-6)
    assert 20 <= blk["size_P"] <= 140
    assert not np.any(np.all(blk["det_P"] == hf[0], axis=1))
    # effective Hamiltonian lowers the HF energy (test_hamiltonian.cpp:260-279)
    blk = c.ham_conn(hf, 7, True, 1e-5)
    hpp = c.ham_diag(blk["det_P"])
    e_hf = c.ham_diag(hf)[0]
    eff = O.ham_eff(blk["H_VV"], blk["H_VP"], hpp, e_hf, "sigma", 1e-10)
    assert eff["val"][0] < blk["H_VV"]["val"][0]


def test_h2_integrals_known_values():
    """tests/cpp/test_integral.cpp:65-117 (H2/6-31G)."""
    c = O.ctx(H2_631G, 4)
    assert c.n_elec == 2 and abs(c.e_nuc - 0.26458860546) < 1e-10
    for (p, q), v in {(0, 0): -0.801833061063, (1, 1): -0.680884820395, (2, 2): 0.202998161132,
                      (3, 3): 0.265010250063, (2, 0): -0.074391365158, (3, 1): 0.120229115266}.items():
        assert abs(c.h1e(p, q) - v) < 1e-10 and c.h1e(q, p) == c.h1e(p, q)
    assert abs(c.h2e(0, 0, 0, 0) - 0.422806268972) < 1e-10
    assert abs(c.h2e(1, 1, 0, 0) - 0.410534684646) < 1e-10
    assert abs(c.h2e(1, 0, 1, 0) - 0.163762656748) < 1e-10
    ref = c.h2e(1, 0, 3, 2)
    assert abs(ref + 0.180315206010) < 1e-10
    for p, q, r, s in [(1, 0, 3, 2), (0, 1, 2, 3), (3, 2, 1, 0), (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0),
                       (0, 1, 3, 2), (1, 0, 2, 3)]:
        assert c.h2e(p, q, r, s) == ref


def test_hb_invariants(c0):
    ints, dets, c, _, _ = c0
    hb = c.hb_export()
    off, a, b, w = hb["row_offsets"], hb["a"], hb["b"], hb["w"]
    nso = 26
    G = ints.h2_dense()
    rid = 0
    for i in range(1, nso):
        for j in range(i):
            s, e = int(off[rid]), int(off[rid + 1])
            assert np.all(np.diff(w[s:e]) <= 0)
            assert np.all(a[s:e] > b[s:e])
            # brute-force membership
            exp = 0
            for x in range(1, nso):
                for y in range(x):
                    if len({x, y, i, j}) < 4:
                        continue
                    if not (((x & 1) == (i & 1) and (y & 1) == (j & 1)) or ((x & 1) == (j & 1) and (y & 1) == (i & 1))):
                        continue
                    exp += 1
            assert e - s == exp  # no weight below 1e-15 in the synthetic integrals
            rid += 1
            if rid > 60:
                return


# ----------------------------------------------------------------------------- slogdet oracle

def test_oracle_slogdet_library_and_closed_forms():
    rng = np.random.default_rng(0)
    for n in (2, 5, 7, 12, 24, 40, 64):
        A = rng.standard_normal((50, n, n))
        s, la = O.slogdet(A)
        ns, nla = np.linalg.slogdet(A)
        assert np.array_equal(s, ns)
        assert np.allclose(la, nla, rtol=1e-11, atol=1e-11)
    T = np.triu(rng.standard_normal((10, 6, 6)))
    s, la = O.slogdet(T)
    d = np.diagonal(T, axis1=1, axis2=2)
    assert np.allclose(la, np.log(np.abs(d)).sum(1), rtol=1e-14) and np.array_equal(s, np.prod(np.sign(d), 1))
    P = np.eye(6)[rng.permutation(6)][None]
    s, la = O.slogdet(P)
    assert la[0] == 0.0 and s[0] == round(np.linalg.det(P[0]))
    Z = np.zeros((1, 4, 4))
    s, la = O.slogdet(Z)
    assert s[0] == 0.0 and la[0] == -np.inf


def test_oracle_slogdet_bwd_inverse_and_fd():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((20, 7, 7)) + 7 * np.eye(7)
    cot = rng.standard_normal(20)
    G = O.slogdet_bwd(A, cot)
    ref = cot[:, None, None] * np.transpose(np.linalg.inv(A), (0, 2, 1))
    assert np.allclose(G, ref, rtol=1e-12, atol=1e-13)
    h = 1e-6
    for b, i, j in [(0, 1, 2), (5, 6, 6), (19, 0, 3)]:
        Ap, Am = A[b].copy(), A[b].copy()
        Ap[i, j] += h
        Am[i, j] -= h
        fd = (O.slogdet(Ap)[1] - O.slogdet(Am)[1]) / (2 * h) * cot[b]
        assert abs(fd - G[b, i, j]) < 1e-7
    S = np.zeros((1, 5, 5))
    assert np.all(O.slogdet_bwd(S, np.ones(1)) == 0)


def test_oracle_ham_eff_dense_closed_form(h2o):
    """H_eff = H_VV + H_VP diag(dinv) H_VP^T with the skip rule inactive."""
    c, F, _ = h2o
    V = F[:40]
    blk = c.ham_conn(V, 7, True, 1e-6)
    nv, npp = V.shape[0], blk["size_P"]
    hpp = c.ham_diag(blk["det_P"])
    e_ref = float(c.ham_diag(V).min() - 0.5)
    eff = O.ham_eff(blk["H_VV"], blk["H_VP"], hpp, e_ref, "linear_shift", 1e-12)
    Hvv = np.zeros((nv, nv))
    Hvv[blk["H_VV"]["row"], blk["H_VV"]["col"]] = blk["H_VV"]["val"]
    Hvp = np.zeros((nv, npp))
    Hvp[blk["H_VP"]["row"], blk["H_VP"]["col"]] = blk["H_VP"]["val"]
    d = e_ref - hpp
    dinv = 1.0 / (d + np.copysign(1e-12, d))
    colnorm = (Hvp ** 2).sum(0)
    keep = np.abs(dinv) * colnorm >= 1e-12
    dense = Hvv + (Hvp[:, keep] * dinv[keep]) @ Hvp[:, keep].T
    E = np.zeros((nv, nv))
    E[eff["row"], eff["col"]] = eff["val"]
    assert np.allclose(E, dense, atol=1e-12, rtol=1e-12)


def test_tolerant_detmap_coefficient_index(h2o):
    """compute_variational_energy with duplicates inside V: the tolerant DetMap
    numbers first occurrences consecutively (det_space.cpp:40-54) and the
    energy reads coeffs[that number] (local_conn.cpp:192-194), so a partner
    after a duplicate takes another row's coefficient.  Dense closed form on
    H2O/STO-3G (H from the reference-pinned ham_vv over the FCI space)."""
    c, F, _ = h2o
    vv = c.ham_vv(F, 7)
    H = np.zeros((F.shape[0],) * 2)
    H[vv["row"], vv["col"]] = vv["val"]
    rng = np.random.default_rng(8)
    perm = rng.permutation(F.shape[0])
    rows = np.concatenate([perm[:60], perm[10:25], perm[60:130], perm[:3]])  # duplicates mid-V and at the end
    V = F[rows]
    coef = rng.standard_normal(rows.size)
    coef[[4, 70]] = 1e-7  # skipped rows
    first = {}
    for k, r in enumerate(rows):
        first.setdefault(int(r), len(first))  # rank among first occurrences
    uniq = np.array(list(first.keys()))
    rank = np.array(list(first.values()))
    e = 0.0
    for k, r in enumerate(rows):
        if coef[k] ** 2 < 1e-12:
            continue
        h = H[r, uniq].copy()
        own = np.flatnonzero(uniq == r)[0]
        h[own] = 0.0
        sig = H[r, r] * coef[k] + h @ coef[rank]
        e += coef[k] * sig
    got = c.variational_energy(V, coef, 7, False, 1e-6)
    assert abs(got - e) <= 1e-12 * abs(e)
    # the row-index rule gives a different energy here, so the test tells them apart
    e_rows = 0.0
    for k, r in enumerate(rows):
        if coef[k] ** 2 < 1e-12:
            continue
        h = H[r, uniq].copy()
        h[np.flatnonzero(uniq == r)[0]] = 0.0
        firstrow = np.array([np.flatnonzero(rows == u)[0] for u in uniq])
        e_rows += coef[k] * (H[r, r] * coef[k] + h @ coef[firstrow])
    assert abs(e_rows - e) > 1e-6 * abs(e)
    if ref_available():
        R = Lib("ref").ctx(H2O_STO3G, 7)
        assert abs(R.variational_energy(V, coef, 7, False, 1e-6) - got) <= 1e-12 * abs(got)


def test_oracle_energy_dense(h2o):
    c, F, _ = h2o
    rng = np.random.default_rng(3)
    V = F[rng.permutation(F.shape[0])[:150]]
    coef = rng.standard_normal(150)
    coef /= np.linalg.norm(coef)
    vv = c.ham_vv(V, 7)
    H = np.zeros((150, 150))
    H[vv["row"], vv["col"]] = vv["val"]
    skip = coef ** 2 < 1e-12
    e = c.variational_energy(V, coef, 7, False, 1e-6)
    c2 = coef.copy()
    assert abs(e - c2[~skip] @ (H @ coef)[~skip]) < 1e-12


def test_sigma_rows_equals_full_sigma(c0):
    """The row-subset sigma is the full energy's sigma on those rows, bit for bit
    (same code per row), with duplicates inside V and skipped rows."""
    ints, dets, c, _, _ = c0
    V = np.concatenate([dets[:700], dets[100:200], dets[700:1200]])
    coef = np.random.default_rng(4).standard_normal(V.shape[0])
    coef[::37] = 1e-7
    for hbm, eps in ((True, 1e-5), (False, 0.0)):
        _, sig = c.variational_energy(V, coef, 13, hbm, eps, want_sigma=True)
        rows = np.array([0, 5, 37, 699, 700, 750, 1299])
        got = c.sigma_rows(V, coef, rows, 13, hbm, eps)
        assert got.tobytes() == sig[rows].tobytes()
    with pytest.raises(RuntimeError, match="out of range"):
        c.sigma_rows(V, coef, np.array([V.shape[0]]), 13, True, 1e-5)
