"""Seeded synthetic inputs for the local-energy path.

Everything here is plain NumPy (PCG64) so the same seed gives the same bits on
any host.  The integrals are stored in the packed layout the DetNQS core uses
(``csrc/detnqs/integral/integral_mo.hpp:30-49``):

* ``h1e[key(p,q)]`` with ``key(p,q) = max(p,q)*(max(p,q)+1)/2 + min(p,q)``;
* ``h2e[key(key(p,q), key(r,s))]`` (chemists' ``(pq|rs)``), 8-fold symmetric.

``write_fcidump`` serialises them with ``%.17e`` so that the reference FCIDUMP
parser (``integral_mo.cpp:157-182``) reads back bit-identical doubles.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Iterable

import numpy as np


def pair_key(p, q):
    """Triangular key of an unordered pair (integral_mo.hpp:30-33)."""
    p = np.asarray(p, dtype=np.int64)
    q = np.asarray(q, dtype=np.int64)
    hi = np.maximum(p, q)
    lo = np.minimum(p, q)
    return hi * (hi + 1) // 2 + lo


def h2e_key(p, q, r, s):
    """Compound key of (pq|rs) (integral_mo.hpp:40-49)."""
    return pair_key(pair_key(p, q), pair_key(r, s))


@dataclasses.dataclass
class Integrals:
    n_orb: int
    n_elec: int
    h1e: np.ndarray  # packed, n1e
    h2e: np.ndarray  # packed, n2e
    e_nuc: float

    @property
    def n1e(self) -> int:
        return self.n_orb * (self.n_orb + 1) // 2

    def h1_dense(self) -> np.ndarray:
        n = self.n_orb
        p, q = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        return self.h1e[pair_key(p, q)]

    def h2_dense(self) -> np.ndarray:
        n = self.n_orb
        idx = np.arange(n)
        p, q, r, s = np.meshgrid(idx, idx, idx, idx, indexing="ij")
        return self.h2e[h2e_key(p, q, r, s)]


def make_integrals(n_orb: int, n_elec: int, n_chol: int, seed: int,
                   sigma: float = 0.3, e_nuc: float | None = None) -> Integrals:
    """Synthetic FP64 integrals.

    h_pq = N(0,1)*exp(-|p-q|/4) (upper triangle, symmetrised), diagonal -|N(2,1)|;
    (pq|rs) = sum_{L<n_chol} L^L_pq L^L_rs with L^L = (M+M^T)/2, M ~ N(0, sigma).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    n = n_orb
    g = rng.standard_normal((n, n))
    pp, qq = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    h = g * np.exp(-np.abs(pp - qq) / 4.0)
    h = np.triu(h, 1)
    h = h + h.T
    h[np.arange(n), np.arange(n)] = -np.abs(rng.normal(2.0, 1.0, size=n))
    M = rng.normal(0.0, sigma, size=(n_chol, n, n))
    L = 0.5 * (M + np.transpose(M, (0, 2, 1)))
    # Packed pair vectors: Lp[L, key(p,q)] for p>=q.
    n1e = n * (n + 1) // 2
    pi, qi = np.tril_indices(n)
    Lp = L[:, pi, qi]  # (n_chol, n1e) ordered by key since tril_indices is row-major
    keys = pair_key(pi, qi)
    order = np.argsort(keys)
    Lp = Lp[:, order]
    assert np.array_equal(np.sort(keys), np.arange(n1e))
    V = Lp.T @ Lp  # (n1e, n1e), symmetric PSD
    a, b = np.tril_indices(n1e)
    h2e = np.empty(n1e * (n1e + 1) // 2)
    h2e[a * (a + 1) // 2 + b] = V[a, b]
    h1e = np.empty(n1e)
    h1e[pair_key(pi, qi)] = h[pi, qi]
    if e_nuc is None:
        e_nuc = float(rng.uniform(5.0, 30.0))
    return Integrals(n_orb=n, n_elec=n_elec, h1e=h1e, h2e=h2e, e_nuc=float(e_nuc))


def write_fcidump(path: str, ints: Integrals) -> None:
    """FCIDUMP writer (1-based indices, %.17e values; integral_mo.cpp:157-182)."""
    n = ints.n_orb
    n1e = ints.n1e
    lines = [f"&FCI NORB={n},NELEC={ints.n_elec},MS2=0,\n", "&END\n"]
    # (pq|rs) with p>=q, r>=s, pq>=rs
    pi, qi = np.tril_indices(n)
    k1 = pair_key(pi, qi)
    order = np.argsort(k1)
    pi, qi = pi[order], qi[order]
    a, b = np.tril_indices(n1e)
    vals = ints.h2e[a * (a + 1) // 2 + b]
    P, Q, R, S = pi[a] + 1, qi[a] + 1, pi[b] + 1, qi[b] + 1
    body = "".join(
        f"{v:.17e} {p} {q} {r} {s}\n" for v, p, q, r, s in zip(vals.tolist(), P.tolist(), Q.tolist(), R.tolist(), S.tolist())
    )
    lines.append(body)
    for p in range(n):
        for q in range(p + 1):
            lines.append(f"{ints.h1e[pair_key(p, q)]:.17e} {p + 1} {q + 1} 0 0\n")
    lines.append(f"{ints.e_nuc:.17e} 0 0 0 0\n")
    with open(path, "w") as f:
        f.writelines(lines)


def read_fcidump(path: str, n_orb: int) -> Integrals:
    """Minimal FCIDUMP reader with the reference's rules (integral_mo.cpp:38-182):
    '&' lines and lines starting with ! * C c # are skipped, Fortran D exponents
    are accepted, '0 0 0 0' is e_nuc, 'p q 0 0' is h_pq, later lines overwrite."""
    n1e = n_orb * (n_orb + 1) // 2
    h1e = np.zeros(n1e)
    h2e = np.zeros(n1e * (n1e + 1) // 2)
    e_nuc = 0.0
    n_elec = 0
    namelist = ""
    in_nl = False
    with open(path) as f:
        text = f.read().splitlines()
    for line in text:
        if not in_nl and ("&FCI" in line or "&fci" in line):
            in_nl = True
        if in_nl:
            namelist += line + " "
            if "&END" in line or "&end" in line:
                break
    up = namelist.upper()
    pos = up.find("NELEC=")
    if pos >= 0:
        pos += 6
        while pos < len(up) and up[pos].isspace():
            pos += 1
        end = pos
        while end < len(up) and (up[end].isdigit() or up[end] in "+-"):
            end += 1
        n_elec = int(up[pos:end])
    for line in text:
        if not line or line[0] in "!*Cc#" or "&" in line:
            continue
        parts = line.replace("D", "E").replace("d", "e").split()
        if len(parts) < 5:
            continue
        try:
            v = float(parts[0])
            p, q, r, s = (int(x) for x in parts[1:5])
        except ValueError:
            continue
        if p == 0 and q == 0 and r == 0 and s == 0:
            e_nuc = v
        elif r == 0 and s == 0:
            h1e[int(pair_key(p - 1, q - 1))] = v
        else:
            h2e[int(h2e_key(p - 1, q - 1, r - 1, s - 1))] = v
    return Integrals(n_orb=n_orb, n_elec=n_elec, h1e=h1e, h2e=h2e, e_nuc=e_nuc)


# ---------------------------------------------------------------------------
# Determinants
# ---------------------------------------------------------------------------

def hf_det(n_alpha: int, n_beta: int) -> np.ndarray:
    """Lowest orbitals occupied per spin (system.py:215-226)."""
    return np.array([(1 << n_alpha) - 1, (1 << n_beta) - 1], dtype=np.uint64)


def _bits(mask: int) -> list[int]:
    out = []
    while mask:
        low = mask & -mask
        out.append(low.bit_length() - 1)
        mask ^= low
    return out


def singles_doubles(det: np.ndarray, n_orb: int) -> np.ndarray:
    """Singles then doubles of ``det`` in the reference generation order
    (det_ops.hpp:123-268): singles a, singles b, doubles aa, bb, ab."""
    a, b = int(det[0]), int(det[1])
    full = (1 << n_orb) - 1
    out: list[tuple[int, int]] = []
    for spin in (0, 1):
        occ = a if spin == 0 else b
        virt = ~occ & full
        for h in _bits(occ):
            for p in _bits(virt):
                nw = occ ^ (1 << h) ^ (1 << p)
                out.append((nw, b) if spin == 0 else (a, nw))
    for spin in (0, 1):
        occ = a if spin == 0 else b
        virt = ~occ & full
        oc, vi = _bits(occ), _bits(virt)
        for i1 in range(len(oc)):
            for i2 in range(i1 + 1, len(oc)):
                for j1 in range(len(vi)):
                    for j2 in range(j1 + 1, len(vi)):
                        nw = occ ^ (1 << oc[i1]) ^ (1 << oc[i2]) ^ (1 << vi[j1]) ^ (1 << vi[j2])
                        out.append((nw, b) if spin == 0 else (a, nw))
    va, vb = ~a & full, ~b & full
    for ha in _bits(a):
        for pa in _bits(va):
            na = a ^ (1 << ha) ^ (1 << pa)
            for hb in _bits(b):
                for pb in _bits(vb):
                    out.append((na, b ^ (1 << hb) ^ (1 << pb)))
    return np.array(out, dtype=np.uint64).reshape(-1, 2)


def _random_excitations(rng, n_orb: int, n_a: int, n_b: int, ranks: np.ndarray) -> np.ndarray:
    """Random dets at the given total excitation ranks from HF; the split of a
    rank over (r_alpha, r_beta) is uniform over the feasible values."""
    N = ranks.shape[0]
    va, vb = n_orb - n_a, n_orb - n_b
    lo = np.maximum(0, ranks - min(n_b, vb))
    hi = np.minimum(ranks, min(n_a, va))
    ra = lo + np.floor(rng.random(N) * (hi - lo + 1)).astype(np.int64)
    rb = ranks - ra

    def spin_words(n_occ, n_virt, r):
        # choose r holes among the n_occ lowest orbitals and r particles among
        # the remaining n_virt, by ranking random keys
        hk = np.argsort(rng.random((N, n_occ)), axis=1)
        pk = np.argsort(rng.random((N, n_virt)), axis=1)
        hol = (np.arange(n_occ)[None, :] < r[:, None])
        par = (np.arange(n_virt)[None, :] < r[:, None])
        word = np.full(N, (1 << n_occ) - 1, dtype=np.uint64)
        hsh = np.left_shift(np.uint64(1), hk.astype(np.uint64))
        psh = np.left_shift(np.uint64(1), (pk + n_occ).astype(np.uint64))
        hbits = np.bitwise_or.reduce(np.where(hol, hsh, np.uint64(0)), axis=1)
        pbits = np.bitwise_or.reduce(np.where(par, psh, np.uint64(0)), axis=1)
        return (word ^ hbits) | pbits

    A = spin_words(n_a, va, ra)
    B = spin_words(n_b, vb, rb)
    return np.stack([A, B], axis=1)


def _det_keys(dets: np.ndarray, n_orb: int):
    """Hashable scalar keys of (alpha, beta) rows (one uint64 when 2*n_orb <= 64)."""
    if 2 * n_orb <= 64:
        return dets[:, 0] | (dets[:, 1] << np.uint64(n_orb))
    return np.ascontiguousarray(dets).view(np.dtype((np.void, 16))).ravel()


def _unique_in_order(dets: np.ndarray, n_orb: int, exclude: np.ndarray | None = None) -> np.ndarray:
    """First occurrences of each det, in input order, minus those in ``exclude``."""
    view = _det_keys(dets, n_orb)
    _, first = np.unique(view, return_index=True)
    keep = np.zeros(dets.shape[0], dtype=bool)
    keep[np.sort(first)] = True
    if exclude is not None and exclude.shape[0]:
        keep &= ~np.isin(view, _det_keys(exclude, n_orb))
    return dets[keep]


def config0_dets(seed: int = 0, n_orb: int = 13, n_a: int = 5, n_b: int = 5,
                 n_high: int = 5000) -> np.ndarray:
    """HF + all singles + all doubles + n_high unique random excitations of rank 3..10."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hf = hf_det(n_a, n_b)
    sd = singles_doubles(hf, n_orb)
    base = np.concatenate([hf[None, :], sd], axis=0)
    out = []
    have = 0
    while have < n_high:
        ranks = rng.integers(3, 11, size=2 * n_high)
        cand = _random_excitations(rng, n_orb, n_a, n_b, ranks)
        prev = np.concatenate([base] + out, axis=0)
        cand = _unique_in_order(cand, n_orb, exclude=prev)
        out.append(cand[: n_high - have])
        have += out[-1].shape[0]
    return np.concatenate([base] + out, axis=0)


def inverse_square_ranks(rng, n: int, r_max: int) -> np.ndarray:
    r = np.arange(1, r_max + 1)
    p = 1.0 / r.astype(np.float64) ** 2
    p /= p.sum()
    return rng.choice(r, size=n, p=p)


def config1_dets(n_dets: int = 1_000_000, seed: int = 1, n_orb: int = 28, n_a: int = 7,
                 n_b: int = 7) -> np.ndarray:
    """HF + unique dets with excitation rank r in [1, n_a+n_b], P(r) ~ 1/r^2."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hf = hf_det(n_a, n_b)[None, :]
    chunks = [hf]
    have = 1
    r_max = min(n_a + n_b, (n_orb - n_a) + (n_orb - n_b))
    n_draw = int(n_dets * 1.3) + 64
    while have < n_dets:
        need = n_dets - have
        ranks = inverse_square_ranks(rng, n_draw, r_max)
        cand = _random_excitations(rng, n_orb, n_a, n_b, ranks)
        cand = _unique_in_order(cand, n_orb, exclude=np.concatenate(chunks, axis=0))
        chunks.append(cand[:need])
        have += chunks[-1].shape[0]
        # low ranks saturate (few unique dets exist there): size the next draw
        # from the yield of this one
        yield_ = max(cand.shape[0], 1) / n_draw
        n_draw = int(min(max((n_dets - have) / yield_ * 1.2, 1024), 50_000_000))
    return np.concatenate(chunks, axis=0)


def config1_dets_device(n_dets: int, seed: int, n_orb: int, n_a: int, n_b: int, device=None):
    """config1_dets drawn on the GPU (torch generator, so a different sequence
    from the NumPy one): HF + unique dets with excitation rank r in
    [1, n_a + n_b], P(r) ~ 1/r^2, the rank split uniformly over the feasible
    (r_alpha, r_beta), holes and particles uniform; first occurrences kept in
    draw order.  Returns an (N, 2) int64 CUDA tensor (uint64 bit patterns)."""
    import torch
    dev = torch.device("cuda") if device is None else torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    va, vb = n_orb - n_a, n_orb - n_b
    r_max = min(n_a + n_b, va + vb)
    p = 1.0 / torch.arange(1, r_max + 1, dtype=torch.float64, device=dev) ** 2

    def spin_words(n_occ, n_virt, r):
        N = r.shape[0]
        word = torch.full((N,), (1 << n_occ) - 1, dtype=torch.int64, device=dev)
        if n_occ == 0 or n_virt == 0:
            return word
        hk = torch.argsort(torch.rand((N, n_occ), generator=g, device=dev), dim=1)
        pk = torch.argsort(torch.rand((N, n_virt), generator=g, device=dev), dim=1) + n_occ
        hol = torch.arange(n_occ, device=dev)[None, :] < r[:, None]
        par = torch.arange(n_virt, device=dev)[None, :] < r[:, None]
        one = torch.ones((), dtype=torch.int64, device=dev)
        hbits = torch.where(hol, torch.bitwise_left_shift(one, hk), 0).sum(dim=1)  # distinct bits: sum = or
        pbits = torch.where(par, torch.bitwise_left_shift(one, pk), 0).sum(dim=1)
        return (word ^ hbits) | pbits

    acc = torch.tensor([[(1 << n_a) - 1, (1 << n_b) - 1]], dtype=torch.int64, device=dev)
    n_draw = int(n_dets * 1.3) + 64
    while acc.shape[0] < n_dets:
        ranks = torch.multinomial(p, n_draw, replacement=True, generator=g) + 1
        lo = torch.clamp(ranks - min(n_b, vb), min=0)
        hi = torch.clamp(ranks, max=min(n_a, va))
        ra = lo + torch.floor(torch.rand(n_draw, generator=g, device=dev, dtype=torch.float64)
                              * (hi - lo + 1).to(torch.float64)).to(torch.int64)
        ra = torch.minimum(ra, hi)
        rb = ranks - ra
        cand = torch.stack([spin_words(n_a, va, ra), spin_words(n_b, vb, rb)], dim=1)
        allc = torch.cat([acc, cand])
        uniq, inv = torch.unique(allc, dim=0, return_inverse=True)
        pos = torch.arange(allc.shape[0], device=dev)
        first = torch.full((uniq.shape[0],), allc.shape[0], dtype=torch.int64, device=dev)
        first = first.scatter_reduce(0, inv, pos, reduce="amin")
        new = torch.sort(first[first >= acc.shape[0]]).values
        take = new[: n_dets - acc.shape[0]]
        yield_ = max(int(new.shape[0]), 1) / n_draw
        acc = torch.cat([acc, allc[take]])
        n_draw = int(min(max((n_dets - acc.shape[0]) / yield_ * 1.2, 1024), 50_000_000))
    return acc.contiguous()


def config2_instance(spec, device=None):
    """Integrals and determinants of one batched-form instance (config2_specs):
    integrals from 2 n_orb Cholesky-like vectors, determinants drawn on the GPU."""
    ne = spec.n_elec // 2
    ints = make_integrals(spec.n_orb, spec.n_elec, 2 * spec.n_orb, seed=spec.seed)
    dets = config1_dets_device(spec.n_dets, spec.seed, spec.n_orb, ne, ne, device)
    return ints, dets


def occ_rows(words: np.ndarray, n_occ: int) -> np.ndarray:
    """Occupied orbital indices (ascending) of each uint64 word: (N, n_occ)."""
    w = words.astype(np.uint64)
    N = w.shape[0]
    out = np.empty((N, n_occ), dtype=np.int64)
    for c in range(n_occ):
        low = w & (~w + np.uint64(1))
        out[:, c] = np.log2(low.astype(np.float64)).astype(np.int64)
        w = w ^ low
    return out


def slater_matrices(dets: np.ndarray, n_orb: int, n_a: int, n_b: int, seed: int,
                    noise: float = 0.0) -> np.ndarray:
    """Occupied-row blocks of a random orbital matrix: for each det, the alpha
    block Phi[occ_a, :n_a] and beta block Phi[occ_b, :n_b] (+ noise*N(0,1)).
    Returns (2*N, n, n) with alpha/beta interleaved per det (requires n_a == n_b)."""
    assert n_a == n_b
    rng = np.random.Generator(np.random.PCG64(seed))
    phi = rng.standard_normal((n_orb, n_a))
    oa = occ_rows(dets[:, 0], n_a)
    ob = occ_rows(dets[:, 1], n_b)
    occ = np.stack([oa, ob], axis=1).reshape(-1, n_a)  # (2N, n)
    A = phi[occ]  # (2N, n, n)
    if noise:
        A = A + noise * rng.standard_normal(A.shape)
    return np.ascontiguousarray(A)


def selection_matrices(batch: int, n: int, n_rows: int, seed: int, noise: float = 0.05,
                       near_singular_frac: float = 0.0) -> np.ndarray:
    """A = Phi[rows] + noise*N(0,1) with Phi (n_rows x n), rows sampled without
    replacement per sample; a fraction of samples has one row scaled by 1e-8."""
    rng = np.random.Generator(np.random.PCG64(seed))
    phi = rng.standard_normal((n_rows, n))
    rows = np.argsort(rng.random((batch, n_rows)), axis=1)[:, :n]
    A = phi[rows] + noise * rng.standard_normal((batch, n, n))
    if near_singular_frac > 0:
        k = int(round(batch * near_singular_frac))
        idx = rng.choice(batch, size=k, replace=False)
        r = rng.integers(0, n, size=k)
        A[idx, r, :] *= 1e-8
    return np.ascontiguousarray(A)


def config0(seed: int = 0):
    ints = make_integrals(13, 10, 40, seed=1000 + seed)
    dets = config0_dets(seed)
    return ints, dets


def config1(n_dets: int = 1_000_000, seed: int = 1):
    ints = make_integrals(28, 14, 80, seed=2000 + seed)
    dets = config1_dets(n_dets, seed=seed)
    return ints, dets


@dataclasses.dataclass
class InstanceSpec:
    """One instance of the batched workload."""
    n_orb: int
    n_elec: int
    n_dets: int
    slog_n: int
    seed: int


def config2_specs(n_inst: int = 64, seed: int = 2002, min_dets: int = 4096, max_dets: int = 524288):
    """Heterogeneous instances: orbitals U{16..60}, electrons U{8..24} (even,
    n_alpha = n_beta), determinant counts log-uniform in [min_dets, max_dets],
    slogdet sizes U{4..24}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(n_inst):
        n_orb = int(rng.integers(16, 61))
        n_elec = 2 * int(rng.integers(4, 13))
        n_dets = int(round(np.exp(rng.uniform(np.log(min_dets), np.log(max_dets)))))
        slog_n = int(rng.integers(4, 25))
        out.append(InstanceSpec(n_orb, n_elec, n_dets, slog_n, seed * 1000 + k))
    return out


def config4_clients(n_clients: int = 16, seed: int = 4004):
    """Serving clients: 42 orbitals, 12 alpha + 12 beta, 120 Cholesky vectors,
    one integral seed per client (a scan geometry each)."""
    return [make_integrals(42, 24, 120, seed=seed * 100 + k) for k in range(n_clients)]
