"""Shared fixtures and comparison helpers for the test suite."""

from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
H2O_STO3G = os.path.join(GOLDEN, "H2O_sto3g.FCIDUMP")
H2_631G = os.path.join(GOLDEN, "H2_631g_2.00.FCIDUMP")
E_FCI_H2O_STO3G = -75.01240139  # tests/cpp/test_hamiltonian.cpp:111


def fci_dets(n_orb: int, n_a: int, n_b: int) -> np.ndarray:
    """All determinants, sorted by (alpha, beta) (DetMap::from_list order)."""
    from itertools import combinations

    def words(k):
        return sorted(sum(1 << i for i in c) for c in combinations(range(n_orb), k))

    A = words(n_a)
    B = words(n_b)
    return np.array([(a, b) for a in A for b in B], dtype=np.uint64)


def write_tmp_fcidump(tmp_path, ints, name="synth.FCIDUMP") -> str:
    from detnqs_amd import synth
    p = os.path.join(str(tmp_path), name)
    synth.write_fcidump(p, ints)
    return p


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def assert_conn_equal(got: dict, exp: dict):
    """Bit-exact comparison of {offsets, dets, values} (order included)."""
    go = np.asarray(got["offsets"], dtype=np.int64)
    eo = np.asarray(exp["offsets"], dtype=np.int64)
    assert np.array_equal(go, eo), "offsets differ"
    assert np.array_equal(got["dets"], exp["dets"]), "connected dets differ"
    assert same_bits(np.asarray(got["values"], dtype=np.float64), np.asarray(exp["values"], dtype=np.float64)), \
        f"values differ: max |d| = {np.abs(got['values'] - exp['values']).max()}"


def assert_coo_equal(got: dict, exp: dict):
    for k in ("row", "col"):
        assert np.array_equal(np.asarray(got[k]).astype(np.int64), np.asarray(exp[k]).astype(np.int64)), f"{k} differs"
    assert tuple(got["shape"]) == tuple(exp["shape"])
    assert same_bits(np.asarray(got["val"], dtype=np.float64), np.asarray(exp["val"], dtype=np.float64)), \
        "values differ"


def explicit_p_sets(V: np.ndarray, det_P: np.ndarray, extra: np.ndarray, seed: int):
    """External sets for get_ham_block with an explicit P, built from a seeded
    shuffle: `disjoint` is half of the connected set det_P (no overlap with V);
    `overlap` mixes rows of V, part of det_P and unrelated dets `extra`, with
    first occurrences kept, so some of P lies in V and some is not connected."""
    rng = np.random.default_rng(seed)
    disjoint = det_P[rng.permutation(det_P.shape[0])[: det_P.shape[0] // 2]]
    mix = np.concatenate([V[V.shape[0] // 3: 5 * V.shape[0] // 6], disjoint[:2000], extra])
    _, first = np.unique(mix, axis=0, return_index=True)
    mix = mix[np.sort(first)]
    overlap = mix[rng.permutation(mix.shape[0])]
    return np.ascontiguousarray(disjoint), np.ascontiguousarray(overlap)
