"""The order of a row's connections without sorting (detnqs_amd/csrc/rowrank.hpp),
restated in Python and checked against brute force.

For a row x = (xa, xb) the kets y that move at most two electrons, listed in
(alpha, beta) integer order, must be numbered 0 .. M-1 by the transition
tables (one lookup per changed orbital, visited from the highest down), and the
same walk must give each string's colex rank.  Brute force: enumerate every
string of the sector, keep those within two moves, sort the pairs.  A table
entry off by one state (a hole taken for a particle, a wrong successor, a
dropped particle term) breaks the numbering for some ket, so every ket of
every case is compared.  The device code itself is checked against the oracle
by the GPU projection tests (get_ham_conn in row order)."""

import itertools
import math
import random

import pytest


def small(n, r):
    return 0 if r < 0 else (1 if r == 0 else (n if r == 1 else n * (n - 1) // 2))


def binom(n, r):
    return math.comb(n, r) if 0 <= r <= n else 0


def popc(v):
    return bin(v).count("1")


def moves(n, k, d):
    return [1, k * (n - k), math.comb(k, 2) * math.comb(n - k, 2)][d]


def tables(x, n, w):
    """Transition tables G[s][q] (s = 3u + h) and Gc[d][q] for one string."""
    # region prefix sums A_s(q) over occupied p < q, particle terms P_s(q)
    A = [[0] * (n + 1) for _ in range(9)]
    for s in range(9):
        u, h = divmod(s, 3)
        acc = 0
        for p in range(n):
            if (x >> p) & 1:
                o = popc(x & ((1 << p) - 1))
                acc += sum(w[d] * small(p - o, d - u) * small(o, d - h - 1) for d in range(3))
            A[s][p + 1] = acc
    R = [[0] * (n + 1) for _ in range(5)]
    for d in range(5):
        acc = 0
        for p in range(n):
            if (x >> p) & 1:
                acc += binom(p, popc(x & ((2 << p) - 1)) + d - 2)
            R[d][p + 1] = acc
    G = [[0] * n for _ in range(9)]
    for s in range(9):
        u, h = divmod(s, 3)
        for q in range(n):
            o = popc(x & ((1 << q) - 1))
            if (x >> q) & 1:
                if h < 2:
                    G[s][q] = A[s + 1][q] - A[s][q + 1]
            elif u < 2:
                part = sum(w[d] * small(q - o, d - u) * small(o, d - h) for d in range(3))
                G[s][q] = A[s + 3][q] - A[s][q + 1] + part
    Gc = [[0] * n for _ in range(5)]
    for d in range(5):
        for q in range(n):
            o = popc(x & ((1 << q) - 1))
            if (x >> q) & 1:
                if d < 4:
                    Gc[d][q] = R[d + 1][q] - R[d][q + 1]
            elif d > 0:
                Gc[d][q] = R[d - 1][q] - R[d][q + 1] + binom(q, o + d - 2)
    return G, Gc, A[0][n], R[2][n]


def spin_rank(T, x, y):
    G, Gc, base, cbase = T
    D, s, d, f, c = x ^ y, 0, 2, base, cbase
    while D:
        q = D.bit_length() - 1
        D ^= 1 << q
        hole = (x >> q) & 1
        f += G[s][q]
        c += Gc[d][q]
        s += 1 if hole else 3
        d += 1 if hole else -1
    return f, c


def colex(y):
    bits = [p for p in range(y.bit_length()) if (y >> p) & 1]
    return sum(binom(p, i + 1) for i, p in enumerate(bits))


def check_row(n, xa, xb):
    na, nb = popc(xa), popc(xb)
    size = [sum(moves(n, nb, d) for d in range(3 - da)) for da in range(3)]
    ta = tables(xa, n, size)
    tb = [tables(xb, n, [1, 1, 1]), tables(xb, n, [1, 1, 0])]

    def near(x, k):
        out = []
        for c in itertools.combinations(range(n), k):
            y = sum(1 << p for p in c)
            if popc(y & ~x) <= 2:
                out.append(y)
        return out

    U = sorted((ya, yb) for ya in near(xa, na) for yb in near(xb, nb)
               if popc(ya & ~xa) + popc(yb & ~xb) <= 2)
    assert len(U) == sum(moves(n, na, da) * size[da] for da in range(3))
    for L, (ya, yb) in enumerate(U):
        da = popc(ya & ~xa)
        fa, ca = spin_rank(ta, xa, ya)
        fb, cb = spin_rank(tb[min(da, 1)], xb, yb)
        if da == 2:
            fb = 0
        assert (fa + fb, ca, cb) == (L, colex(ya), colex(yb)), (hex(xa), hex(xb), hex(ya), hex(yb))


@pytest.mark.parametrize("n,na,nb", [(8, 3, 4), (10, 5, 2), (10, 0, 3), (7, 7, 3), (12, 4, 4),
                                     (6, 1, 1), (9, 8, 1), (11, 2, 9), (13, 5, 5)])
def test_order_and_colex_against_brute_force(n, na, nb):
    rng = random.Random(1000 * n + 10 * na + nb)
    for _ in range(2):
        xa = sum(1 << p for p in rng.sample(range(n), na))
        xb = sum(1 << p for p in rng.sample(range(n), nb))
        check_row(n, xa, xb)


def test_colex_orders_like_integers():
    """Colex ranks of one popcount are a bijection onto 0 .. C(n, k)-1 that
    keeps integer order (the ket keys sort like the packed words)."""
    for n, k in [(10, 4), (12, 6), (9, 1), (8, 8)]:
        ys = sorted(sum(1 << p for p in c) for c in itertools.combinations(range(n), k))
        assert [colex(y) for y in ys] == list(range(math.comb(n, k)))
