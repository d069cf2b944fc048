// Order of one row's connections without sorting them (n_orb <= 32).
//
// The connections of a row x = (xa, xb) are the kets y = (ya, yb) that move
// at most two electrons in all.  Listed in (alpha, beta) integer order -- the
// order of P and of every row of H_VP (det.hpp:30-36, ham_utils.cpp:41-83) --
// they are numbered 0 .. M-1 by
//
//   L(y) = Phi_A(ya) + Phi_B[da](yb),   Phi_w(y) = sum_d w_d #{z in S_d(x) : z < y}
//
// where S_d(x) is the set of strings that move d electrons of x (d = 0, 1, 2),
// da the number of alpha moves of y, the alpha weights w_d the number of beta
// strings allowed after d alpha moves, and the beta weights 1 for the d that
// da leaves room for.  #{z in S_d(x) : z < y} is a sum over the set bits p of
// y of the strings that agree with y above p and clear p: the particles still
// to place go to the empty positions of x below p, the holes to its occupied
// positions below p, so each term is C(empty below p, d - u) C(occupied below
// p, d - h - x_p) with u particles and h holes of y above p.  Between two
// changed positions of y these terms depend on p through x alone: per-row
// prefix tables over p reduce the sum to two lookups per changed position.
//
// The same walk gives the colex rank of each string (the combinatorial number
// system, sum over the set bits of C(p, index of p)), which orders strings of
// one popcount like their integer values: C(n, na) C(n, nb) ket keys of
// ceil(log2) bits instead of 2 n, for the radix sort that ranks P.
#pragma once
#include "dq_common.hpp"

namespace dq {

// C(n, r) for 0 <= n, r <= 32 (C(32, 16) < 2^32).
struct BinomTable {
  u32 c[33][33];
  constexpr BinomTable() : c() {
    for (int n = 0; n <= 32; ++n) {
      c[n][0] = 1;
      for (int r = 1; r <= n; ++r) c[n][r] = c[n - 1][r - 1] + (r <= n - 1 ? c[n - 1][r] : 0u);
    }
  }
};
inline constexpr BinomTable kBinomHost{};
// Device copy; kernels stage it in LDS.
static __device__ const BinomTable kBinomDev{};

inline u64 binom_host(int n, int r) { return (r < 0 || r > n || n > 32) ? 0ull : kBinomHost.c[n][r]; }

// C(n, r) for r <= 2.
__device__ __forceinline__ u32 small_binom(int n, int r) {
  return r < 0 ? 0u : (r == 0 ? 1u : (r == 1 ? static_cast<u32>(n) : static_cast<u32>(n * (n - 1) / 2)));
}

// Number of strings of n orbitals and k electrons that move d of them.
__host__ __device__ __forceinline__ u64 moves(int n, int k, int d) {
  const u64 o = static_cast<u64>(k), v = static_cast<u64>(n - k);
  return d == 0 ? 1ull : (d == 1 ? o * v : (o * (o - (o > 0)) / 2) * (v * (v - (v > 0)) / 2));
}

// Connections of a row with na alpha and nb beta electrons (the row included).
__host__ __device__ inline u64 row_universe(int n, int na, int nb) {
  u64 m = 0;
  for (int da = 0; da <= 2; ++da) {
    u64 size = 0;
    for (int d = 0; d <= 2 - da; ++d) size += moves(n, nb, d);
    m += moves(n, na, da) * size;
  }
  return m;
}

// Per-wave tables of one row (LDS).  With the changed positions of y visited
// from the highest down, state s = 3 u + h (particles and holes seen so far)
// and colex index shift d = 2 + h - u, each change at q adds one entry:
//   Phi   += G[v][s][q] = A_{s'}(q) - A_s(q + 1) + (particle ? P_s(q) : 0)
//   colex += Gc[sp][d][q] = R_{d'}(q) - R_d(q + 1) + (particle ? C(q, o(q) + d - 2) : 0)
// where A_s(q), R_d(q) are the region prefix sums over the occupied p < q,
// P_s(q) the term of a particle at q and s', d' the states after the change
// (hole: s + 1, d + 1; particle: s + 3, d - 1); the region below the lowest
// change is carried by the next entry's A_{s'}(q) term, the region above the
// highest by the base A_0(n).
struct RowRank {
  u32 G[3][9][32];  // [0] alpha weights, [1] beta after no alpha move, [2] after one
  u32 Gc[2][5][32];
  u32 base[3];      // Phi of x itself
  u32 cbase[2];     // colex rank of xa, xb
};

// The colex part alone (for callers that need only the ket ranks).
struct ColexRank {
  u32 Gc[2][5][32];
  u32 cbase[2];
};

// Fills T for row (xa, xb) from all 64 lanes of a wave (read after a barrier):
// with PHI the Phi and colex tables of a RowRank, without only the colex
// tables (a RowRank or a ColexRank).  B: binomial table.
template <bool PHI, class TT>
__device__ inline void build_rank_tables(TT& T, const BinomTable& B, u32 xa, u32 xb, int n) {
  const int lane = __lane_id(), half = lane >> 5, p = lane & 31;
  const int nb = __popc(xb);
  const u32 s1 = 1u + static_cast<u32>(moves(n, nb, 1));
  const u32 s0 = s1 + static_cast<u32>(moves(n, nb, 2));
  // five tasks, two per round (one per half-wave): Phi v = 0, 1, 2, colex sp = 0, 1
  for (int round = 0; round < (PHI ? 3 : 1); ++round) {
    const int task = PHI ? 2 * round + half : 3 + half;
    if (task >= 5) break;  // half-wave uniform
    const bool colex = task >= 3;
    const u32 x = (task == 0 || task == 3) ? xa : xb;
    const bool occ = p < n && ((x >> p) & 1u);
    const bool vir = p < n && !occ;
    const int o = __popc(x & ((1u << p) - 1u)), vl = p - o;
    const u32 w0 = task == 0 ? s0 : 1u, w1 = task == 0 ? s1 : 1u, w2 = task == 2 ? 0u : 1u;
    u32 incl[9], part[9];
#pragma unroll
    for (int st = 0; st < 9; ++st) {
      u32 t = 0, pt = 0;
      if (!colex) {
        const int u = st / 3, h = st % 3;
        if (occ)
          t = w0 * small_binom(vl, -u) * small_binom(o, -h - 1) + w1 * small_binom(vl, 1 - u) * small_binom(o, -h) +
              w2 * small_binom(vl, 2 - u) * small_binom(o, 1 - h);
        if (vir)
          pt = w0 * small_binom(vl, -u) * small_binom(o, -h) + w1 * small_binom(vl, 1 - u) * small_binom(o, 1 - h) +
               w2 * small_binom(vl, 2 - u) * small_binom(o, 2 - h);
      } else if (st < 5) {
        const int dl = st - 2;
        if (occ) {
          const int idx = o + 1 + dl;  // 1-based index of p among the set bits, shifted
          t = (idx >= 0 && idx <= p) ? B.c[p][idx] : 0u;
        }
        if (vir) {
          const int idx = o + dl;
          pt = (idx >= 0 && idx <= p) ? B.c[p][idx] : 0u;
        }
      }
      u32 a = t;
#pragma unroll
      for (int off = 1; off < 32; off <<= 1) {
        const u32 y = __shfl_up(a, off, 32);
        if (p >= off) a += y;
      }
      incl[st] = a;
      part[st] = pt;
    }
    // exclusive prefix (occupied p' < p) of each state
    u32 excl[9];
#pragma unroll
    for (int st = 0; st < 9; ++st) {
      const u32 prev = __shfl_up(incl[st], 1, 32);
      excl[st] = p > 0 ? prev : 0u;
    }
    if (!colex) {
      if constexpr (PHI) {
        u32* G = &T.G[task][0][0];
#pragma unroll
        for (int st = 0; st < 9; ++st) {
          const int u = st / 3, h = st % 3;
          u32 g = 0;
          if (occ && h < 2) g = excl[st + 1] - incl[st];
          if (vir && u < 2) g = excl[st + 3] - incl[st] + part[st];
          G[st * 32 + p] = g;
        }
        const u32 tot = __shfl(incl[0], 31, 32);
        if (p == 0) T.base[task] = tot;
      }
    } else {
      u32* G = &T.Gc[task - 3][0][0];
#pragma unroll
      for (int st = 0; st < 5; ++st) {
        u32 g = 0;
        if (occ && st < 4) g = excl[st + 1] - incl[st];
        if (vir && st > 0) g = excl[st - 1] - incl[st] + part[st];
        G[st * 32 + p] = g;
      }
      const u32 tot = __shfl(incl[2], 31, 32);
      if (p == 0) T.cbase[task - 3] = tot;
    }
  }
}

__device__ inline void build_row_rank(RowRank& T, const BinomTable& B, u32 xa, u32 xb, int n) {
  build_rank_tables<true>(T, B, xa, xb, n);
}

// Phi_v(y) (and, with CR, the colex rank of y) for one spin string y near x.
template <bool CR>
__device__ __forceinline__ void spin_rank(const RowRank& T, int v, int sp, u32 x, u32 y, u32& phi, u32& crank) {
  u32 D = x ^ y;
  int s = 0, d = 2;
  u32 f = T.base[v], c = T.cbase[sp];
  const u32* G = &T.G[v][0][0];
  const u32* Gc = &T.Gc[sp][0][0];
  while (D) {
    const int q = 31 - __clz(D);
    D ^= 1u << q;
    const bool hole = (x >> q) & 1u;
    f += G[s * 32 + q];
    if (CR) c += Gc[d * 32 + q];
    s += hole ? 1 : 3;
    d += hole ? 1 : -1;
  }
  phi = f;
  crank = c;
}

// Colex rank of spin string y (sp 0 alpha, 1 beta) of the same popcount near
// the row's string x (at most two holes and two particles), from T's Gc table.
template <class TT>
__device__ __forceinline__ u32 crank_near(const TT& T, int sp, u32 x, u32 y) {
  u32 D = x ^ y;
  int d = 2;
  u32 c = T.cbase[sp];
  const u32* Gc = &T.Gc[sp][0][0];
  while (D) {
    const int q = 31 - __clz(D);
    D ^= 1u << q;
    const bool hole = (x >> q) & 1u;
    c += Gc[d * 32 + q];
    d += hole ? 1 : -1;
  }
  return c;
}

// Position of ket (ya, yb) among the row's connections; with CR also the
// colex ranks of ya and yb.
template <bool CR>
__device__ __forceinline__ u32 ket_order(const RowRank& T, u32 xa, u32 xb, u32 ya, u32 yb, u32& ra, u32& rb) {
  const int da = __popc(ya & ~xa);
  u32 fa, fb = 0;
  spin_rank<CR>(T, 0, 0, xa, ya, fa, ra);
  if (da < 2) {
    spin_rank<CR>(T, 1 + da, 1, xb, yb, fb, rb);
  } else {
    rb = T.cbase[1];  // yb = xb
  }
  return fa + fb;
}

// Colex rank r of k electrons -> string (largest C(c, i) <= r, i = k .. 1).
__device__ __forceinline__ u32 colex_unrank(const BinomTable& B, u32 r, int k, int n) {
  u32 s = 0;
  int c = n - 1;
  for (int i = k; i >= 1; --i) {
    while (B.c[c][i] > r) --c;
    s |= 1u << c;
    r -= B.c[c][i];
    --c;
  }
  return s;
}

// Colex rank of string s: sum over its set bits p_1 < p_2 < ... of C(p_i, i).
__device__ __forceinline__ u32 colex_rank(const BinomTable& B, u32 s) {
  u32 r = 0;
  int i = 1;
  for (u32 m = s; m; m &= m - 1, ++i) r += B.c[__builtin_ctz(m)][i];
  return r;
}

namespace {

// Popcount range of the rows: mm = {min na, max na, min nb, max nb}.
__global__ void pop_range(const u64* __restrict__ dets, int64_t n, int* __restrict__ mm) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int lo_a = 64, hi_a = -1, lo_b = 64, hi_b = -1;
  if (i < n) {
    lo_a = hi_a = __popcll(dets[2 * i]);
    lo_b = hi_b = __popcll(dets[2 * i + 1]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo_a = min(lo_a, __shfl_xor(lo_a, off));
    hi_a = max(hi_a, __shfl_xor(hi_a, off));
    lo_b = min(lo_b, __shfl_xor(lo_b, off));
    hi_b = max(hi_b, __shfl_xor(hi_b, off));
  }
  if (__lane_id() == 0 && hi_a >= 0) {
    atomicMin(mm, lo_a);
    atomicMax(mm + 1, hi_a);
    atomicMin(mm + 2, lo_b);
    atomicMax(mm + 3, hi_b);
  }
}

// colex rank -> string for every string of k electrons (decode table).
__global__ void colex_strings(u32* __restrict__ out, u32 count, int k, int n) {
  const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < count) out[r] = colex_unrank(kBinomDev, r, k, n);
}

}  // namespace

}  // namespace dq
