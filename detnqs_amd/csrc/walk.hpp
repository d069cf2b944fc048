// Row walker shared by the assembly, projection and energy kernels.
//
// One wavefront walks the connections of one determinant x in the reference's
// generation order (local_conn.cpp:36-94, build_ham.cpp:256-312):
//   diagonal; singles alpha then beta (det_ops.hpp:123-159, hole-major);
//   doubles either in generation order (det_ops.hpp:161-246: aa, bb, ab) or
//   heat-bath order (ham_utils.hpp:280-313: occupied SO pairs p>q of the
//   alpha-then-beta list, each table row walked in its stored order).
// Each step hands the sink 64 consecutive candidates (one per lane):
//   sink.diag(d)                     uniform diagonal value
//   sink.single(ok, ka, kb, v)       ok = candidate exists
//   sink.dbl(ok, ka, kb, v)          ok = candidate exists (heat-bath: both
//                                    added orbitals virtual)
// v is the signed matrix element <x|H|x'> (phase included) computed with the
// reference's operands and order.  Sinks that do not need x' or the sign
// (Sink::kNeedKet == false) get them unspecified.  Count-only sinks
// (Sink::kCountOnly) receive the doubles of every pair whose kept set is the
// whole pair-mask set as one number, sink.add_doubles(count), counted with
// popcounts; only the other pairs are walked entry by entry.
#pragma once

#include "ctx.hpp"

#ifndef DQ_HB_BATCH
#define DQ_HB_BATCH 8
#endif

namespace dq {

constexpr int kWave = 64;

struct RowLists {
  int occA[64], occB[64], virA[64], virB[64];
  int occSO[128];
};

// Marks a value the whole wave holds as wave-uniform (scalar registers).
__device__ __forceinline__ u64 uniform_u64(u64 x) {
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(x));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(x >> 32));
  return (static_cast<u64>(hi) << 32) | lo;
}
__device__ __forceinline__ int64_t uniform_i64(int64_t x) {
  return static_cast<int64_t>(uniform_u64(static_cast<u64>(x)));
}

__device__ __forceinline__ u64 lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
}

// ham_eval.cpp:53-86: e1 = sum_p h_pp over the occupied SOs (alpha list then
// beta list), e2 = sum_{i>j} <pq||pq> over that list, returned as e1 + e2.
// hd(p) = h_pp, jd(p, q) = (pp|qq), kd(p, q) = (pq|qp): tables in HBM or LDS.
template <class HD, class JD, class KD>
__device__ __forceinline__ double diag_value_with(u64 xa, u64 xb, HD hd, JD jd, KD kd) {
  double e1 = 0.0;
  for (u64 m = xa; m; m &= m - 1) e1 += hd(ctz(m));
  for (u64 m = xb; m; m &= m - 1) e1 += hd(ctz(m));
  double e2 = 0.0;
  for (u64 m = xa; m; m &= m - 1) {  // p alpha: q alpha below p
    const int p = ctz(m);
    for (u64 q = xa & low_mask(p); q; q &= q - 1) {
      const int k = ctz(q);
      e2 += jd(p, k) - kd(p, k);
    }
  }
  for (u64 m = xb; m; m &= m - 1) {  // p beta: all alpha, then beta below p
    const int p = ctz(m);
    for (u64 q = xa; q; q &= q - 1) e2 += jd(p, ctz(q)) - 0.0;
    for (u64 q = xb & low_mask(p); q; q &= q - 1) {
      const int k = ctz(q);
      e2 += jd(p, k) - kd(p, k);
    }
  }
  return e1 + e2;
}

// diag_value_with with each same-spin J - K read from a table of the
// differences (each difference is one rounding either way, and x - 0.0 == x,
// so the result is bit-identical); one table read per pair instead of two.
template <class HD, class JD, class DD>
__device__ __forceinline__ double diag_value_jdd(u64 xa, u64 xb, HD hd, JD jd, DD dd) {
  double e1 = 0.0;
  for (u64 m = xa; m; m &= m - 1) e1 += hd(ctz(m));
  for (u64 m = xb; m; m &= m - 1) e1 += hd(ctz(m));
  double e2 = 0.0;
  for (u64 m = xa; m; m &= m - 1) {
    const int p = ctz(m);
    for (u64 q = xa & low_mask(p); q; q &= q - 1) e2 += dd(p, ctz(q));
  }
  for (u64 m = xb; m; m &= m - 1) {
    const int p = ctz(m);
    for (u64 q = xa; q; q &= q - 1) e2 += jd(p, ctz(q));
    for (u64 q = xb & low_mask(p); q; q &= q - 1) e2 += dd(p, ctz(q));
  }
  return e1 + e2;
}

__device__ inline double diag_value(const DevView& V, u64 xa, u64 xb) {
  const int n = V.n;
  return diag_value_with(
      xa, xb, [&](int p) { return V.h1[p * n + p]; }, [&](int p, int q) { return V.Jd[p * n + q]; },
      [&](int p, int q) { return V.Kd[p * n + q]; });
}

// ham_eval.cpp:88-115 for removing r and adding a (spin `spin`): h_ra plus, over
// the occupied orbitals of the excited det except a (i.e. those of x except r),
// alpha then beta ascending, <r k||a k> = (ra|kk) - [(rk|ka) if same spin].
__device__ inline double single_value(const DevView& V, u64 xa, u64 xb, int spin, int r, int a) {
  const int n = V.n;
  double v = V.h1[r * n + a];
  const double* J = V.Js + static_cast<size_t>(r) * n * n + a;
  const double* K = V.Ks + static_cast<size_t>(r) * n * n + a;
  const u64 ma = spin == 0 ? (xa & ~(1ULL << r)) : xa;
  const u64 mb = spin == 1 ? (xb & ~(1ULL << r)) : xb;
  for (u64 m = ma; m; m &= m - 1) {
    const int k = ctz(m);
    v += J[k * n] - (spin == 0 ? K[k * n] : 0.0);
  }
  for (u64 m = mb; m; m &= m - 1) {
    const int k = ctz(m);
    v += J[k * n] - (spin == 1 ? K[k * n] : 0.0);
  }
  return v;
}

__device__ __forceinline__ double g2(const DevView& V, int p, int q, int r, int s) {
  return V.h2[key2(p, q, r, s)];
}

__device__ __forceinline__ bool so_occupied(u64 xa, u64 xb, int so, int n_orb) {
  const int mo = so >> 1;
  if (mo < 0 || mo >= n_orb) return true;  // ham_utils.cpp:470-479
  return (((so & 1) ? xb : xa) >> mo) & 1ULL;
}

// The same test with the row's blocked masks (occupied, or at or above n_orb)
// precomputed: ba = xa | ~low_mask(n_orb), bb likewise; so < 128.
__device__ __forceinline__ bool so_blocked(u64 ba, u64 bb, int so) { return (((so & 1) ? bb : ba) >> (so >> 1)) & 1ULL; }

// Pair index t -> (u, w), u < w < v, lexicographic.
__device__ __forceinline__ void pair_from_index(int t, int v, int& u, int& w) {
  int start = 0;
  u = 0;
  while (t >= start + (v - 1 - u)) {
    start += v - 1 - u;
    ++u;
  }
  w = u + 1 + (t - start);
}

// Prefix of a heat-bath row with w >= cut (RowView::with_cutoff,
// hb_table.hpp:131-146; cut <= 0 keeps the row), capped at w > 1e-15 when
// `cap` is set (callers that then drop |element| <= 1e-15 anyway).
__device__ inline u32 hb_prefix(const double* w, u32 len, double cut, bool cap = true) {
  u32 lo = 0, hi = len;
  if (cut > 0.0) {
    while (lo < hi) {
      const u32 mid = lo + (hi - lo) / 2;
      if (w[mid] >= cut)
        lo = mid + 1;
      else
        hi = mid;
    }
    hi = lo;
    lo = 0;
  }
  if (!cap) return hi;
  u32 keep = hi;
  lo = 0;
  hi = keep;
  while (lo < hi) {
    const u32 mid = lo + (hi - lo) / 2;
    if (w[mid] > kMatThr)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Number of kept doubles of SO pair row `rid` (occupied i, j) for a det with
// virtual masks va, vb, from the pair masks.
__device__ __forceinline__ int mask_count(const DevView& V, size_t rid, int i, int j, u64 va, u64 vb) {
  const u64* m = V.pm + rid * V.n;
  const bool same = ((i ^ j) & 1) == 0;
  const u64 outer = same ? ((i & 1) ? vb : va) : va;
  const u64 inner = same ? outer : vb;
  // Eight mask words per batch, their loads issued together (the words are
  // independent; a one-at-a-time loop waits for each).
  int c = 0;
  u64 w = outer;
  while (w) {
    u64 mw[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      mw[t] = w ? m[ctz(w)] : 0ull;
      w &= w - 1;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) c += __popcll(mw[t] & inner);
  }
  return c;
}

__device__ __forceinline__ int tri_row(int t) {  // largest p with p(p-1)/2 <= t
  int p = static_cast<int>((1.0f + sqrtf(1.0f + 8.0f * static_cast<float>(t))) * 0.5f);
  while (p * (p - 1) / 2 > t) --p;
  while ((p + 1) * p / 2 <= t) ++p;
  return p;
}

struct RowEnv {
  DevView V;
  u64 xa, xb;
  int n_orb;
  int na, nb, nva, nvb;
  const RowLists* L;
};

// Fills the per-wave orbital lists (call from all lanes, then barrier).
__device__ __forceinline__ void build_lists(RowLists& L, u64 xa, u64 xb, int n_orb) {
  const int lane = __lane_id();
  const u64 full = low_mask(n_orb);
  const u64 va = ~xa & full, vb = ~xb & full;
  const int na = __popcll(xa);
  const u64 bit = 1ULL << lane;
  const u64 below = low_mask(lane);
  if (xa & bit) L.occA[__popcll(xa & below)] = lane;
  if (xb & bit) L.occB[__popcll(xb & below)] = lane;
  if (va & bit) L.virA[__popcll(va & below)] = lane;
  if (vb & bit) L.virB[__popcll(vb & below)] = lane;
  if (xa & bit) L.occSO[__popcll(xa & below)] = 2 * lane;
  if (xb & bit) L.occSO[na + __popcll(xb & below)] = 2 * lane + 1;
}

__device__ __forceinline__ RowEnv make_env(const DevView& V, u64 xa, u64 xb, int n_orb, const RowLists& L) {
  const u64 full = low_mask(n_orb);
  RowEnv R;
  R.V = V;
  R.xa = xa;
  R.xb = xb;
  R.n_orb = n_orb;
  R.na = __popcll(xa);
  R.nb = __popcll(xb);
  R.nva = __popcll(~xa & full);
  R.nvb = __popcll(~xb & full);
  R.L = &L;
  return R;
}

// dbl_mode: 0 = generation order (no table), 1 = heat-bath with per-row
// lengths hb_len[rid], 2 = heat-bath with cutoff `cut` searched per row,
// 3 = as 2 without the 1e-15 cap (every stored entry with w >= cut, as
// for_each_double_hb visits them, ham_utils.hpp:280-313), 4 = no doubles
// (the caller walks them itself).
template <class Sink>
__device__ void walk_row(const RowEnv& R, Sink& S, int dbl_mode, const u32* hb_len, double cut) {
  constexpr bool NK = Sink::kNeedKet;
  const DevView& V = R.V;
  const RowLists& L = *R.L;
  const int lane = __lane_id();
  const u64 xa = R.xa, xb = R.xb;

  S.diag(diag_value(V, xa, xb));

  for (int spin = 0; spin < 2; ++spin) {
    const int no = spin ? R.nb : R.na, nv = spin ? R.nvb : R.nva;
    const int* O = spin ? L.occB : L.occA;
    const int* Vl = spin ? L.virB : L.virA;
    const u64 xs = spin ? xb : xa;
    const int Sn = no * nv;
    for (int base = 0; base < Sn; base += kWave) {
      const int s = base + lane;
      bool ok = false;
      u64 ka = xa, kb = xb;
      double v = 0.0;
      if (s < Sn) {
        const int r = O[s / nv], a = Vl[s % nv];
        const double e = single_value(V, xa, xb, spin, r, a);
        v = (between(xs, a, r) & 1) ? -e : e;  // (+-1.0) * e
        ok = true;
        const u64 flip = (1ULL << r) | (1ULL << a);
        if (spin == 0)
          ka ^= flip;
        else
          kb ^= flip;
      }
      S.single(ok, ka, kb, v);
    }
  }

  if (dbl_mode == 4) return;

  if constexpr (Sink::kCountOnly) {
    const u64 full = low_mask(R.n_orb);
    const u64 va = ~xa & full, vb = ~xb & full;
    const int ne = R.na + R.nb, npairs = ne * (ne - 1) / 2;
    int cnt = 0;
    for (int t = lane; t < npairs; t += kWave) {
      const int p = tri_row(t), q = t - p * (p - 1) / 2;
      const int i = L.occSO[p], j = L.occSO[q];
      const int hi = i > j ? i : j, lo = i > j ? j : i;
      const size_t rid = static_cast<size_t>(hi) * (hi - 1) / 2 + lo;
      if (dbl_mode == 0 || (hb_len[rid] & kFullRow)) cnt += mask_count(V, rid, i, j, va, vb);
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, kWave);
    S.add_doubles(cnt);
    if (dbl_mode == 0) return;
  }

  if (dbl_mode != 0) {
    const int ne = R.na + R.nb;
    for (int p = 1; p < ne; ++p) {
      // wave-uniform pair: its row, bounds and spin case live in scalar registers
      const int i = __builtin_amdgcn_readfirstlane(L.occSO[p]);
      for (int q = 0; q < p; ++q) {
        const int j = __builtin_amdgcn_readfirstlane(L.occSO[q]);
        const int hi = i > j ? i : j, lo = i > j ? j : i;
        const size_t rid = static_cast<size_t>(hi) * (hi - 1) / 2 + lo;
        const u32 off = V.hb_off[rid];
        u32 len;
        if (dbl_mode == 1) {
          const u32 lf = hb_len[rid];
          if (Sink::kCountOnly && (lf & kFullRow)) continue;  // counted from the masks
          len = lf & ~kFullRow;
        } else {
          len = hb_prefix(V.hb_w + off, V.hb_off[rid + 1] - off, cut, dbl_mode == 2);
        }
        const bool same = ((i ^ j) & 1) == 0;
        const int mi = i >> 1, mj = j >> 1;
        const u64 ba = xa | ~low_mask(R.n_orb), bb = xb | ~low_mask(R.n_orb);
        constexpr int kBatch = DQ_HB_BATCH;  // chunks whose loads are issued together
        for (u32 base = 0; base < len; base += kBatch * kWave) {
          u32 abv[kBatch];
          double valv[kBatch];
#pragma unroll
          for (int c = 0; c < kBatch; ++c) {
            const u32 k = base + c * kWave + lane;
            abv[c] = k < len ? V.hb_ab[off + k] : 0u;
            valv[c] = k < len ? V.hb_v[off + k] : 0.0;
          }
#pragma unroll
          for (int c = 0; c < kBatch; ++c) {
            if (base + c * kWave >= len) break;
            const u32 k = base + c * kWave + lane;
            bool ok = false;
            u64 ka = xa, kb = xb;
            double v = 0.0;
            if (k < len) {
              const u32 ab = abv[c];
              const int a = static_cast<int>(ab & 0xFFFFu), b = static_cast<int>(ab >> 16);
              const double val = valv[c];
              ok = !(static_cast<int>(so_blocked(ba, bb, a)) | static_cast<int>(so_blocked(ba, bb, b)));
              v = val;
              if (NK && ok) {
                int perm;
                const int ma = a >> 1, mb = b >> 1;
                if (same) {
                  const u64 xs = (i & 1) ? xb : xa;
                  perm = parity2(xs, mi < mj ? mi : mj, mi < mj ? mj : mi, ma < mb ? ma : mb, ma < mb ? mb : ma);
                  const u64 flip = (1ULL << mi) | (1ULL << mj) | (1ULL << ma) | (1ULL << mb);
                  if (i & 1)
                    kb ^= flip;
                  else
                    ka ^= flip;
                } else {
                  const int ra = (i & 1) ? mj : mi, rb = (i & 1) ? mi : mj;
                  const int aa = (a & 1) ? mb : ma, ab_ = (a & 1) ? ma : mb;
                  perm = between(xa, aa, ra) + between(xb, ab_, rb);
                  ka ^= (1ULL << ra) | (1ULL << aa);
                  kb ^= (1ULL << rb) | (1ULL << ab_);
                }
                v = (perm & 1) ? -val : val;
              }
            }
            S.dbl(ok, ka, kb, v);
          }
        }
      }
    }
    return;
  }

  // generation order: same-spin doubles alpha then beta (det_ops.hpp:161-215)
  for (int spin = 0; spin < 2; ++spin) {
    const int no = spin ? R.nb : R.na, nv = spin ? R.nvb : R.nva;
    if (no < 2 || nv < 2) continue;
    const int* O = spin ? L.occB : L.occA;
    const int* Vl = spin ? L.virB : L.virA;
    const u64 xs = spin ? xb : xa;
    const int NP = nv * (nv - 1) / 2;
    for (int h1i = 0; h1i < no; ++h1i)
      for (int h2i = h1i + 1; h2i < no; ++h2i) {
        const int h0 = __builtin_amdgcn_readfirstlane(O[h1i]), h1 = __builtin_amdgcn_readfirstlane(O[h2i]);
        for (int base = 0; base < NP; base += kWave) {
          const int t = base + lane;
          bool ok = false;
          u64 ka = xa, kb = xb;
          double v = 0.0;
          if (t < NP) {
            int u, w;
            pair_from_index(t, nv, u, w);
            const int p0 = Vl[u], p1 = Vl[w];
            const double val = g2(V, h0, p0, h1, p1) - g2(V, h0, p1, h1, p0);
            ok = true;
            v = val;
            if (NK) {
              v = (parity2(xs, h0, h1, p0, p1) & 1) ? -val : val;
              const u64 flip = (1ULL << h0) | (1ULL << h1) | (1ULL << p0) | (1ULL << p1);
              if (spin)
                kb ^= flip;
              else
                ka ^= flip;
            }
          }
          S.dbl(ok, ka, kb, v);
        }
      }
  }
  // opposite-spin doubles, nested (ha, pa, hb, pb) (det_ops.hpp:217-246)
  if (R.na && R.nva && R.nb && R.nvb) {
    const int NB = R.nb * R.nvb;
    for (int hai = 0; hai < R.na; ++hai)
      for (int pai = 0; pai < R.nva; ++pai) {
        const int ha = __builtin_amdgcn_readfirstlane(L.occA[hai]);
        const int pa = __builtin_amdgcn_readfirstlane(L.virA[pai]);
        const int pa_perm = between(xa, pa, ha);
        for (int base = 0; base < NB; base += kWave) {
          const int t = base + lane;
          bool ok = false;
          u64 ka = xa, kb = xb;
          double v = 0.0;
          if (t < NB) {
            const int hb = L.occB[t / R.nvb], pb = L.virB[t % R.nvb];
            const double val = g2(V, ha, pa, hb, pb);
            ok = true;
            v = val;
            if (NK) {
              v = ((pa_perm + between(xb, pb, hb)) & 1) ? -val : val;
              ka ^= (1ULL << ha) | (1ULL << pa);
              kb ^= (1ULL << hb) | (1ULL << pb);
            }
          }
          S.dbl(ok, ka, kb, v);
        }
      }
  }
}

}  // namespace dq
