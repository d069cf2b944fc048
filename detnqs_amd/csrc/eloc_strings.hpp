// V-driven local energy of one row (one wavefront): sigma_x = sum over the
// determinants y of V connected to x of H_xy c_y, found through the string
// index instead of generating every connected determinant.
//
// Matrix elements, phases and screening are those of the candidate walk
// (walk.hpp, after ham_eval.cpp:33-147 and local_conn.cpp:148-229):
//   singles   v = +-single_value(r -> a), kept if |v| > sthr;
//   doubles   same spin: (h0 p0|h1 p1) - (h0 p1|h1 p0), h0<h1 removed,
//             p0<p1 added, phase parity2; opposite spin: (r a|s b), phase
//             between(alpha) + between(beta); kept if |v| > 1e-15 and, with a
//             heat-bath table, |v| >= dmin (the table keeps an excitation iff its
//             weight |<ij||ab>| = |v| is >= the build threshold and >= eps1);
//   diagonal  kept if |d| > 1e-15, weighted by the row's own coefficient;
//   y counts only as the first occurrence of its determinant (DetMap); the
//   groups hold first occurrences only, and coefficients are read in member
//   order (StrCoef), so a scan streams words and coefficients contiguously.
#pragma once
#include "hash.hpp"
#include "strings.hpp"
#include "walk.hpp"

namespace dq {

// Wave-scope barrier for LDS written and read by lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ u64 word_hash(u64 w) { return mix64(w ^ 0x2545F4914F6CDD1DULL); }

// Copy of a struct at a wave-uniform address into scalar registers.  A plain
// load through the pointer stays in vector registers when the compiler cannot
// prove that the kernel's stores leave the struct unchanged.
template <class T>
__device__ __forceinline__ T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "struct size must be a multiple of 4 bytes");
  constexpr int kWords = sizeof(T) / 4;
  const unsigned* src = reinterpret_cast<const unsigned*>(p);
  unsigned w[kWords];
#pragma unroll
  for (int i = 0; i < kWords; ++i) w[i] = __builtin_amdgcn_readfirstlane(src[i]);
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// Group of a word, or -1.
__device__ __forceinline__ long str_find(const u64* __restrict__ hkey, const u32* __restrict__ hval, u64 mask,
                                         u64 w) {
  if (w == ~0ULL) return -1;  // reserved word (empty slots, later duplicates)
  u64 s = word_hash(w) & mask;
  while (true) {
    const u64 k = hkey[s];
    if (k == w) return static_cast<long>(hval[s]);
    if (k == ~0ULL) return -1;
    s = (s + 1) & mask;
  }
}

struct StrWave {
  RowLists L;
  u32 hit_beg[kWave];  // first member of the hit's alpha group
  u32 hit_pre[kWave];  // members of the earlier hits of the round
  u32 hit_ra[kWave];   // r | a << 8 of the hit's alpha single excitation
};

// Members y of V connected to x, visited one wavefront per row.  The visitor
// gets every kept (y, H_xy) pair in a fixed per-lane order:
//   sink.diag(d)            lane 0 only, first, if |d| > 1e-15;
//   sink.member(v, side, k) member k of an alpha group (side 1: rows X.a_row,
//                           amplitudes C.ca) or of a beta group (side 0:
//                           X.b_row, C.cb), v = H_xy.
// The energy sink adds v * c_y into the lane's partial sum, so a lane's sum
// is fixed by that order; the layout sinks (count, emit) record the same
// order, which lets a product over the recorded layout reproduce the sum bit
// for bit.

// Members of one group that differ from x in the other spin word only (spin 0:
// x's beta group, alpha words differ; spin 1: x's alpha group, beta words
// differ).  The member's side is the group's: side = spin.
//
// Items of consecutive scans are dealt to the lanes as one round-robin
// sequence: `cur` (the same in every lane) is the lane that takes the scan's
// first item, and moves past the scan's last one.
__device__ __forceinline__ u32 first_item(u32 cur) { return (static_cast<u32>(__lane_id()) - cur) & (kWave - 1); }

template <class Sink>
__device__ __forceinline__ void str_same_group(const DevView& V, u64 xa, u64 xb, int spin,
                                               const u64* __restrict__ word, u32 beg, u32 end, double sthr,
                                               double dmin, u32& cur, Sink& sink) {
  const u64 xs = spin ? xb : xa;
  const int side = spin;
  const u32 m = end - beg;
  for (u32 i = first_item(cur); i < m; i += kWave) {
    const u32 k = beg + i;
    const u64 ys = word[k];
    const u64 rm = xs & ~ys, ad = ys & ~xs;
    const int nr = popc(rm);
    if (nr == 0 || nr > 2 || nr != popc(ad)) continue;
    if (nr == 1) {
      const int r = ctz(rm), a = ctz(ad);
      const double e = single_value(V, xa, xb, spin, r, a);
      const double v = (between(xs, a, r) & 1) ? -e : e;
      if (fabs(v) > sthr) sink.member(v, side, k);
    } else {
      const int h0 = ctz(rm), h1 = ctz(rm & (rm - 1));
      const int p0 = ctz(ad), p1 = ctz(ad & (ad - 1));
      const double val = g2(V, h0, p0, h1, p1) - g2(V, h0, p1, h1, p0);
      if (fabs(val) > kMatThr && fabs(val) >= dmin) {
        const double v = (parity2(xs, h0, h1, p0, p1) & 1) ? -val : val;
        sink.member(v, side, k);
      }
    }
  }
  cur = (cur + m) & (kWave - 1);
}

// All lanes call it; W is the wave's LDS (W.L holds x's orbital lists).
template <class Sink>
__device__ inline void str_visit_row(const DevView& V, const StrIndex& X, u64 xa, u64 xb, int n_orb, double sthr,
                                     double dmin, StrWave& W, Sink& sink) {
  const int lane = __lane_id();
  if (lane == 0) {
    const double d = diag_value(V, xa, xb);
    if (fabs(d) > kMatThr) sink.diag(d);
  }
  u32 cur = 1;  // the diagonal was lane 0's
  // beta group of x: alpha singles and alpha-alpha doubles
  const long gb = str_find(X.b_hkey, X.b_hval, X.hmask, xb);
  if (gb >= 0) str_same_group(V, xa, xb, 0, X.b_word, X.b_off[gb], X.b_off[gb + 1], sthr, dmin, cur, sink);
  // alpha group of x: beta singles and beta-beta doubles
  const long ga = str_find(X.a_hkey, X.a_hval, X.hmask, xa);
  if (ga >= 0) str_same_group(V, xa, xb, 1, X.a_word, X.a_off[ga], X.a_off[ga + 1], sthr, dmin, cur, sink);
  // alpha-beta doubles: alpha groups of the alpha single excitations of x.
  // A round probes up to 64 excitations; the members of all its hit groups
  // are numbered consecutively (hit order, then member order) and dealt to the
  // lanes round-robin, so the lanes share the scan whatever the group sizes.
  const u64 full = low_mask(n_orb);
  const int na = popc(xa), nva = popc(~xa & full);
  const int ns = na * nva;
  for (int base = 0; base < ns; base += kWave) {
    const int t = base + lane;
    long g = -1;
    int r = 0, a = 0;
    if (t < ns) {
      r = W.L.occA[t / nva];
      a = W.L.virA[t % nva];
      g = str_find(X.a_hkey, X.a_hval, X.hmask, xa ^ (1ULL << r) ^ (1ULL << a));
    }
    const u64 hits = __ballot(g >= 0);
    u32 beg = 0, sz = 0;
    if (g >= 0) {
      beg = X.a_off[g];
      sz = X.a_off[g + 1] - beg;
    }
    u32 inc = sz;  // inclusive prefix over lanes (hit lanes are in hit order)
    for (int d = 1; d < kWave; d <<= 1) {
      const u32 y = static_cast<u32>(__shfl_up(static_cast<int>(inc), d, kWave));
      if (lane >= d) inc += y;
    }
    const u32 m = static_cast<u32>(__shfl(static_cast<int>(inc), kWave - 1, kWave));
    if (g >= 0) {
      const int pos = __popcll(hits & lanemask_lt());
      W.hit_beg[pos] = beg;
      W.hit_pre[pos] = inc - sz;
      W.hit_ra[pos] = static_cast<u32>(r) | (static_cast<u32>(a) << 8);
    }
    wave_sync();
    const int nh = __popcll(hits);
    for (u32 f = first_item(cur); f < m; f += kWave) {
      int lo = 0, hi = nh;  // last hit whose members start at or before f
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (W.hit_pre[mid] <= f)
          lo = mid;
        else
          hi = mid;
      }
      const u32 k = W.hit_beg[lo] + (f - W.hit_pre[lo]);
      const u32 ra = W.hit_ra[lo];
      const int rr = static_cast<int>(ra & 0xFFu), aa = static_cast<int>(ra >> 8);
      const u64 yb = X.a_word[k];
      const u64 rm = xb & ~yb, ad = yb & ~xb;
      if (popc(rm) != 1 || popc(ad) != 1) continue;
      const int s = ctz(rm), b = ctz(ad);
      const double val = g2(V, rr, aa, s, b);
      if (fabs(val) > kMatThr && fabs(val) >= dmin) {
        const double v = ((between(xa, aa, rr) + between(xb, b, s)) & 1) ? -val : val;
        sink.member(v, 1, k);
      }
    }
    cur = (cur + m) & (kWave - 1);
    wave_sync();  // the hit list is rewritten by the next round
  }
}

// Lane partial sum of sigma_x (the caller reduces over the wave).
struct StrSumSink {
  const double* ca;
  const double* cb;
  double ci;
  double acc;
  __device__ __forceinline__ void diag(double d) { acc += d * ci; }
  __device__ __forceinline__ void member(double v, int side, u32 k) { acc += v * (side ? ca[k] : cb[k]); }
};

__device__ inline double str_energy_row(const DevView& V, const StrIndex& X, const StrCoef& C, u64 xa, u64 xb,
                                        int n_orb, double ci, double sthr, double dmin, StrWave& W) {
  StrSumSink S{C.ca, C.cb, ci, 0.0};
  str_visit_row(V, X, xa, xb, n_orb, sthr, dmin, W, S);
  return S.acc;
}

// ------------------------------------------------------------------ lane-ordered H_VV
// H_VV of V in the order the energy visits it: row x owns rounds
// [off[x], off[x+1]); slot (round t, lane l) is col[64 t + l], val[64 t + l],
// and lane l's j-th kept pair sits in round off[x] + j.  Unused slots hold
// kVvHole.  Columns are rows of V (the first occurrence of y; the diagonal
// names row x itself).  sigma_x = sum over lanes (butterfly, as the energy)
// of the lane's pairs in round order, each v * c[col], which is the energy
// kernel's sum bit for bit.
constexpr u32 kVvHole = 0xFFFFFFFFu;

struct VvCountSink {
  u32 n = 0;
  __device__ __forceinline__ void diag(double) { ++n; }
  __device__ __forceinline__ void member(double, int, u32) { ++n; }
};

struct VvEmitSink {
  const u32* a_row;
  const u32* b_row;
  u32* col;     // this lane's first slot; its j-th pair goes to col[64 j]
  double* val;
  u32 row;
  u32 n = 0;
  __device__ __forceinline__ void diag(double d) {
    col[0] = row;
    val[0] = d;
    n = 1;
  }
  __device__ __forceinline__ void member(double v, int side, u32 k) {
    col[static_cast<size_t>(n) * kWave] = side ? a_row[k] : b_row[k];
    val[static_cast<size_t>(n) * kWave] = v;
    ++n;
  }
};

// Lane partial sum of sigma_x over the lane-ordered layout.  Every lane runs
// all of the row's rounds (r0, r1 are the same in every lane), so the loop
// has no lane-dependent exit; a lane's holes all follow its pairs.  Four
// rounds' slots and amplitudes are loaded before any of them is added, and the
// adds keep round order.
__device__ __forceinline__ double vv_row_sum(const u32* __restrict__ col, const double* __restrict__ val,
                                             int64_t r0, int64_t r1, const double* __restrict__ c) {
  constexpr int kU = 4;
  double acc = 0.0;
  const int lane = __lane_id();
  int64_t t = r0;
  for (; t + kU <= r1; t += kU) {
    u32 j[kU];
    double v[kU], cj[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const size_t slot = static_cast<size_t>(t + u) * kWave + lane;
      j[u] = col[slot];
      v[u] = val[slot];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) cj[u] = j[u] != kVvHole ? c[j[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (j[u] != kVvHole) acc += v[u] * cj[u];
  }
  for (; t < r1; ++t) {
    const size_t slot = static_cast<size_t>(t) * kWave + lane;
    const u32 j = col[slot];
    if (j != kVvHole) acc += val[slot] * c[j];
  }
  return acc;
}

}  // namespace dq
