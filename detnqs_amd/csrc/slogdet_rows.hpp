// Row-distributed batched slogdet for 8 < n <= 64 (gfx950, wave64, FP64).
//
// A group of G lanes owns one matrix; lane q of the group holds rows
// q, q + G, q + 2G, ... (R = ceil(NC / G) rows of NC register columns), so a
// wave factors 64 / G matrices at once.  Fewer lanes per matrix means more
// matrices share each wave-wide instruction (pivot search rounds, the pivot
// row broadcast, the reciprocal) while the row updates, which are the only
// work that grows with R, stay a small part of the instruction stream.
//
// The factorisation is LAPACK getf2 (Golub & Van Loan Alg. 3.4.1), exactly the
// operation sequence of the reference semantics (csrc/ffi/logdet_cuda.cu:
// 187-291): at step k the pivot is the first row, in the current row order, of
// maximal |a_rk|; multipliers are a_rk * (1 / pivot) when |pivot| >= DBL_MIN
// and a_rk / pivot otherwise; the update is a_rj - l * u_kj with no FMA
// contraction.  Rows are never moved: each keeps its current position `pos`
// (a row swap exchanges two positions) and the step at which it was chosen.
// The pivot search is a local pass over the lane's rows followed by DPP rounds
// across the group (quad_perm xor 1 and xor 2, half-row and row mirrors; lane
// shuffles beyond 16 lanes); the pivot row is broadcast with ds_bpermute.
// log|u_kk| is taken once, by the lane that holds step k, after the loop, and
// the logs are added in step order (the same sum as adding them step by step).
//
// Backward (cot * A^{-T}, logdet_cuda.cu:293-386): L\U in position order and
// the permutation are staged in LDS; lane q solves the R contiguous columns
// q R .. q R + R - 1 of A^{-T} (U^T y = e_c, then L^T w = y, getrs 'T' order)
// and writes them through the permutation, so for each row of the result the
// group stores one contiguous run.
#pragma once

#include <cfloat>

#include "slogdet.hpp"

namespace dq {
namespace slog {

constexpr int kWave = 64;

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i32<CTRL>(static_cast<int>(b)), hi = dpp_i32<CTRL>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// All-reduce of the candidate (|a|, key) over the G lanes of a group: larger
// |a| wins, then the smaller key (position in the high bits, so the first row
// in the current order wins a tie).  The order is total (keys are distinct),
// so every lane ends with the same winner.
__device__ __forceinline__ void take(double& bc, int& key, double oc, int ok) {
  const bool w = oc > bc || (oc == bc && ok < key);
  bc = w ? oc : bc;
  key = w ? ok : key;
}

template <int CTRL>
__device__ __forceinline__ void round_dpp(double& bc, int& key) {
  take(bc, key, dpp_f64<CTRL>(bc), dpp_i32<CTRL>(key));
}

template <int G>
__device__ __forceinline__ void group_best(double& bc, int& key) {
  round_dpp<0xB1>(bc, key);                          // quad_perm [1,0,3,2]
  round_dpp<0x4E>(bc, key);                          // quad_perm [2,3,0,1]
  if constexpr (G >= 8) round_dpp<0x141>(bc, key);   // row_half_mirror
  if constexpr (G >= 16) round_dpp<0x140>(bc, key);  // row_mirror
  if constexpr (G >= 32) take(bc, key, __shfl_xor(bc, 16, kWave), __shfl_xor(key, 16, kWave));
  if constexpr (G >= 64) take(bc, key, __shfl_xor(bc, 32, kWave), __shfl_xor(key, 32, kWave));
}

// The same reduction carrying the candidate's signed value v (the rolled
// factorisation takes its pivot from here instead of from the broadcast row).
template <int CTRL>
__device__ __forceinline__ void round_dpp_v(double& bc, int& key, double& v) {
  const double oc = dpp_f64<CTRL>(bc), ov = dpp_f64<CTRL>(v);
  const int ok = dpp_i32<CTRL>(key);
  const bool w = oc > bc || (oc == bc && ok < key);
  bc = w ? oc : bc;
  key = w ? ok : key;
  v = w ? ov : v;
}

__device__ __forceinline__ void round_xor_v(double& bc, int& key, double& v, int off) {
  const double oc = __shfl_xor(bc, off, kWave), ov = __shfl_xor(v, off, kWave);
  const int ok = __shfl_xor(key, off, kWave);
  const bool w = oc > bc || (oc == bc && ok < key);
  bc = w ? oc : bc;
  key = w ? ok : key;
  v = w ? ov : v;
}

template <int G>
__device__ __forceinline__ void group_best_v(double& bc, int& key, double& v) {
  round_dpp_v<0xB1>(bc, key, v);
  round_dpp_v<0x4E>(bc, key, v);
  if constexpr (G >= 8) round_dpp_v<0x141>(bc, key, v);
  if constexpr (G >= 16) round_dpp_v<0x140>(bc, key, v);
  if constexpr (G >= 32) round_xor_v(bc, key, v, 16);
  if constexpr (G >= 64) round_xor_v(bc, key, v, 32);
}

// Value of lane (base + i) of the group, i = k % G known at compile time after
// unrolling; DPP quad broadcast when the group is a quad.
template <int G>
__device__ __forceinline__ double from_lane(double v, int base, int i) {
  if constexpr (G == 4) {
    switch (i) {
      case 0: return dpp_f64<0x00>(v);
      case 1: return dpp_f64<0x55>(v);
      case 2: return dpp_f64<0xAA>(v);
      default: return dpp_f64<0xFF>(v);
    }
  } else {
    return __shfl(v, base + i, kWave);
  }
}

template <int NC, class T>
__device__ __forceinline__ void load_row(double (&row)[NC], const T* __restrict__ p, int n, bool vec) {
  if (vec) {
    if constexpr (sizeof(T) == 8) {
#pragma unroll
      for (int j = 0; j < NC; j += 2)
        if (j < n) {
          const double2 v = *reinterpret_cast<const double2*>(p + j);
          row[j] = v.x;
          row[j + 1] = v.y;
        }
    } else {
#pragma unroll
      for (int j = 0; j < NC; j += 4)
        if (j < n) {
          const float4 v = *reinterpret_cast<const float4*>(p + j);
          row[j] = v.x;
          row[j + 1] = v.y;
          row[j + 2] = v.z;
          row[j + 3] = v.w;
        }
    }
  } else {
#pragma unroll
    for (int j = 0; j < NC; ++j)
      if (j < n) row[j] = static_cast<double>(p[j]);
  }
}

// Stores n values produced by get(j) at p (contiguous).
template <int NC, class T, class F>
__device__ __forceinline__ void store_row(T* __restrict__ p, int n, bool vec, F&& get) {
  if (vec) {
    if constexpr (sizeof(T) == 8) {
#pragma unroll
      for (int j = 0; j < NC; j += 2)
        if (j < n) *reinterpret_cast<double2*>(p + j) = make_double2(get(j), get(j + 1));
    } else {
#pragma unroll
      for (int j = 0; j < NC; j += 4)
        if (j < n)
          *reinterpret_cast<float4*>(p + j) =
              make_float4(static_cast<float>(get(j)), static_cast<float>(get(j + 1)),
                          static_cast<float>(get(j + 2)), static_cast<float>(get(j + 3)));
    }
  } else {
#pragma unroll
    for (int j = 0; j < NC; ++j)
      if (j < n) p[j] = static_cast<T>(get(j));
  }
}

// Extra doubles after each matrix's L\U tile in LDS so that the rows the
// lanes of one 16- or 32-lane LDS pass store land on distinct banks
// (32 or 64 banks of 4 bytes; checked for every built tile).
constexpr int lds_pad(int NC, int G) {
  return (NC == 12 && G == 8) ? 12 : (NC == 16 && G <= 8) ? G : (NC == 24 && G == 16) ? 24
                                     : (NC == 32 && G == 16) ? 16 : 0;
}

// One step-by-step getf2 factorisation of the group's matrix held in a.
// EXACT = false takes every multiplier as a * (1 / pivot) and reports in
// `tiny` whether some nonzero pivot had |pivot| < DBL_MIN, where the reference
// divides instead; the caller then factors the batch again with EXACT = true,
// which applies that rule.  (A division branch inside every step measured
// slower; so did carrying the signed pivot through the search and a
// branch-free update with selects.)
template <int NC, int G, bool EXACT>
__device__ __forceinline__ void factor(double (&a)[(NC + G - 1) / G][NC], int (&stp)[(NC + G - 1) / G],
                                       double (&pv)[(NC + G - 1) / G], int n, int q, int base, int& swaps,
                                       double& sg, bool& bad, bool& sing, bool& tiny) {
  constexpr int R = (NC + G - 1) / G;
  int pos[R];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const int r = q + G * s;
    act[s] = r < n;
    pos[s] = r;
    stp[s] = r;
    pv[s] = 1.0;
  }
  swaps = 0;
  sg = 1.0;
  bad = sing = tiny = false;
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    if (k < n) {
      // pivot search: the lane's rows, then the group
      double bc = -1.0;
      int key = 0x7FFFFFFF;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        double t = act[s] ? fabs(a[s][k]) : -1.0;
        if (t != t) t = -0.5;  // NaN never wins over a number
        const int ks = (pos[s] << 16) | (q << 8) | s;
        const bool w = t > bc || (t == bc && ks < key);
        bc = w ? t : bc;
        key = w ? ks : key;
      }
      group_best<G>(bc, key);
      const int bp = key >> 16, bq = (key >> 8) & 0xFF, bs = key & 0xFF;
      // the pivot row's columns k.. from lane bq, slot bs
      double prow[NC];
      const int src = base + bq;
#pragma unroll
      for (int j = 0; j < NC; ++j)
        if (j >= k) {  // columns past n hold zeros in the rows of the matrix (never read back)
          double v = a[0][j];
#pragma unroll
          for (int s = 1; s < R; ++s) v = (bs == s) ? a[s][j] : v;
          prow[j] = __shfl(v, src, kWave);
        }
      const double piv = prow[k];
      swaps += (bp != k);
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const bool me = (q == bq) && (s == bs);
        act[s] = act[s] && !me;
        stp[s] = me ? k : stp[s];
        pos[s] = me ? k : (pos[s] == k ? bp : pos[s]);
      }
      const double ad = fabs(piv);
      bad = bad || !(ad > 0.0 && isfinite(ad));
      sg = (piv < 0.0) ? -sg : sg;
      pv[k / G] = (q == k % G) ? piv : pv[k / G];
      if (piv == 0.0) {
        sing = true;
      } else {
        const double rp = 1.0 / piv;
        const bool small = !(ad >= DBL_MIN);
        tiny = tiny || small;
        double l[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
          l[s] = a[s][k] * rp;
          if constexpr (EXACT) l[s] = small ? a[s][k] / piv : l[s];
        }
#pragma unroll
        for (int s = 0; s < R; ++s) {
          if (act[s]) {
            a[s][k] = l[s];
#pragma unroll
            for (int j = 0; j < NC; ++j)
              if (j > k) a[s][j] = a[s][j] - l[s] * prow[j];
          }
        }
      }
    }
  }
}

// The same factorisation with the step loop rolled, for n > 32 where the
// unrolled steps (quadratic in n) no longer fit the instruction cache.  k is
// wave-uniform; register columns keep static indices and the step touches
// every column with selects (column k takes the multiplier, columns past k
// the update, the rest keep their value), so a step is one basic block in
// which the pivot-row broadcasts pipeline.  The pivot's signed value comes
// with the search, so the multipliers are ready before the row arrives.
template <int NC, int G, bool EXACT>
__device__ __forceinline__ void factor_rolled(double (&a)[(NC + G - 1) / G][NC], int (&stp)[(NC + G - 1) / G],
                                              double (&pv)[(NC + G - 1) / G], int n, int q, int base, int& swaps,
                                              double& sg, bool& bad, bool& sing, bool& tiny) {
  constexpr int R = (NC + G - 1) / G;
  int pos[R];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const int r = q + G * s;
    act[s] = r < n;
    pos[s] = r;
    stp[s] = r;
    pv[s] = 1.0;
  }
  swaps = 0;
  sg = 1.0;
  bad = sing = tiny = false;
#pragma unroll 1
  for (int k = 0; k < n; ++k) {
    double ak[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      ak[s] = a[s][0];
#pragma unroll
      for (int j = 1; j < NC; ++j) ak[s] = (j == k) ? a[s][j] : ak[s];
    }
    double bc = -1.0, bv = 0.0;
    int key = 0x7FFFFFFF;
#pragma unroll
    for (int s = 0; s < R; ++s) {
      double t = act[s] ? fabs(ak[s]) : -1.0;
      if (t != t) t = -0.5;  // NaN never wins over a number
      const int ks = (pos[s] << 16) | (q << 8) | s;
      const bool w = t > bc || (t == bc && ks < key);
      bc = w ? t : bc;
      key = w ? ks : key;
      bv = w ? ak[s] : bv;
    }
    group_best_v<G>(bc, key, bv);
    const int bp = key >> 16, bq = (key >> 8) & 0xFF, bs = key & 0xFF;
    const double piv = bv;
    const int src = base + bq;
    swaps += (bp != k);
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const bool me = (q == bq) && (s == bs);
      act[s] = act[s] && !me;
      stp[s] = me ? k : stp[s];
      pos[s] = me ? k : (pos[s] == k ? bp : pos[s]);
      pv[s] = (q + G * s == k) ? piv : pv[s];
    }
    const double ad = fabs(piv);
    bad = bad || !(ad > 0.0 && isfinite(ad));
    sg = (piv < 0.0) ? -sg : sg;
    const bool zero = piv == 0.0;
    sing = sing || zero;
    const bool small = !(ad >= DBL_MIN);
    tiny = tiny || (small && !zero);
    const double rp = 1.0 / piv;
    double l[R];
    bool upd[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      l[s] = ak[s] * rp;
      if constexpr (EXACT) l[s] = small ? ak[s] / piv : l[s];
      upd[s] = act[s] && !zero;
    }
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      double v = a[0][j];
#pragma unroll
      for (int s = 1; s < R; ++s) v = (bs == s) ? a[s][j] : v;
      const double p = __shfl(v, src, kWave);
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const double u = a[s][j] - l[s] * p;
        const double w = (j == k) ? l[s] : u;
        a[s][j] = (upd[s] && j >= k && j < n) ? w : a[s][j];
      }
    }
  }
}

// Columns c0 .. c0 + R - 1 of A^{-T} from L\U in LDS (row stride RS): U^T y =
// e_c, then L^T w = y, in the reference's operation order.
// PAD: rows n.. of L (LDS) are identity rows and zt is zero past n, so the
// back substitution runs over every k > i (a k past n subtracts 0 * 0 = +0,
// which leaves the sum unchanged) instead of testing k < n.
template <int NC, int R, int RS, bool PAD>
__device__ __forceinline__ void solve_cols(const double* L, double (&zt)[R][NC], int n, int c0) {
  if constexpr (PAD) {
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int k = 0; k < NC; ++k)
        if (k >= n) zt[t][k] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    if (i < n) {
      double acc[R];
#pragma unroll
      for (int t = 0; t < R; ++t) acc[t] = (i == c0 + t) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < NC; ++k)
        if (k < i) {
          const double u = L[k * RS + i];
#pragma unroll
          for (int t = 0; t < R; ++t) acc[t] -= u * zt[t][k];
        }
      const double d = L[i * RS + i];
#pragma unroll
      for (int t = 0; t < R; ++t) zt[t][i] = acc[t] / d;
    }
  }
#pragma unroll
  for (int i = NC - 1; i >= 0; --i) {
    if (i < n) {
      double acc[R];
#pragma unroll
      for (int t = 0; t < R; ++t) acc[t] = zt[t][i];
#pragma unroll
      for (int k = 0; k < NC; ++k)
        if (k > i && (PAD || k < n)) {
          const double l = L[k * RS + i];
#pragma unroll
          for (int t = 0; t < R; ++t) acc[t] -= l * zt[t][k];
        }
#pragma unroll
      for (int t = 0; t < R; ++t) zt[t][i] = acc[t];
    }
  }
}

// The same solve with the row loops rolled (n > 32): i is wave-uniform and
// every column takes part through selects (a column outside the sum
// subtracts nothing: the select picks +0, and x - (+0) = x), so each row of
// the solve is one basic block.
template <int NC, int R, int RS>
__device__ __forceinline__ void solve_cols_rolled(const double* L, double (&zt)[R][NC], int n, int c0) {
#pragma unroll
  for (int t = 0; t < R; ++t)
#pragma unroll
    for (int k = 0; k < NC; ++k) zt[t][k] = 0.0;
#pragma unroll 1
  for (int i = 0; i < n; ++i) {
    double acc[R];
#pragma unroll
    for (int t = 0; t < R; ++t) acc[t] = (i == c0 + t) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const double u = L[k * RS + i];
#pragma unroll
      for (int t = 0; t < R; ++t) acc[t] -= (k < i) ? u * zt[t][k] : 0.0;
    }
    const double d = L[i * RS + i];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const double v = acc[t] / d;
#pragma unroll
      for (int k = 0; k < NC; ++k) zt[t][k] = (k == i) ? v : zt[t][k];
    }
  }
#pragma unroll 1
  for (int i = n - 1; i >= 0; --i) {
    double acc[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      acc[t] = zt[t][0];
#pragma unroll
      for (int k = 1; k < NC; ++k) acc[t] = (k == i) ? zt[t][k] : acc[t];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const double l = L[k * RS + i];
#pragma unroll
      for (int t = 0; t < R; ++t) acc[t] -= (k > i && k < n) ? l * zt[t][k] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int k = 0; k < NC; ++k) zt[t][k] = (k == i) ? acc[t] : zt[t][k];
  }
}

// MODE: 0 forward, 1 forward + LU/perm out, 2 backward refactorising A,
// 3 backward from LU/perm.  vec_in / vec_out: rows start on 16-byte
// boundaries and n is a multiple of the 16-byte vector width.
// FULL: n == NC, so every bound on n is known to the compiler (no guards on
// the columns and steps past n).
// The fixed-bound backward kernels with four lanes per matrix sit at the
// 256-register edge of two waves per SIMD; the bound keeps them there.
template <int NC, int G, int MODE, class T, bool FULL>
__global__ __launch_bounds__(kWave, (FULL && MODE >= 2 && G == 4) ? 2 : 1) void slogdet_rows(const T* __restrict__ A, const int32_t* __restrict__ perm_in,
                                                      int64_t B, int n_in, T* __restrict__ sign,
                                                      T* __restrict__ logabs, const T* __restrict__ cot,
                                                      T* __restrict__ out_mat, int32_t* __restrict__ perm_out,
                                                      bool vec_in, bool vec_out) {
  const int n = FULL ? NC : n_in;
  constexpr int R = (NC + G - 1) / G;  // rows per lane
  constexpr int M = kWave / G;         // matrices per wave
  constexpr int RS = NC + 1;           // LDS row stride in doubles (odd: a group's rows on distinct banks)
  constexpr int STR = NC * RS + lds_pad(NC, G);  // LDS doubles per matrix
  constexpr bool ROLL = NC >= 48;      // rolled step loops (instruction cache)
  __shared__ double tile[(MODE >= 2) ? M * STR : 1];
  __shared__ int perm_s[(MODE >= 2) ? M * NC : 1];
  // forward + LU output through LDS (measured: direct stores from the lanes'
  // rows put one 16-byte piece per lane and instruction on rows far apart).
  // At 12 columns the wave's L\U tiles and permutations, dense and in output
  // order, leave in contiguous stores (up to 20 KB per wave); at 16 one row
  // slot at a time, each row a 128-byte line stored by consecutive lanes,
  // so that the buffer stays at 8.5 KB.  Longer rows store well enough
  // directly.
  constexpr bool STAGE = MODE == 1 && NC < 16 && M * NC * NC * 8 <= 20480;
  constexpr bool STAGE_ROWS = MODE == 1 && NC == 16;
  constexpr int RSR = NC + 1;  // odd row stride: one column of the lanes' rows on distinct banks
  __shared__ __attribute__((aligned(16))) double lu_s[STAGE ? M * NC * NC : 2];
  __shared__ int perm1_s[STAGE ? M * NC : 1];
  __shared__ double row_s[STAGE_ROWS ? kWave * RSR : 1];
  __shared__ int64_t dst_s[STAGE_ROWS ? kWave : 1];
  const int lane = threadIdx.x;
  const int g = lane / G, q = lane % G, base = g * G;
  const int64_t m = static_cast<int64_t>(blockIdx.x) * M + g;
  const bool valid = m < B;
  const size_t nn = static_cast<size_t>(n) * n;
  const T* Am = A + (valid ? m : 0) * nn;

  double a[R][NC];
  int stp[R];
  auto load = [&]() {
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int r = q + G * s;
#pragma unroll
      for (int j = 0; j < NC; ++j) a[s][j] = (j == r) ? 1.0 : 0.0;
      stp[s] = r;
      if (valid && r < n) load_row<NC>(a[s], Am + static_cast<size_t>(r) * n, n, vec_in);
    }
  };
  load();

  bool sing = false;
  if constexpr (MODE != 3) {
    int swaps;
    double sg;
    bool bad, tiny;
    double pv[R];  // pv[s]: pivot of step q + G s
    if constexpr (ROLL)
      factor_rolled<NC, G, false>(a, stp, pv, n, q, base, swaps, sg, bad, sing, tiny);
    else
      factor<NC, G, false>(a, stp, pv, n, q, base, swaps, sg, bad, sing, tiny);
    if (__builtin_expect(__ballot(tiny) != 0, 0)) {
      // a pivot below DBL_MIN: factor again with the reference's division rule
      load();
      if constexpr (ROLL)
        factor_rolled<NC, G, true>(a, stp, pv, n, q, base, swaps, sg, bad, sing, tiny);
      else
        factor<NC, G, true>(a, stp, pv, n, q, base, swaps, sg, bad, sing, tiny);
    }
    if constexpr (MODE <= 1) {
      double lg[R];
#pragma unroll
      for (int s = 0; s < R; ++s) lg[s] = (q + G * s < n) ? log(fabs(pv[s])) : 0.0;
      double la = 0.0;
#pragma unroll
      for (int k = 0; k < NC; ++k)
        if (k < n) la += from_lane<G>(lg[k / G], base, k % G);
      if (valid && q == 0) {
        const bool b = bad || sing;
        sign[m] = static_cast<T>(b ? 0.0 : ((swaps & 1) ? -sg : sg));
        logabs[m] = static_cast<T>(b ? -INFINITY : la);
      }
      if constexpr (STAGE_ROWS) {
        const int C = vec_out ? n / 2 : n;  // 16-byte (or 8-byte) pieces per row
        const int rpi = kWave / C;          // rows per store instruction
        const int lr = lane / C, lc = lane % C;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int r = q + G * s;
          const bool has = valid && r < n;
          if (has) {
#pragma unroll
            for (int j = 0; j < NC; ++j)
              if (j < n) row_s[lane * RSR + j] = a[s][j];
            perm_out[m * n + stp[s]] = r;
          }
          dst_s[lane] = has ? m * static_cast<int64_t>(nn) + stp[s] * n : -1;
          __syncthreads();
          for (int r0 = 0; r0 < kWave; r0 += rpi) {
            const int row = r0 + lr;
            if (lr < rpi && row < kWave) {
              const int64_t d = dst_s[row];
              if (d >= 0) {
                if (vec_out)
                  *reinterpret_cast<double2*>(out_mat + d + 2 * lc) =
                      make_double2(row_s[row * RSR + 2 * lc], row_s[row * RSR + 2 * lc + 1]);
                else
                  out_mat[d + lc] = row_s[row * RSR + lc];
              }
            }
          }
          __syncthreads();
        }
      }
      if constexpr (MODE == 1 && !STAGE && !STAGE_ROWS) {
        if (valid) {
#pragma unroll
          for (int s = 0; s < R; ++s) {
            const int r = q + G * s;
            if (r < n) {
              store_row<NC>(out_mat + m * nn + static_cast<size_t>(stp[s]) * n, n, vec_out,
                            [&](int j) { return a[s][j]; });
              perm_out[m * n + stp[s]] = r;
            }
          }
        }
      }
      if constexpr (STAGE) {
        // Row stp[s] of matrix g goes to lu_s[g nn + stp[s] n ..]; then the
        // wave's valid matrices (consecutive in the batch) leave as one
        // contiguous run of mv nn doubles and mv n permutation entries.
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int r = q + G * s;
          if (r < n) {
            double* row = lu_s + g * static_cast<int>(nn) + stp[s] * n;
#pragma unroll
            for (int j = 0; j < NC; ++j)
              if (j < n) row[j] = a[s][j];
            perm1_s[g * n + stp[s]] = r;
          }
        }
        __syncthreads();
        const int64_t m0 = static_cast<int64_t>(blockIdx.x) * M;
        const int mv = static_cast<int>(B - m0 < M ? B - m0 : M);
        const int tot = mv * static_cast<int>(nn);
        T* dst = out_mat + m0 * static_cast<int64_t>(nn);
        if (vec_out) {  // n even and dst 16-byte aligned
          for (int e = 2 * lane; e < tot; e += 2 * kWave)
            *reinterpret_cast<double2*>(dst + e) = *reinterpret_cast<const double2*>(lu_s + e);
        } else {
          for (int e = lane; e < tot; e += kWave) dst[e] = lu_s[e];
        }
        for (int e = lane; e < mv * n; e += kWave) perm_out[m0 * n + e] = perm1_s[e];
      }
      return;
    }
  } else {
    // L\U given in position order: singular iff some u_kk is zero
    bool z = false;
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int r = q + G * s;
      double d = 1.0;
#pragma unroll
      for (int j = 0; j < NC; ++j)
        if (j == r) d = a[s][j];
      z |= (r < n && d == 0.0);
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) z |= __shfl_xor(static_cast<int>(z), off, kWave) != 0;
    sing = z;
  }

  // ---- backward: L\U and the permutation to LDS, position order
  double* L = tile + g * STR;
  int* P = perm_s + g * NC;
  // with n below the bucket every row (rows n.. are identity rows: their stp
  // is their index, and the unrolled solve reads them as zeros below the
  // diagonal)
  constexpr bool ALL_ROWS = !FULL && !ROLL;
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const int r = q + G * s;
    if (r < (ALL_ROWS ? NC : n)) {
      const int at = stp[s];
#pragma unroll
      for (int j = 0; j < NC; ++j)
        if (ALL_ROWS || j < n) L[at * RS + j] = a[s][j];
      if (!ALL_ROWS || r < n) {
        if constexpr (MODE == 3)
          P[at] = valid ? perm_in[m * n + r] : r;
        else
          P[at] = r;
      }
    }
  }
  __syncthreads();
  // lane q solves columns q R .. q R + R - 1 of A^{-T}; the group's L\U reads
  // are one LDS address per group (a broadcast)
  const int c0 = q * R;
  double zt[R][NC];
  if constexpr (ROLL)
    solve_cols_rolled<NC, R, RS>(L, zt, n, c0);
  else
    solve_cols<NC, R, RS, ALL_ROWS>(L, zt, n, c0);
  if (!valid) return;
  // X[perm[k]][c] = cot * w_c[k]: per k the group writes one contiguous row
  const double cc = static_cast<double>(cot[m]);
  T* Xm = out_mat + m * nn;
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    if (k < n) {
      T* row = Xm + static_cast<size_t>(P[k]) * n;
#pragma unroll
      for (int t = 0; t < R; ++t)
        if (c0 + t < n) row[c0 + t] = static_cast<T>(sing ? 0.0 : cc * zt[t][k]);
    }
  }
}

template <int NC, int G, int MODE, class T>
void launch_rows(const SlogArgs& x) {
  constexpr int M = kWave / G;
  const int64_t blocks = (x.B + M - 1) / M;
  DQ_REQUIRE(blocks < (1LL << 31), DQ_EINVAL, "batch too large");
  constexpr int VW = 16 / static_cast<int>(sizeof(T));
  const bool nvec = x.n % VW == 0;
  const bool vec_in = nvec && (reinterpret_cast<uintptr_t>(x.A) % 16 == 0);
  const bool vec_out = nvec && (reinterpret_cast<uintptr_t>(x.out_mat) % 16 == 0);
  auto* kern = slogdet_rows<NC, G, MODE, T, false>;
  if constexpr (NC <= 32) {  // the unrolled buckets
    if (x.fixed_n && x.n == NC) kern = slogdet_rows<NC, G, MODE, T, true>;
  }
  kern<<<dim3(static_cast<unsigned>(blocks)), kWave, 0, x.s>>>(
      static_cast<const T*>(x.A), x.perm_in, x.B, x.n, static_cast<T*>(x.sign), static_cast<T*>(x.logabs),
      static_cast<const T*>(x.cot), static_cast<T*>(x.out_mat), x.perm_out, vec_in, vec_out);
}

template <int NC, int G>
void launch_rows_mode(int mode, bool f32, const SlogArgs& x) {
  if (f32) {
    if (mode == 0)
      launch_rows<NC, G, 0, float>(x);
    else if (mode == 2)
      launch_rows<NC, G, 2, float>(x);
    else
      throw Error(DQ_EINVAL, "internal: FP32 slogdet has modes 0 and 2 only");
    return;
  }
  switch (mode) {
    case 0: launch_rows<NC, G, 0, double>(x); break;
    case 1: launch_rows<NC, G, 1, double>(x); break;
    case 2: launch_rows<NC, G, 2, double>(x); break;
    case 3: launch_rows<NC, G, 3, double>(x); break;
    default: throw Error(DQ_EINVAL, "internal: bad slogdet mode");
  }
}

}  // namespace slog
}  // namespace dq
