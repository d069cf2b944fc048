// Projection of H onto the selected space V (H_VV, H_VP, P), the fused
// variational energy, and the H_eff SpGEMM down-folding, on gfx950.
//
//  * DetMap (det_space.cpp:29-65) is an open-addressing hash in HBM: 128-bit
//    keys, u32 values, load factor <= 1/2.  Insertion claims slots with a CAS
//    on the value word; a second pass finds, for every key, the first slot
//    holding it in probe order and lowers its value to the smallest index
//    with that key (first occurrence, as from_ordered keeps), flagging
//    duplicates for the strict mode.
//  * stream_build (build_ham.cpp:213-314) with the KnownSets / StaticHB /
//    DynamicAmp policies (build_ham.cpp:78-196) is a count pass + scan + fill
//    pass over the shared row walker; COO output is canonicalised by a stable
//    radix sort on (row << 32 | col) (sort_and_merge_coo, ham_utils.cpp:41-83);
//    P is the sorted unique set of external keys (finalize_vp_deferred,
//    build_ham.cpp:346-380).  For n_orb <= 32 each row's H_VP entries reach
//    their (row, column) order without a sort by row: a ket's place in its
//    row follows from its position among the row's connections (rowrank.hpp),
//    and one radix sort of the ket keys (colex ranks, 41 bits at 28 orbitals
//    7a 7b) ranks P and gives each entry its column.
//  * compute_variational_energy (local_conn.cpp:148-229): each lane sums its
//    own candidates in generation order, the wave combines the 64 partial sums
//    by a fixed butterfly, rows are combined by a fixed tree: the result does
//    not depend on launch shape or on what else is in the batch.
//  * get_ham_eff (ham_eff.cpp:35-233): each row of V accumulates its terms in
//    an LDS window over j, one column of H_VP at a time in ascending k, so
//    each element is summed in the reference's order; V wider than four
//    windows emits every contribution (i, j, k) in ascending k, stably sorts
//    them by (i, j) and folds them sequentially.
#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "assemble.hpp"
#include "batch.hpp"
#include "eloc_strings.hpp"
#include "hash.hpp"
#include "rowrank.hpp"
#include "sort.hpp"
#include "walk.hpp"

struct dq_blocks {
  dq::DevBuf<dq::u32> vv_row, vv_col, vp_row, vp_col;
  dq::DevBuf<double> vv_val, vp_val;
  dq::DevBuf<dq::u64> det_p;
  int64_t nnz_vv = 0, nnz_vp = 0, n_p = 0;
};

namespace dq {

namespace {

constexpr int kWavesPerBlock = 4;

__global__ void hash_insert(const u64* __restrict__ dets, int64_t n, u64* __restrict__ keys, u32* __restrict__ vals,
                            u64 mask) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 a = dets[2 * i], b = dets[2 * i + 1];
  u64 s = det_hash(a, b) & mask;
  while (true) {
    if (atomicCAS(&vals[s], kEmpty, static_cast<u32>(i)) == kEmpty) {
      keys[2 * s] = a;
      keys[2 * s + 1] = b;
      return;
    }
    s = (s + 1) & mask;
  }
}

__global__ void hash_resolve(const u64* __restrict__ dets, int64_t n, const u64* __restrict__ keys,
                             u32* __restrict__ vals, u64 mask, int* __restrict__ dup) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 a = dets[2 * i], b = dets[2 * i + 1];
  u64 s = det_hash(a, b) & mask;
  while (true) {
    if (keys[2 * s] == a && keys[2 * s + 1] == b) {
      if (vals[s] != static_cast<u32>(i)) {
        if (dup) atomicOr(dup, 1);
        atomicMin(&vals[s], static_cast<u32>(i));
      }
      return;
    }
    s = (s + 1) & mask;
  }
}

}  // namespace

__global__ void fill_u32_kernel(u32* __restrict__ p, size_t n, u32 v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

void fill_u32(void* p, size_t words, u32 v, hipStream_t s) {
  if (words == 0) return;
  const size_t blocks = std::min<size_t>((words + 255) / 256, 4096);
  fill_u32_kernel<<<static_cast<unsigned>(blocks), 256, 0, s>>>(static_cast<u32*>(p), words, v);
  DQ_HIP(hipGetLastError());
}

u64 hash_capacity(int64_t n) {
  u64 cap = 16;
  while (cap < static_cast<u64>(2 * n)) cap <<= 1;
  return cap;
}

namespace {
__global__ void filter_bits(const u64* __restrict__ pairs, int64_t n, int lg, u32* __restrict__ bits) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 h = filt_slot(pairs[2 * i], pairs[2 * i + 1], lg);
  atomicOr(bits + (h >> 5), 1u << (h & 31));
}
}  // namespace

void filter_build(const u64* pairs, int64_t n, int lg, u32* bits, hipStream_t s) {
  DQ_HIP(hipMemsetAsync(bits, 0, (size_t(1) << lg) / 8, s));
  if (n > 0) filter_bits<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(pairs, n, lg, bits);
  DQ_HIP(hipGetLastError());
}

void hash_build_async(const u64* dets, int64_t n, u64* keys, u32* vals, u64 mask, hipStream_t s) {
  DQ_REQUIRE(n < static_cast<int64_t>(kEmpty), DQ_ERUNTIME, "DetMap size exceeds u32 index capacity");
  const u64 cap = mask + 1;
  DQ_REQUIRE(cap >= static_cast<u64>(2 * n), DQ_EINVAL, "hash capacity below twice the key count");
  fill_u32(vals, cap, 0xFFFFFFFFu, s);
  fill_u32(keys, 4 * cap, 0xFFFFFFFFu, s);
  if (n == 0) return;
  const unsigned g = static_cast<unsigned>((n + 255) / 256);
  hash_insert<<<g, 256, 0, s>>>(dets, n, keys, vals, mask);
  hash_resolve<<<g, 256, 0, s>>>(dets, n, keys, vals, mask, nullptr);
  DQ_HIP(hipGetLastError());
}

bool DetHash::build(const u64* dets, int64_t n, hipStream_t s, bool ordered) {
  if (n == 0) return false;
  DQ_REQUIRE(n < static_cast<int64_t>(kEmpty), DQ_ERUNTIME, "DetMap size exceeds u32 index capacity");
  const u64 cap = hash_capacity(n);
  mask = cap - 1;
  if (ordered) {
    keys.alloc(2 * cap, s);
    vals.alloc(cap, s);
  } else {
    keys.alloc(2 * cap);
    vals.alloc(cap);
  }
  DQ_HIP(hipMemsetAsync(vals.p, 0xFF, cap * sizeof(u32), s));
  DQ_HIP(hipMemsetAsync(keys.p, 0xFF, 2 * cap * sizeof(u64), s));
  DevBuf<int> dup(1, s);
  DQ_HIP(hipMemsetAsync(dup.p, 0, sizeof(int), s));
  const unsigned g = static_cast<unsigned>((n + 255) / 256);
  hash_insert<<<g, 256, 0, s>>>(dets, n, keys.p, vals.p, mask);
  hash_resolve<<<g, 256, 0, s>>>(dets, n, keys.p, vals.p, mask, dup.p);
  DQ_HIP(hipGetLastError());
  int h = 0;
  DQ_HIP(hipMemcpyAsync(&h, dup.p, sizeof(int), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  return h != 0;
}

namespace {

// ------------------------------------------------------------------ projected assembly
enum Mode { kKnown = 0, kStatic = 1, kDynamic = 2 };

// H_VP in row order without a sort by row (n_orb <= 32, no external set).
// Each row owns the range [ub[row], ub[row + 1]) of its kept connections (a
// count that does not split H_VV from H_VP); the fill writes H_VP entries from
// the front -- the ket's key, its position among the row's connections
// (rowrank.hpp) and the value -- and H_VV entries from the back.  ord_permute
// turns positions into each entry's place in its row and moves the values
// there through LDS; vv_gather compacts H_VV.
struct ProjOrd {
  u64* ket;         // fill: ket key (H_VP) or row << 32 | column (H_VV)
  u32* pos;         // fill: the ket's position among the row's connections
  const int64_t* ub;
  u64 cb;           // C(n_orb, nb): colex keys ra * cb + rb; 0: packed alpha << n_orb | beta
};

// The projection's prefilter of V's map: 2^17 bits (16 KB, read through the
// vector cache); a ket whose bit is clear is not in V and skips the probe.
constexpr int kProjFiltLg = 17;

template <bool FILL, bool ORD = false>
struct ProjSink {
  static constexpr bool kNeedKet = true;
  static constexpr bool kCountOnly = false;
  __device__ __forceinline__ void add_doubles(int) {}
  int mode;
  double psi_i, eps;
  u64 row, xa, xb;
  HashView HV, HE;
  const u32* vfilt;  // prefilter of HV (2^kProjFiltLg bits), or null
  int64_t pv, pp;
  u64* vv_key;
  double* vv_val;
  u64* vp_key;
  u64* vp_det;
  double* vp_val;
  // row order (ORD)
  const RowRank* T;
  ProjOrd O;
  int64_t e0, e1;  // the row's range
  int n_orb;

  __device__ __forceinline__ void put(bool keep, u64 ka, u64 kb, double v) {
    long jv = -1, je = -1;
    if (keep) {
      if (!vfilt || filt_test(vfilt, filt_slot(ka, kb, kProjFiltLg))) jv = hash_find(HV, ka, kb);
      if (jv < 0 && mode == kKnown) je = hash_find(HE, ka, kb);
    }
    const bool kv = keep && jv >= 0;
    const bool kp = keep && jv < 0 && (mode != kKnown || je >= 0);
    const u64 mv = __ballot(kv), mp = __ballot(kp);
    if (FILL && ORD) {
      const u64 lt = lanemask_lt();
      if (kv) {
        const int64_t at = e1 - 1 - (pv + __popcll(mv & lt));
        if (at >= e0) {
          O.ket[at] = (row << 32) | static_cast<u64>(jv);
          vp_val[at] = v;
        }
      }
      if (kp) {
        const int64_t at = e0 + pp + __popcll(mp & lt);
        u32 ra, rb;
        const u32 L = ket_order<true>(*T, static_cast<u32>(xa), static_cast<u32>(xb), static_cast<u32>(ka),
                                      static_cast<u32>(kb), ra, rb);
        if (at < e1) {
          O.ket[at] = O.cb ? static_cast<u64>(ra) * O.cb + rb : (ka << n_orb) | kb;
          O.pos[at] = L;
          vp_val[at] = v;
        }
      }
    } else if (FILL) {
      const u64 lt = lanemask_lt();
      if (kv) {
        const int64_t at = pv + __popcll(mv & lt);
        vv_key[at] = (row << 32) | static_cast<u64>(jv);
        vv_val[at] = v;
      }
      if (kp) {
        const int64_t at = pp + __popcll(mp & lt);
        if (mode == kKnown) {
          vp_key[at] = (row << 32) | static_cast<u64>(je);
        } else {
          vp_key[at] = row;
          vp_det[2 * at] = ka;
          vp_det[2 * at + 1] = kb;
        }
        vp_val[at] = v;
      }
    }
    pv += __popcll(mv);
    pp += __popcll(mp);
  }
  __device__ __forceinline__ void diag(double d) {
    const bool keep = __lane_id() == 0 && fabs(d) > kMatThr;
    const u64 mv = __ballot(keep);
    if (FILL && ORD && keep && e1 - 1 - pv >= e0) {
      O.ket[e1 - 1 - pv] = (row << 32) | row;
      vp_val[e1 - 1 - pv] = d;
    } else if (FILL && !ORD && keep) {
      vv_key[pv] = (row << 32) | row;
      vv_val[pv] = d;
    }
    pv += __popcll(mv);
  }
  __device__ __forceinline__ void single(bool ok, u64 ka, u64 kb, double v) {
    bool keep = ok && fabs(v) > kMatThr;
    if (mode == kDynamic) keep = keep && (fabs(v * psi_i) >= eps);  // build_ham.cpp:186-188
    put(keep, ka, kb, v);
  }
  __device__ __forceinline__ void dbl(bool ok, u64 ka, u64 kb, double v) { put(ok && fabs(v) > kMatThr, ka, kb, v); }
};

// Dynamic LDS of the row-order fill: a RowRank per wave.
constexpr size_t kOrdLds = sizeof(RowRank) * kWavesPerBlock;

template <bool FILL, bool ORD = false>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void proj_kernel(
    DevView V, const uint64_t* __restrict__ dets, int64_t n_rows, int n_orb, int mode, int dbl_mode, double eps,
    const double* __restrict__ psi, const u32* __restrict__ hb_len, HashView HV, HashView HE,
    const u32* __restrict__ vfilt, int64_t* __restrict__ cnt_vv, int64_t* __restrict__ cnt_vp,
    const int64_t* __restrict__ off_vv,
    const int64_t* __restrict__ off_vp, u64* vv_key, double* vv_val, u64* vp_key, u64* vp_det, double* vp_val,
    ProjOrd O, int* __restrict__ err) {
  __shared__ RowLists lists[kWavesPerBlock];
  extern __shared__ u32 ord_lds[];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wid;
  RowLists& L = lists[wid];
  const bool active = row < n_rows;
  u64 xa = 0, xb = 0;
  if (active) {
    xa = dets[2 * row];
    xb = dets[2 * row + 1];
  }
  const bool valid = active && !((xa | xb) & ~low_mask(n_orb));
  build_lists(L, xa, xb, n_orb);
  RowRank* T = nullptr;
  if constexpr (ORD) {
    T = reinterpret_cast<RowRank*>(ord_lds) + wid;
    if (valid) build_row_rank(*T, kBinomDev, static_cast<u32>(xa), static_cast<u32>(xb), n_orb);
  }
  __syncthreads();
  if (!active) return;
  if (!valid) {
    if (!FILL && lane == 0) {
      cnt_vv[row] = 0;
      cnt_vp[row] = 0;
      atomicOr(err, 1);
    }
    return;
  }
  ProjSink<FILL, ORD> S;
  S.mode = mode;
  S.psi_i = psi ? psi[row] : 0.0;
  S.eps = eps;
  S.row = static_cast<u64>(row);
  S.xa = xa;
  S.xb = xb;
  S.HV = HV;
  S.HE = HE;
  S.vfilt = vfilt;
  S.pv = (FILL && !ORD) ? off_vv[row] : 0;
  S.pp = (FILL && !ORD) ? off_vp[row] : 0;
  S.vv_key = vv_key;
  S.vv_val = vv_val;
  S.vp_key = vp_key;
  S.vp_det = vp_det;
  S.vp_val = vp_val;
  S.T = T;
  S.O = O;
  S.e0 = ORD ? O.ub[row] : 0;
  S.e1 = ORD ? O.ub[row + 1] : 0;
  S.n_orb = n_orb;
  double cut = 0.0;
  if (dbl_mode == 2) cut = eps / fmax(fabs(S.psi_i), 1e-12);  // build_ham.cpp:181-183
  const RowEnv R = make_env(V, xa, xb, n_orb, L);
  walk_row(R, S, dbl_mode, hb_len, cut);
  if ((!FILL || ORD) && lane == 0) {
    cnt_vv[row] = S.pv;
    cnt_vp[row] = S.pp;
    if (ORD && S.pv + S.pp != S.e1 - S.e0) atomicOr(err, 4);  // the range count disagrees with the walk
  }
}

__global__ void add_i64(const int64_t* __restrict__ a, const int64_t* __restrict__ b, int64_t n,
                        int64_t* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// H_VV entries from the back of each row's range to their compact place.
// (their slots then hold `sentinel`, above every ket key, so the ket sort can
// run over the whole range array and leave them at its end).
__global__ void vv_gather(const int64_t* __restrict__ ub, const int64_t* __restrict__ cnt_vv,
                          const int64_t* __restrict__ off_vv, int64_t n_rows, u64* __restrict__ key,
                          const double* __restrict__ val, u64* __restrict__ okey, double* __restrict__ oval,
                          u64 sentinel) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (row >= n_rows) return;
  const int64_t c = cnt_vv[row], src = ub[row + 1] - c, dst = off_vv[row];
  for (int64_t i = __lane_id(); i < c; i += kWave) {
    okey[dst + i] = key[src + i];
    oval[dst + i] = val[src + i];
    key[src + i] = sentinel;
  }
}

// One block per row: the row's H_VP entries (emission order from ub[row])
// mark their positions in an LDS bitmap; a prefix count over it gives each
// entry's place in the row.  The place goes out as the entry's sort payload
// (emission layout), and the values go to their places through an LDS stage
// of kPermStage places per sweep (a 28-orbital 7a 7b row, up to 30,723
// entries, takes two sweeps).  Staging the values themselves measured 4.0 ms
// on 6.1e8 entries against 6.6 ms for staging indices and gathering the
// values.  A row whose marks fall short of its count had two kets on one
// position.
constexpr int kPermThreads = 1024;
constexpr int kPermStage = 16384;
constexpr int kPermUnroll = 8;

inline size_t perm_lds_bytes(int words) {
  return sizeof(double) * kPermStage + 2 * sizeof(u32) * static_cast<size_t>(words);
}

__global__ __launch_bounds__(kPermThreads) void ord_permute(
    const int64_t* __restrict__ ub, const int64_t* __restrict__ cnt_vp, const int64_t* __restrict__ off, int words,
    const u32* __restrict__ pos, const double* __restrict__ val_em, u32* __restrict__ opos,
    double* __restrict__ val, u32* __restrict__ row_out, int* __restrict__ err) {
  extern __shared__ double perm_lds[];
  double* stage = perm_lds;
  u32* bm = reinterpret_cast<u32*>(perm_lds + kPermStage);
  u32* pre = bm + words;
  __shared__ u32 wsum[kPermThreads / kWave];
  const int64_t row = blockIdx.x;
  const int64_t e0 = ub[row], cnt = cnt_vp[row], o0 = off[row];
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  for (int i = tid; i < words; i += kPermThreads) bm[i] = 0u;
  __syncthreads();
  const u32 nbits = static_cast<u32>(words) * 32u;
  // loops over the row take kPermUnroll entries per thread at a time, their
  // loads issued together
  for (int64_t b = 0; b < cnt; b += static_cast<int64_t>(kPermThreads) * kPermUnroll) {
    u32 L[kPermUnroll];
#pragma unroll
    for (int k = 0; k < kPermUnroll; ++k) {
      const int64_t e = b + k * kPermThreads + tid;
      L[k] = e < cnt ? pos[e0 + e] : nbits;
    }
#pragma unroll
    for (int k = 0; k < kPermUnroll; ++k)
      if (L[k] < nbits) atomicOr(&bm[L[k] >> 5], 1u << (L[k] & 31u));
  }
  __syncthreads();
  // exclusive prefix of the word popcounts: chunks of kPermThreads words
  u32 carry = 0;
  for (int base = 0; base < words; base += kPermThreads) {
    const int i = base + tid;
    const u32 c = i < words ? __popc(bm[i]) : 0u;
    u32 incl = c;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const u32 y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[wid] = incl;
    __syncthreads();
    u32 before = carry, total = 0;
    for (int w = 0; w < kPermThreads / kWave; ++w) {
      const u32 t = wsum[w];
      if (w < wid) before += t;
      total += t;
    }
    if (i < words) pre[i] = before + incl - c;
    carry += total;
    __syncthreads();
  }
  if (tid == 0 && static_cast<int64_t>(carry) != cnt) atomicOr(err, 2);
  for (int64_t lo = 0; lo < cnt; lo += kPermStage) {
    const int64_t hi = lo + kPermStage < cnt ? lo + kPermStage : cnt;
    for (int64_t b = 0; b < cnt; b += static_cast<int64_t>(kPermThreads) * kPermUnroll) {
      u32 L[kPermUnroll];
#pragma unroll
      for (int k = 0; k < kPermUnroll; ++k) {
        const int64_t e = b + k * kPermThreads + tid;
        L[k] = e < cnt ? pos[e0 + e] : 0u;
      }
      double v[kPermUnroll];
#pragma unroll
      for (int k = 0; k < kPermUnroll; ++k) {
        const int64_t e = b + k * kPermThreads + tid;
        v[k] = e < cnt ? val_em[e0 + e] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < kPermUnroll; ++k) {
        const int64_t e = b + k * kPermThreads + tid;
        if (e >= cnt) continue;
        const u32 l = L[k] < nbits ? L[k] : nbits - 1u;
        const u32 w = l >> 5;
        const int64_t at = static_cast<int64_t>(pre[w]) + __popc(bm[w] & ((1u << (l & 31u)) - 1u));
        if (lo == 0) opos[e0 + e] = static_cast<u32>(o0 + at);
        if (at >= lo && at < hi) stage[at - lo] = v[k];
      }
    }
    __syncthreads();
    for (int64_t b = lo; b < hi; b += static_cast<int64_t>(kPermThreads) * kPermUnroll) {
      double v[kPermUnroll];
#pragma unroll
      for (int k = 0; k < kPermUnroll; ++k) {
        const int64_t i = b + k * kPermThreads + tid;
        v[k] = i < hi ? stage[i - lo] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < kPermUnroll; ++k) {
        const int64_t i = b + k * kPermThreads + tid;
        if (i < hi) {
          val[o0 + i] = v[k];
          row_out[o0 + i] = static_cast<u32>(row);
        }
      }
    }
    __syncthreads();
  }
}

__global__ void split_keys(const u64* __restrict__ key, const double* __restrict__ val, int64_t n,
                           u32* __restrict__ row, u32* __restrict__ col, double* __restrict__ v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  row[i] = static_cast<u32>(key[i] >> 32);
  col[i] = static_cast<u32>(key[i] & 0xFFFFFFFFull);
  v[i] = val[i];
}

__global__ void rank_keys(const u64* __restrict__ rows, const u64* __restrict__ det, int64_t n, HashView HP,
                          u64* __restrict__ key) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = (rows[i] << 32) | static_cast<u64>(hash_find(HP, det[2 * i], det[2 * i + 1]));
}

// ---- row-order H_VP: P ranks from the sorted ket keys
// Sorted keys in tiles of 1024, one wave per tile: heads (first of each run
// of equal keys) per tile.
constexpr int kRankTile = 1024;

__global__ void head_tiles(const u64* __restrict__ k, int64_t n, int64_t ntiles, int64_t* __restrict__ cnt) {
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (tile >= ntiles) return;
  const int lane = __lane_id();
  int c = 0;
  for (int j = 0; j < kRankTile / kWave; ++j) {
    const int64_t i = tile * kRankTile + j * kWave + lane;
    if (i < n) c += (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (lane == 0) cnt[tile] = c;
}

// Each sorted entry's ket rank (heads at or before it, minus one) goes to its
// output position as the column; each head writes its ket into P.
__global__ void rank_scatter(const u64* __restrict__ k, const u32* __restrict__ pos, int64_t n, int64_t ntiles,
                             const int64_t* __restrict__ toff, int n_orb, int na, int nb, u64 cb,
                             const u32* __restrict__ str_a, u64 n_str_a, const u32* __restrict__ str_b,
                             u32* __restrict__ col, u64* __restrict__ det_p) {
  __shared__ u32 bt_words[33 * 33];
  for (int i = threadIdx.x; i < 33 * 33; i += blockDim.x) bt_words[i] = (&kBinomDev.c[0][0])[i];
  const BinomTable& Bt = *reinterpret_cast<const BinomTable*>(bt_words);
  __syncthreads();
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (tile >= ntiles) return;
  const int lane = __lane_id();
  const u64 le = lanemask_lt() | (1ull << lane);
  const double inv_cb = cb ? 1.0 / static_cast<double>(cb) : 0.0;
  int64_t run = toff[tile];
  const int64_t t0 = tile * kRankTile;
  // key before the tile: the previous step's last key carries on from there
  u64 last = t0 > 0 ? k[t0 - 1] : ~0ull;
  constexpr int kSteps = kRankTile / kWave, kBatch = 4;
  for (int j0 = 0; j0 < kSteps; j0 += kBatch) {
    u64 key[kBatch];
    u32 ps[kBatch];
#pragma unroll
    for (int c = 0; c < kBatch; ++c) {
      const int64_t i = t0 + (j0 + c) * kWave + lane;
      key[c] = i < n ? k[i] : ~0ull;
      ps[c] = i < n ? pos[i] : 0u;
    }
#pragma unroll
    for (int c = 0; c < kBatch; ++c) {
      const int64_t i = t0 + (j0 + c) * kWave + lane;
      const bool in = i < n;
      u64 prev = __shfl_up(key[c], 1);
      if (lane == 0) prev = last;
      last = __shfl(key[c], kWave - 1);
      const bool head = in && (i == 0 || key[c] != prev);
      const u64 m = __ballot(head);
      const int64_t r = run + __popcll(m & le) - 1;
      if (in) col[ps[c]] = static_cast<u32>(r);
      if (head) {
        const u64 kk = key[c];
        u64 a, b;
        if (cb) {
          u64 ra = static_cast<u64>(static_cast<double>(kk) * inv_cb);
          while (ra * cb > kk) --ra;
          while ((ra + 1) * cb <= kk) ++ra;
          const u32 rb = static_cast<u32>(kk - ra * cb);
          a = (str_a && ra < n_str_a) ? str_a[ra] : colex_unrank(Bt, static_cast<u32>(ra), na, n_orb);
          b = (str_b && rb < cb) ? str_b[rb] : colex_unrank(Bt, rb, nb, n_orb);
        } else {
          a = kk >> n_orb;
          b = kk & low_mask(n_orb);
        }
        det_p[2 * r] = a;
        det_p[2 * r + 1] = b;
      }
      run += __popcll(m);
    }
  }
}

int key_bits(int64_t n_rows) {
  int b = 1;
  while ((1LL << b) < n_rows) ++b;
  return 32 + b;
}

// Sort (key, val) by key and split into COO arrays.
void finish_coo(DevBuf<u64>& key, DevBuf<double>& val, int64_t n, int end_bit, DevBuf<u32>& row, DevBuf<u32>& col,
                DevBuf<double>& v, hipStream_t s) {
  row.alloc(n);
  col.alloc(n);
  v.alloc(n);
  if (n == 0) return;
  DevBuf<u64> ks(n, s);
  DevBuf<double> vs(n, s);
  sort_pairs_u64_f64(key.p, val.p, ks.p, vs.p, static_cast<size_t>(n), end_bit, s);
  const unsigned g = static_cast<unsigned>((n + 255) / 256);
  split_keys<<<g, 256, 0, s>>>(ks.p, vs.p, n, row.p, col.p, v.p);
  DQ_HIP(hipGetLastError());
}

int64_t read_i64(const int64_t* d, hipStream_t s) {
  int64_t h = 0;
  DQ_HIP(hipMemcpyAsync(&h, d, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  return h;
}

// ------------------------------------------------------------------ energy
struct EnergySink {
  static constexpr bool kNeedKet = true;
  static constexpr bool kCountOnly = false;
  __device__ __forceinline__ void add_doubles(int) {}
  HashView H;
  const double* c;
  double ci, sthr, acc;
  __device__ __forceinline__ void diag(double d) {
    if (__lane_id() == 0 && fabs(d) > kMatThr) acc += d * ci;
  }
  __device__ __forceinline__ void add(bool keep, u64 ka, u64 kb, double v) {
    if (keep) {
      const long j = hash_find(H, ka, kb);
      if (j >= 0) acc += v * c[j];
    }
  }
  __device__ __forceinline__ void single(bool ok, u64 ka, u64 kb, double v) { add(ok && fabs(v) > sthr, ka, kb, v); }
  __device__ __forceinline__ void dbl(bool ok, u64 ka, u64 kb, double v) { add(ok && fabs(v) > kMatThr, ka, kb, v); }
};

__global__ __launch_bounds__(kWave* kWavesPerBlock) void energy_kernel(
    DevView V, const uint64_t* __restrict__ dets, const double* __restrict__ coeffs, int64_t n_rows, int n_orb,
    int use_hb, double sthr, const u32* __restrict__ hb_len, HashView H, double* __restrict__ sigma,
    int* __restrict__ err) {
  __shared__ RowLists lists[kWavesPerBlock];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wid;
  RowLists& L = lists[wid];
  const bool active = row < n_rows;
  u64 xa = 0, xb = 0;
  if (active) {
    xa = dets[2 * row];
    xb = dets[2 * row + 1];
  }
  build_lists(L, xa, xb, n_orb);
  __syncthreads();
  if (!active) return;
  if ((xa | xb) & ~low_mask(n_orb)) {
    if (lane == 0) {
      sigma[row] = 0.0;
      atomicOr(err, 1);
    }
    return;
  }
  const double ci = coeffs[row];
  if (ci * ci < 1e-12) {  // local_conn.cpp:190
    if (lane == 0) sigma[row] = 0.0;
    return;
  }
  EnergySink S{H, coeffs, ci, sthr, 0.0};
  const RowEnv R = make_env(V, xa, xb, n_orb, L);
  walk_row(R, S, use_hb ? 1 : 0, hb_len, 0.0);
  double a = S.acc;
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, kWave);
  if (lane == 0) sigma[row] = a;
}

// V-driven variant of energy_kernel (same row contract), one wave per row.
__global__ __launch_bounds__(kWave* kWavesPerBlock) void energy_str_kernel(
    DevView V, const uint64_t* __restrict__ dets, const double* __restrict__ coeffs, int64_t n_rows, int n_orb,
    double sthr, double dmin, StrIndex X, StrCoef C, double* __restrict__ sigma, int* __restrict__ err) {
  __shared__ StrWave lds[kWavesPerBlock];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wid;
  if (row >= n_rows) return;  // whole waves leave; the rest sync per wave only
  StrWave& W = lds[wid];
  const u64 xa = dets[2 * row], xb = dets[2 * row + 1];
  build_lists(W.L, xa, xb, n_orb);
  wave_sync();
  if ((xa | xb) & ~low_mask(n_orb)) {
    if (lane == 0) {
      sigma[row] = 0.0;
      atomicOr(err, 1);
    }
    return;
  }
  const double ci = coeffs[row];
  if (ci * ci < 1e-12) {  // local_conn.cpp:190
    if (lane == 0) sigma[row] = 0.0;
    return;
  }
  double a = str_energy_row(V, X, C, xa, xb, n_orb, ci, sthr, dmin, W);
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, kWave);
  if (lane == 0) sigma[row] = a;
}

// Fixed-order reduction of c_i * sigma_i: blocks of 256 by a tree, then one
// thread over the block sums.
__global__ void dot_blocks(const double* __restrict__ c, const double* __restrict__ s, int64_t n,
                           double* __restrict__ part) {
  __shared__ double buf[256];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  buf[threadIdx.x] = i < n ? c[i] * s[i] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) buf[threadIdx.x] += buf[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = buf[0];
}

__global__ void sum_serial(const double* __restrict__ part, int64_t m, double* __restrict__ out) {
  double t = 0.0;
  for (int64_t i = 0; i < m; ++i) t += part[i];
  *out = t;
}

// ------------------------------------------------------------------ SpGEMM (H_eff)
__global__ void make_keys(const u32* __restrict__ r, const u32* __restrict__ c, int64_t n, u64 nr, u64 nc,
                          int swap, u64* __restrict__ key, int* __restrict__ err) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (r[i] >= nr || c[i] >= nc) atomicOr(err, 1);
  key[i] = swap ? ((static_cast<u64>(c[i]) << 32) | r[i]) : ((static_cast<u64>(r[i]) << 32) | c[i]);
}

// flag = 1 when keys are not strictly increasing or a value is exactly 0.0,
// i.e. when the COO is not already canonical.
__global__ void check_canonical(const u64* __restrict__ key, const double* __restrict__ val, int64_t n,
                                int* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool bad = i < n && (val[i] == 0.0 || (i > 0 && key[i] <= key[i - 1]));
  if (__ballot(bad) != 0 && __lane_id() == 0) atomicOr(flag, 1);  // one atomic per wave
}

// (hi << 32 | lo) -> (lo << 32 | hi).
__global__ void swap_halves(const u64* __restrict__ in, int64_t n, u64* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (in[i] << 32) | (in[i] >> 32);
}

// Canonical merge of sorted keys: duplicates summed in order, exact zeros
// dropped (coo_to_csr / coo_to_csc, ham_utils.cpp:183-302).  keep[i]=1 marks
// the first entry of a key whose sum is nonzero; sums are written at heads.
__global__ void merge_heads(const u64* __restrict__ key, const double* __restrict__ val, int64_t n,
                            double* __restrict__ sum, int64_t* __restrict__ keep) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i > 0 && key[i] == key[i - 1]) {
    keep[i] = 0;
    return;
  }
  double acc = val[i];
  for (int64_t t = i + 1; t < n && key[t] == key[i]; ++t) acc += val[t];
  sum[i] = acc;
  keep[i] = acc != 0.0 ? 1 : 0;
}

__global__ void compact_kv(const u64* __restrict__ key, const double* __restrict__ sum,
                           const int64_t* __restrict__ keep, const int64_t* __restrict__ pos, int64_t n,
                           u64* __restrict__ okey, double* __restrict__ oval) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  okey[pos[i]] = key[i];
  oval[pos[i]] = sum[i];
}

// One column of H_VP for the row kernel's entry pass: start and length in the
// CSC (u32: the LDS path requires nnz < 2^32), the skip flag, dinv; 16 bytes,
// so an entry gathers its column with one load.
struct __align__(16) ColRec {
  u32 q0;
  u32 len_skip;  // length, skip flag in bit 31
  double dinv;
};
constexpr u32 kColSkip = 0x80000000u;

// dinv and the column skip rule (ham_eff.cpp:35-56, 95-108); rec, when
// given, also receives the column records.
__global__ void dinv_kernel(const double* __restrict__ hpp, int64_t np, double e_ref, double eps, int sigma,
                            const int64_t* __restrict__ cptr, const double* __restrict__ cval,
                            double* __restrict__ dinv, unsigned char* __restrict__ skip, ColRec* __restrict__ rec) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const double d = e_ref - hpp[j];
  const double di = sigma ? d / (d * d + eps * eps) : 1.0 / (d + copysign(eps, d));
  double s = 0.0;
  for (int64_t q = cptr[j]; q < cptr[j + 1]; ++q) s += cval[q] * cval[q];
  const bool sk = fabs(di) * s < 1e-12;
  dinv[j] = di;
  skip[j] = sk ? 1 : 0;
  if (rec) {
    ColRec r;
    r.q0 = static_cast<u32>(cptr[j]);
    r.len_skip = static_cast<u32>(cptr[j + 1] - cptr[j]) | (sk ? kColSkip : 0u);
    r.dinv = di;
    rec[j] = r;
  }
}

// Triple counts per H_VP entry p = (i, k), row-major: the entry contributes
// one triple per j of column k unless skip[k], a == 0 or dinv[k] == 0
// (ham_eff.cpp:173-205).  Emission in entry order keeps, for each (i, j),
// the reference's ascending-k accumulation order.
__global__ void triple_count(const u32* __restrict__ rcol, const double* __restrict__ rval,
                             const int64_t* __restrict__ cptr, const double* __restrict__ dinv,
                             const unsigned char* __restrict__ skip, int64_t m, int64_t* __restrict__ cnt) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= m) return;
  const u32 k = rcol[p];
  cnt[p] = (skip[k] || rval[p] == 0.0 || dinv[k] == 0.0) ? 0 : cptr[k + 1] - cptr[k];
}

// Triple emission balanced over a block: the output ranges of 256
// consecutive entries lie end to end, and each output slot finds its entry
// by binary search over the block's offsets in LDS, so a long column costs
// no more than many short ones and stores stay coalesced.
constexpr int kTripleBlock = 256;

__global__ __launch_bounds__(kTripleBlock) void triple_fill(
    const u32* __restrict__ rrow, const u32* __restrict__ rcol, const double* __restrict__ rval,
    const int64_t* __restrict__ cptr, const u32* __restrict__ crow, const double* __restrict__ cval,
    const double* __restrict__ dinv, int64_t m, const int64_t* __restrict__ off, u64* __restrict__ key,
    double* __restrict__ val) {
  __shared__ int64_t s_off[kTripleBlock + 1];
  __shared__ int64_t s_q[kTripleBlock];
  __shared__ u64 s_row[kTripleBlock];
  __shared__ double s_as[kTripleBlock];
  const int t = threadIdx.x;
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * kTripleBlock;
  const int64_t p = p0 + t;
  const int64_t o = off[p < m ? p : m];
  s_off[t] = o;
  if (t == 0) s_off[kTripleBlock] = off[p0 + kTripleBlock < m ? p0 + kTripleBlock : m];
  if (p < m) {
    const u32 k = rcol[p];
    s_q[t] = cptr[k] - o;
    s_row[t] = static_cast<u64>(rrow[p]) << 32;
    s_as[t] = rval[p] * dinv[k];
  }
  __syncthreads();
  const int64_t o1 = s_off[kTripleBlock];
  for (int64_t x = s_off[0] + t; x < o1; x += kTripleBlock) {
    int lo = 0, hi = kTripleBlock;  // s_off[lo] <= x < s_off[hi]
#pragma unroll
    for (int step = 0; step < 8; ++step) {
      const int mid = (lo + hi) >> 1;
      if (s_off[mid] <= x)
        lo = mid;
      else
        hi = mid;
    }
    const int64_t q = x + s_q[lo];
    key[x] = s_row[lo] | crow[q];
    val[x] = s_as[lo] * cval[q];
  }
}

// Run heads of a sorted key array: flag[i] = 1 where key[i] starts a run.
__global__ void head_flags(const u64* __restrict__ key, int64_t n, int64_t* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// hpos[r] = start of run r; hpos[runs] = n.
__global__ void head_scatter(const int64_t* __restrict__ flag, const int64_t* __restrict__ pos, int64_t n,
                             int64_t* __restrict__ hpos) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) hpos[pos[i]] = i;
  if (i == 0) hpos[pos[n]] = n;
}

// Sequential fold of each key's run, starting from 0.0 (ham_eff.cpp:161,214),
// one thread per run with its bounds known, so loads run ahead of the adds.
__global__ void fold_heads(const u64* __restrict__ key, const double* __restrict__ val,
                           const int64_t* __restrict__ hpos, int64_t runs, u64* __restrict__ okey,
                           double* __restrict__ oval) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= runs) return;
  const int64_t b = hpos[r], e = hpos[r + 1];
  double acc = 0.0;
#pragma unroll 8
  for (int64_t t = b; t < e; ++t) acc += val[t];
  okey[r] = key[b];
  oval[r] = acc;
}

// csr_add(VV, ADD, 1e-15) on a stable sort of [VV ; ADD]: a key present in
// both gives VV + ADD; |v| <= 1e-15 is dropped (ham_utils.cpp:330-442).
__global__ void add_runs(const u64* __restrict__ key, const double* __restrict__ val, int64_t n,
                         double* __restrict__ sum, int64_t* __restrict__ keep) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i > 0 && key[i] == key[i - 1]) {
    keep[i] = 0;
    return;
  }
  double v = val[i];
  if (i + 1 < n && key[i + 1] == key[i]) v = val[i] + val[i + 1];
  sum[i] = v;
  keep[i] = fabs(v) > kMatThr ? 1 : 0;
}

// ---------------------------------------------------------------- H_eff by row
// ADD[i, :] is accumulated one row of V per wavefront in an LDS window of
// FP64 accumulators over j.  The row's kept H_VP entries (steps) are taken in
// ascending k and each step's column is applied by the lanes at once; a
// column holds every j once, so lanes never collide, and each element
// receives its terms in the reference's order, 0.0 + t_k1 + t_k2 + ...
// (ham_eff.cpp:189-219).  H_VV is then added (csr_add, ham_utils.cpp:330-442)
// and the window is emitted in ascending j with |v| > 1e-15.
//
// Every kept column holds row i itself, so ADD[i, i] takes a term from every
// step: it is summed by its own sequential chain (eff_diag), and the row
// kernel visits only columns with two or more rows, skipping j = i.  A column
// longer than a wavefront is split into consecutive steps of at most 64 rows
// (distinct j, so their relative order is immaterial).

constexpr int kEffStep = 32;       // steps per half of a meta group
constexpr int kEffMaxCols = 8128;  // widest window; + 64 dummy slots = 64 KB of LDS
constexpr int kEffMaxPasses = 4;   // wider V goes through the sort path

__device__ __forceinline__ int lane_i32(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ double lane_f64(double x, int l) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const unsigned hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double(static_cast<long long>((static_cast<u64>(hi) << 32) | lo));
}

// acc += x as one LDS atomic without return: the adds of one wavefront reach
// LDS in issue order, so successive steps on one j are applied in sequence
// without waiting for a read.
__device__ __forceinline__ void lds_add(double* p, double x) {
  __hip_atomic_fetch_add(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Per H_VP entry p = (i, k), in row-major order: kept when skip[k] = 0,
// a != 0 and dinv[k] != 0 (ham_eff.cpp:191-201).  d = (a * dinv[k]) * a is
// its ADD[i, i] term, +0.0 when not kept (adding +0.0 to a sum that starts at
// +0.0 changes nothing); nsub = ceil(rows / 64) row-kernel steps when its
// column holds two or more rows, else 0.
__global__ void eff_entries(const u32* __restrict__ rcol, const double* __restrict__ rval,
                            const ColRec* __restrict__ crec, int64_t m, double* __restrict__ e_d,
                            double* __restrict__ e_as, u32* __restrict__ e_q0, u32* __restrict__ e_len,
                            int64_t* __restrict__ nsub) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= m) return;
  const ColRec c = crec[rcol[p]];
  const double a = rval[p], s = c.dinv;
  const u32 len = c.len_skip & ~kColSkip;
  const bool kept = !((c.len_skip & kColSkip) || a == 0.0 || s == 0.0);
  const double as = a * s;
  e_d[p] = kept ? as * a : 0.0;
  e_as[p] = as;
  e_q0[p] = c.q0;
  e_len[p] = len;
  nsub[p] = (kept && len > 1) ? (static_cast<int64_t>(len) + kWave - 1) / kWave : 0;
}

// Step records of the multi-row entries: start of the step's rows in the
// column, rows - 1, a * dinv[k] (ham_eff.cpp:203).
__global__ void eff_steps(const double* __restrict__ e_as, const u32* __restrict__ e_q0,
                          const u32* __restrict__ e_len, const int64_t* __restrict__ nsub,
                          const int64_t* __restrict__ mpos, int64_t m, u32* __restrict__ st_q,
                          u32* __restrict__ st_lm1, double* __restrict__ st_as) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= m) return;
  const int64_t n = nsub[p];
  if (n == 0) return;
  const int64_t o = mpos[p];
  const u32 q0 = e_q0[p];
  const int64_t len = e_len[p];
  const double as = e_as[p];
  for (int64_t t = 0; t < n; ++t) {
    const int64_t rest = len - t * kWave;
    st_q[o + t] = q0 + static_cast<u32>(t * kWave);
    st_lm1[o + t] = static_cast<u32>((rest < kWave ? rest : kWave) - 1);
    st_as[o + t] = as;
  }
}

__global__ void eff_row_ptr(const int64_t* __restrict__ rptr, const int64_t* __restrict__ mpos, int64_t nv,
                            int64_t* __restrict__ sptr) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i <= nv) sptr[i] = mpos[rptr[i]];
}

// ADD[i, i]: 0.0 + d_1 + d_2 + ... over the row's kept entries in order.
__global__ void eff_diag(const int64_t* __restrict__ rptr, const double* __restrict__ e_d, int64_t nv,
                         double* __restrict__ diag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int64_t e = rptr[i + 1];
  int64_t t = rptr[i];
  double acc = 0.0;
  // 64 loads in flight per wait: one thread per row is the whole parallelism.
  constexpr int kBatch = 64;
  double buf[kBatch];
  for (; t + kBatch <= e; t += kBatch) {
#pragma unroll
    for (int u = 0; u < kBatch; ++u) buf[u] = e_d[t + u];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) acc += buf[u];
  }
  for (; t < e; ++t) acc += e_d[t];
  diag[i] = acc;
}

// Output bound per row: min(n_V, nnz(H_VV row) + 1 + rows of its steps).
__global__ void eff_caps(const int64_t* __restrict__ sptr, const u32* __restrict__ st_lm1,
                         const int64_t* __restrict__ vvptr, int64_t nv, int64_t* __restrict__ cap) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (i >= nv) return;
  const int lane = __lane_id();
  int64_t t = 0;
  for (int64_t s = sptr[i] + lane; s < sptr[i + 1]; s += kWave) t += st_lm1[s] + 1;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d);
  if (lane == 0) {
    const int64_t c = t + 1 + vvptr[i + 1] - vvptr[i];
    cap[i] = c < nv ? c : nv;
  }
}

__global__ __launch_bounds__(kWave) void eff_rows(
    const int64_t* __restrict__ sptr, const u32* __restrict__ st_q, const u32* __restrict__ st_lm1,
    const double* __restrict__ st_as, const u32* __restrict__ crow, const double* __restrict__ cval,
    const double* __restrict__ diag, const int64_t* __restrict__ vvptr, const u64* __restrict__ vvkey,
    const double* __restrict__ vvval, int64_t nv, int win, const int64_t* __restrict__ soff,
    u32* __restrict__ scol, double* __restrict__ sval, int64_t* __restrict__ cnt) {
  extern __shared__ double acc[];  // win accumulators, then one dummy slot per lane
  const int lane = __lane_id();
  const int64_t i = blockIdx.x;
  const int64_t s0 = uniform_i64(sptr[i]);
  const int64_t S = uniform_i64(sptr[i + 1]) - s0;
  const int64_t v0 = uniform_i64(vvptr[i]), v1 = uniform_i64(vvptr[i + 1]);
  const int64_t out0 = uniform_i64(soff[i]);
  int64_t out = out0;
  const int64_t groups = (S + 2 * kEffStep - 1) / (2 * kEffStep);
  const u32 dummy = static_cast<u32>(win) + lane;

  // Meta group G: lane l holds step 64G + l; past the row's last step, an
  // empty step (lm1 = -1) at the last step's start.
  auto meta = [&](int64_t G, u32& q, int& lm1, double& as) {
    const int64_t st = G * (2 * kEffStep) + lane;
    const int64_t c = s0 + (st < S ? st : S - 1);
    q = st_q[c];
    lm1 = st < S ? static_cast<int>(st_lm1[c]) : -1;
    as = st_as[c];
  };
  // The rows of each step of half h; every lane loads (clamped to the step),
  // so the load count is fixed.
  auto cols = [&](u32 mq, int ml, int h, u32 (&J)[kEffStep], double (&X)[kEffStep]) {
#pragma unroll
    for (int u = 0; u < kEffStep; ++u) {
      const u32 qb = static_cast<u32>(lane_i32(static_cast<int>(mq), h * kEffStep + u));
      const int lm1 = lane_i32(ml, h * kEffStep + u);
      const u32 q = qb + static_cast<u32>(lane < lm1 ? lane : (lm1 > 0 ? lm1 : 0));
      J[u] = crow[q];
      X[u] = cval[q];
    }
  };

  for (int64_t w0 = 0; w0 < nv; w0 += win) {
    const u32 wn = static_cast<u32>(nv - w0 < win ? nv - w0 : win);
    const u32 wb = static_cast<u32>(w0);
    const u32 iw = static_cast<u32>(i) - wb;  // row i's own slot (wraps when outside)
    for (u32 x = lane; x < wn; x += kWave) acc[x] = 0.0;
    auto apply = [&](int ml, double ma, int h, const u32 (&J)[kEffStep], const double (&X)[kEffStep]) {
#pragma unroll
      for (int u = 0; u < kEffStep; ++u) {
        const int lm1 = lane_i32(ml, h * kEffStep + u);
        const double as = lane_f64(ma, h * kEffStep + u);
        const u32 jj = J[u] - wb;
        const bool ok = lane <= lm1 && jj < wn && jj != iw;
        lds_add(&acc[ok ? jj : dummy], as * X[u]);
      }
    };
    if (groups > 0) {
      u32 qc, qn, qnn;
      int lc, ln, lnn;
      double ac, an, ann;
      u32 J0[kEffStep], J1[kEffStep];
      double X0[kEffStep], X1[kEffStep];
      meta(0, qc, lc, ac);
      meta(1, qn, ln, an);
      cols(qc, lc, 0, J0, X0);
      for (int64_t G = 0; G < groups; ++G) {
        meta(G + 2, qnn, lnn, ann);
        cols(qc, lc, 1, J1, X1);
        apply(lc, ac, 0, J0, X0);
        cols(qn, ln, 0, J0, X0);
        apply(lc, ac, 1, J1, X1);
        qc = qn, lc = ln, ac = an;
        qn = qnn, ln = lnn, an = ann;
      }
    }
    if (lane == 0 && iw < wn) acc[iw] = diag[i];
    // csr_add(VV, ADD): VV + ADD
    for (int64_t t = v0 + lane; t < v1; t += kWave) {
      const u32 jj = static_cast<u32>(vvkey[t] & 0xFFFFFFFFull) - wb;
      if (jj < wn) acc[jj] = vvval[t] + acc[jj];
    }
    for (u32 x0 = 0; x0 < wn; x0 += kWave) {
      const u32 x = x0 + lane;
      const double v = x < wn ? acc[x] : 0.0;
      const bool keep = fabs(v) > kMatThr;
      const u64 m = __ballot(keep);
      if (keep) {
        const int64_t o = out + __popcll(m & lanemask_lt());
        scol[o] = wb + x;
        sval[o] = v;
      }
      out += __popcll(m);
    }
  }
  if (lane == 0) cnt[i] = out - out0;
}

// Rows of the staging buffer -> contiguous COO.
__global__ void eff_gather(const int64_t* __restrict__ soff, const int64_t* __restrict__ ooff,
                           const u32* __restrict__ scol, const double* __restrict__ sval, int64_t nv,
                           u32* __restrict__ row, u32* __restrict__ col, double* __restrict__ val) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (i >= nv) return;
  const int lane = __lane_id();
  const int64_t a = soff[i], o = ooff[i], n = ooff[i + 1] - o;
  for (int64_t t = lane; t < n; t += kWave) {
    row[o + t] = static_cast<u32>(i);
    col[o + t] = scol[a + t];
    val[o + t] = sval[a + t];
  }
}

unsigned grid_of(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

enum Reduce { kMerge, kFold, kAdd };

// Sorted keys/values -> one result per key (kMerge: merge_heads, kFold:
// fold_heads, kAdd: add_runs), compacted; returns the count.
int64_t reduce_sorted(Reduce how, const u64* ks, const double* vs, int64_t n, DevBuf<u64>& okey,
                      DevBuf<double>& oval, hipStream_t s) {
  if (n == 0) {
    okey.alloc(1);
    oval.alloc(1);
    return 0;
  }
  if (how == kFold) {
    DevBuf<int64_t> flag(n, s), pos(n + 1, s);
    head_flags<<<grid_of(n), 256, 0, s>>>(ks, n, flag.p);
    exclusive_scan_i64(flag.p, n, pos.p, s);
    const int64_t runs = read_i64(pos.p + n, s);
    DevBuf<int64_t> hpos(runs + 1, s);
    head_scatter<<<grid_of(n), 256, 0, s>>>(flag.p, pos.p, n, hpos.p);
    okey.alloc(runs);
    oval.alloc(runs);
    fold_heads<<<grid_of(runs), 256, 0, s>>>(ks, vs, hpos.p, runs, okey.p, oval.p);
    DQ_HIP(hipGetLastError());
    return runs;
  }
  okey.alloc(n);
  oval.alloc(n);
  DevBuf<double> sum(n, s);
  DevBuf<int64_t> keep(n, s), pos(n + 1, s);
  if (how == kMerge)
    merge_heads<<<grid_of(n), 256, 0, s>>>(ks, vs, n, sum.p, keep.p);
  else
    add_runs<<<grid_of(n), 256, 0, s>>>(ks, vs, n, sum.p, keep.p);
  exclusive_scan_i64(keep.p, n, pos.p, s);
  compact_kv<<<grid_of(n), 256, 0, s>>>(ks, sum.p, keep.p, pos.p, n, okey.p, oval.p);
  DQ_HIP(hipGetLastError());
  return read_i64(pos.p + n, s);
}

__global__ void split_kv(const u64* __restrict__ key, int64_t n, u32* __restrict__ hi, u32* __restrict__ lo) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  hi[i] = static_cast<u32>(key[i] >> 32);
  lo[i] = static_cast<u32>(key[i] & 0xFFFFFFFFull);
}

// (major << 32 | minor) sorted keys -> pointer array over `nmaj` majors:
// slot i in [0, n] writes ptr[m] = i for every major m in (major(i-1),
// major(i)], with major(-1) = -1 and major(n) = nmaj, which covers
// [0, nmaj] once.
__global__ void ptr_from_sorted(const u64* __restrict__ key, int64_t n, int64_t nmaj, int64_t* __restrict__ ptr) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int64_t cur = i < n ? static_cast<int64_t>(key[i] >> 32) : nmaj;
  const int64_t prev = i > 0 ? static_cast<int64_t>(key[i - 1] >> 32) : -1;
  for (int64_t mj = prev + 1; mj <= cur; ++mj) ptr[mj] = i;
}

void build_ptr(const u64* key, int64_t n, int64_t nmaj, DevBuf<int64_t>& ptr, hipStream_t s) {
  ptr.alloc(nmaj + 1);
  ptr_from_sorted<<<grid_of(n + 1), 256, 0, s>>>(key, n, nmaj, ptr.p);
  DQ_HIP(hipGetLastError());
}

}  // namespace

void energy_reduce(const double* c, const double* sigma, int64_t n, double* part, double* out, hipStream_t s) {
  const int64_t nb = (n + 255) / 256;
  if (nb > 0) dot_blocks<<<static_cast<unsigned>(nb), 256, 0, s>>>(c, sigma, n, part);
  sum_serial<<<1, 1, 0, s>>>(part, nb, out);
  DQ_HIP(hipGetLastError());
}

}  // namespace dq

// ----------------------------------------------------------------------------- C-ABI
using dq::DevBuf;
using dq::u32;
using dq::u64;

namespace {

void blocks_from_keys(dq_blocks& B, DevBuf<u64>& key, DevBuf<double>& val, int64_t n, int end_bit, bool vp,
                      hipStream_t s) {
  if (vp) {
    dq::finish_coo(key, val, n, end_bit, B.vp_row, B.vp_col, B.vp_val, s);
    B.nnz_vp = n;
  } else {
    dq::finish_coo(key, val, n, end_bit, B.vv_row, B.vv_col, B.vv_val, s);
    B.nnz_vv = n;
  }
}

}  // namespace

extern "C" {

int dq_ham_blocks(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, const uint64_t* d_ext, int64_t n_ext,
                  int n_orb, int mode, int use_hb, double eps1, const double* d_psi, void* stream,
                  dq_blocks** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx && out, DQ_EINVAL, "null argument");
    DQ_REQUIRE(n >= 0 && (n == 0 || d_dets), DQ_EINVAL, "bad dets");
    DQ_REQUIRE(n_orb >= 1 && n_orb <= ctx->n_orb, DQ_EINVAL, "n_orb must be in [1, n_orb of the integral context]");
    DQ_REQUIRE(mode >= 0 && mode <= 2, DQ_EINVAL, "mode must be 0 (known sets), 1 (static HB) or 2 (dynamic)");
    auto* c = const_cast<dq_ctx*>(ctx);
    if (mode == dq::kStatic && use_hb) DQ_REQUIRE(c->hb, DQ_EINVAL, "Heat-Bath table not initialized");
    if (mode == dq::kDynamic) {
      DQ_REQUIRE(c->hb, DQ_EINVAL, "get_ham_conn_amp: Heat-Bath table is required");
      DQ_REQUIRE(n == 0 || d_psi, DQ_EINVAL, "get_ham_conn_amp: psi_V size must match |V|");
    }
    hipStream_t s = dq::as_stream(stream);
    auto B = std::make_unique<dq_blocks>();
    if (n == 0) {
      *out = B.release();
      return;
    }
    dq::DetHash HV, HE;
    if (HV.build(d_dets, n, s)) throw dq::Error(DQ_EINVAL, "DetMap::from_ordered: duplicate determinant found");
    DevBuf<dq::u32> vf((size_t(1) << dq::kProjFiltLg) / 32, s);
    dq::filter_build(d_dets, n, dq::kProjFiltLg, vf.p, s);
    const bool has_ext = mode == dq::kKnown && d_ext && n_ext > 0;
    if (has_ext && HE.build(d_ext, n_ext, s))
      throw dq::Error(DQ_EINVAL, "DetMap::from_ordered: duplicate determinant found");
    int dbl_mode = 0;
    const dq::u32* len = nullptr;
    if (mode == dq::kStatic && use_hb) {
      dbl_mode = 1;
      len = c->hb_lengths(eps1, s);
    } else if (mode == dq::kDynamic) {
      dbl_mode = 2;
    }
    const auto V = c->view();
    // Row order (kets in static / dynamic modes, n_orb <= 32): H_VP reaches
    // its (row, column) order without a sort by row, and one sort of the ket
    // keys ranks P.
    dq::ProjOrd O{nullptr, nullptr, 0};
    bool ord = mode != dq::kKnown && n_orb <= 32;
    int kbits = 2 * n_orb, na = 0, nb = 0, words = 0;
    if (ord) {
      int mm[4] = {64, -1, 64, -1};
      DevBuf<int> dmm(4, s);
      DQ_HIP(hipMemcpyAsync(dmm.p, mm, sizeof(mm), hipMemcpyHostToDevice, s));
      dq::pop_range<<<dq::grid_of(n), 256, 0, s>>>(d_dets, n, dmm.p);
      DQ_HIP(hipGetLastError());
      DQ_HIP(hipMemcpyAsync(mm, dmm.p, sizeof(mm), hipMemcpyDeviceToHost, s));
      DQ_HIP(hipStreamSynchronize(s));
      u64 umax = 0;  // (rows with orbitals at or above n_orb are rejected by the count pass)
      for (int a = std::min(mm[0], n_orb); a <= std::min(mm[1], n_orb); ++a)
        for (int b = std::min(mm[2], n_orb); b <= std::min(mm[3], n_orb); ++b)
          umax = std::max(umax, dq::row_universe(n_orb, a, b));
      words = static_cast<int>((umax + 31) / 32);
      if (mm[0] == mm[1] && mm[2] == mm[3]) {  // one sector: colex keys
        na = mm[0];
        nb = mm[2];
        O.cb = dq::binom_host(n_orb, nb);
        const u64 keys = dq::binom_host(n_orb, na) * O.cb;  // colex keys 0 .. keys - 1
        kbits = 1;
        while (kbits < 63 && (1ull << kbits) <= keys) ++kbits;  // all-ones of kbits bits: above every key
      }
    }
    DevBuf<int64_t> cnt_vv(n, s), cnt_vp(n, s), off_vv(n + 1, s), off_vp(n + 1, s);
    const unsigned g = static_cast<unsigned>((n + dq::kWavesPerBlock - 1) / dq::kWavesPerBlock);
    const unsigned tb = dq::kWave * dq::kWavesPerBlock;
    const int bits = dq::key_bits(n);
    if (ord) {
      // each row's kept connections, H_VV and H_VP together: in the static
      // mode the local-connection count with singles kept above 1e-15 (the
      // projection's own screening), otherwise the projection's count pass
      DevBuf<int64_t> rc(n, s), ub(n + 1, s);
      if (mode == dq::kStatic) {
        dq::launch_local_conn(*c, V, d_dets, n, n_orb, use_hb != 0, eps1, len, rc.p, nullptr, nullptr, nullptr,
                              false, s, dq::kMatThr);
      } else {
        dq::proj_kernel<false><<<g, tb, 0, s>>>(V, d_dets, n, n_orb, mode, dbl_mode, eps1, d_psi, len, HV.view(),
                                                HE.view(), vf.p, cnt_vv.p, cnt_vp.p, nullptr, nullptr, nullptr, nullptr,
                                                nullptr, nullptr, nullptr, O, c->d_err.p);
        DQ_HIP(hipGetLastError());
        dq::add_i64<<<dq::grid_of(n), 256, 0, s>>>(cnt_vv.p, cnt_vp.p, n, rc.p);
      }
      DQ_HIP(hipGetLastError());
      dq::exclusive_scan_i64(rc.p, n, ub.p, s);
      c->check_rows(s);
      const int64_t U = dq::read_i64(ub.p + n, s);
      ord = U < (int64_t(1) << 32);  // positions are u32 in the ranking sort
      if (ord) {
        const size_t U1 = U > 0 ? static_cast<size_t>(U) : 1;
        dq::PoolBuf<u64> ket(U1, s);
        dq::PoolBuf<u32> lpos(U1, s);
        dq::PoolBuf<double> val_em(U1, s);
        O.ket = ket.p;
        O.pos = lpos.p;
        O.ub = ub.p;
        dq::proj_kernel<true, true><<<g, tb, dq::kOrdLds, s>>>(
            V, d_dets, n, n_orb, mode, dbl_mode, eps1, d_psi, len, HV.view(), HE.view(), vf.p, cnt_vv.p, cnt_vp.p,
            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, val_em.p, O, c->d_err.p);
        DQ_HIP(hipGetLastError());
        dq::exclusive_scan_i64(cnt_vv.p, n, off_vv.p, s);
        dq::exclusive_scan_i64(cnt_vp.p, n, off_vp.p, s);
        c->check_rows(s);
        const int64_t nvv = dq::read_i64(off_vv.p + n, s), nvp = dq::read_i64(off_vp.p + n, s);
        DevBuf<u64> vv_key(nvv > 0 ? nvv : 1, s);
        DevBuf<double> vv_val(nvv > 0 ? nvv : 1, s);
        const u64 sentinel = dq::low_mask(kbits);
        dq::vv_gather<<<static_cast<unsigned>((n + 3) / 4), 256, 0, s>>>(ub.p, cnt_vv.p, off_vv.p, n, ket.p,
                                                                        val_em.p, vv_key.p, vv_val.p, sentinel);
        DQ_HIP(hipGetLastError());
        blocks_from_keys(*B, vv_key, vv_val, nvv, bits, false, s);
        const size_t N = static_cast<size_t>(nvp);
        B->nnz_vp = nvp;
        if (nvp > 0) {
          B->vp_row.alloc(N);
          B->vp_val.alloc(N);
          dq::PoolBuf<u32> opos(U1, s);
          const size_t plds = dq::perm_lds_bytes(words);
          DQ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&dq::ord_permute),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(plds)));
          dq::ord_permute<<<static_cast<unsigned>(n), dq::kPermThreads, plds, s>>>(
              ub.p, cnt_vp.p, off_vp.p, words, lpos.p, val_em.p, opos.p, B->vp_val.p, B->vp_row.p, c->d_err.p);
          DQ_HIP(hipGetLastError());
          c->check_rows(s);
          val_em.release();
          lpos.release();
          // the sort runs over the range array: H_VV slots hold the sentinel
          const size_t UN = static_cast<size_t>(U);
          dq::PoolBuf<u64> ket_s(UN, s);
          dq::PoolBuf<u32> pos_s(UN, s);
          size_t tmp = 0;
          DQ_HIP(rocprim::radix_sort_pairs(nullptr, tmp, ket.p, ket_s.p, opos.p, pos_s.p, UN, 0u,
                                                              static_cast<unsigned>(kbits), s));
          {
            dq::PoolBuf<unsigned char> t(tmp + 1, s);
            DQ_HIP(rocprim::radix_sort_pairs(t.p, tmp, ket.p, ket_s.p, opos.p, pos_s.p, UN, 0u,
                                                                static_cast<unsigned>(kbits), s));
          }
          ket.release();
          opos.release();
          const int64_t ntiles = (nvp + dq::kRankTile - 1) / dq::kRankTile;
          dq::PoolBuf<int64_t> tcnt(ntiles, s), toff(ntiles + 1, s);
          const unsigned tg = static_cast<unsigned>((ntiles + 3) / 4);
          dq::head_tiles<<<tg, 256, 0, s>>>(ket_s.p, nvp, ntiles, tcnt.p);
          DQ_HIP(hipGetLastError());
          dq::exclusive_scan_i64(tcnt.p, ntiles, toff.p, s);
          const int64_t np = dq::read_i64(toff.p + ntiles, s);
          B->det_p.alloc(2 * np);
          B->n_p = np;
          B->vp_col.alloc(N);
          // decode tables of the colex ranks when they are small
          const u64 ca = O.cb ? dq::binom_host(n_orb, na) : 0, cbb = O.cb;
          const bool tables = O.cb && ca + cbb <= (u64(1) << 25);
          dq::PoolBuf<u32> sa(tables ? ca : 1, s), sb(tables ? cbb : 1, s);
          if (tables) {
            dq::colex_strings<<<static_cast<unsigned>((ca + 255) / 256), 256, 0, s>>>(sa.p, static_cast<u32>(ca),
                                                                                      na, n_orb);
            dq::colex_strings<<<static_cast<unsigned>((cbb + 255) / 256), 256, 0, s>>>(sb.p, static_cast<u32>(cbb),
                                                                                       nb, n_orb);
            DQ_HIP(hipGetLastError());
          }
          dq::rank_scatter<<<tg, 256, 0, s>>>(ket_s.p, pos_s.p, nvp, ntiles, toff.p, n_orb, na, nb, O.cb,
                                              tables ? sa.p : nullptr, ca, tables ? sb.p : nullptr, B->vp_col.p,
                                              B->det_p.p);
          DQ_HIP(hipGetLastError());
        }
        DQ_HIP(hipStreamSynchronize(s));
        *out = B.release();
        return;
      }
    }
    dq::proj_kernel<false><<<g, tb, 0, s>>>(V, d_dets, n, n_orb, mode, dbl_mode, eps1, d_psi, len, HV.view(),
                                            HE.view(), vf.p, cnt_vv.p, cnt_vp.p, nullptr, nullptr, nullptr, nullptr,
                                            nullptr, nullptr, nullptr, O, c->d_err.p);
    DQ_HIP(hipGetLastError());
    dq::exclusive_scan_i64(cnt_vv.p, n, off_vv.p, s);
    dq::exclusive_scan_i64(cnt_vp.p, n, off_vp.p, s);
    c->check_rows(s);
    const int64_t nvv = dq::read_i64(off_vv.p + n, s), nvp = dq::read_i64(off_vp.p + n, s);
    DevBuf<u64> vv_key(nvv > 0 ? nvv : 1, s), vp_key(nvp > 0 ? nvp : 1, s);
    DevBuf<double> vv_val(nvv > 0 ? nvv : 1, s), vp_val(nvp > 0 ? nvp : 1, s);
    DevBuf<u64> vp_det(mode == dq::kKnown ? 1 : 2 * (nvp > 0 ? nvp : 1), s);
    dq::proj_kernel<true><<<g, tb, 0, s>>>(V, d_dets, n, n_orb, mode, dbl_mode, eps1, d_psi, len, HV.view(),
                                           HE.view(), vf.p, nullptr, nullptr, off_vv.p, off_vp.p, vv_key.p, vv_val.p,
                                           vp_key.p, vp_det.p, vp_val.p, O, c->d_err.p);
    DQ_HIP(hipGetLastError());
    blocks_from_keys(*B, vv_key, vv_val, nvv, bits, false, s);
    if (mode == dq::kKnown) {
      if (has_ext) blocks_from_keys(*B, vp_key, vp_val, nvp, bits, true, s);
      if (has_ext) {
        B->det_p.alloc(2 * n_ext);
        DQ_HIP(hipMemcpyAsync(B->det_p.p, d_ext, 2 * n_ext * sizeof(u64), hipMemcpyDeviceToDevice, s));
        B->n_p = n_ext;
      }
    } else {
      // P = sorted unique external keys; columns are their ranks.
      B->det_p.alloc(2 * (nvp > 0 ? nvp : 1));
      const int64_t np = static_cast<int64_t>(dq::sort_unique_dets(vp_det.p, static_cast<size_t>(nvp), B->det_p.p, s));
      B->n_p = np;
      if (np > 0) {
        dq::DetHash HP;
        HP.build(B->det_p.p, np, s);
        dq::rank_keys<<<dq::grid_of(nvp), 256, 0, s>>>(vp_key.p, vp_det.p, nvp, HP.view(), vp_key.p);
        DQ_HIP(hipGetLastError());
      }
      blocks_from_keys(*B, vp_key, vp_val, nvp, bits, true, s);
    }
    DQ_HIP(hipStreamSynchronize(s));
    *out = B.release();
  });
}

int dq_blocks_sizes(const dq_blocks* b, int64_t* nnz_vv, int64_t* nnz_vp, int64_t* n_p) {
  return dq::guarded([&] {
    DQ_REQUIRE(b, DQ_EINVAL, "null blocks");
    if (nnz_vv) *nnz_vv = b->nnz_vv;
    if (nnz_vp) *nnz_vp = b->nnz_vp;
    if (n_p) *n_p = b->n_p;
  });
}

int dq_blocks_copy(const dq_blocks* b, uint32_t* vv_row, uint32_t* vv_col, double* vv_val, uint32_t* vp_row,
                   uint32_t* vp_col, double* vp_val, uint64_t* det_p, int dst_on_device, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(b, DQ_EINVAL, "null blocks");
    hipStream_t s = dq::as_stream(stream);
    const hipMemcpyKind k = dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      if (dst && bytes) DQ_HIP(hipMemcpyAsync(dst, src, bytes, k, s));
    };
    cp(vv_row, b->vv_row.p, b->nnz_vv * sizeof(u32));
    cp(vv_col, b->vv_col.p, b->nnz_vv * sizeof(u32));
    cp(vv_val, b->vv_val.p, b->nnz_vv * sizeof(double));
    cp(vp_row, b->vp_row.p, b->nnz_vp * sizeof(u32));
    cp(vp_col, b->vp_col.p, b->nnz_vp * sizeof(u32));
    cp(vp_val, b->vp_val.p, b->nnz_vp * sizeof(double));
    cp(det_p, b->det_p.p, b->n_p * 2 * sizeof(u64));
    DQ_HIP(hipStreamSynchronize(s));
  });
}

int dq_blocks_free(dq_blocks* b) {
  return dq::guarded([&] { delete b; });
}

int dq_variational_energy(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs, int64_t n, int n_orb,
                          int use_hb, double eps1, double* d_sigma, double* h_energy, void* stream) {
  return dq_variational_energy_method(ctx, d_dets, d_coeffs, n, n_orb, use_hb, eps1, DQ_ELOC_STRINGS, d_sigma,
                                      h_energy, stream);
}

int dq_variational_energy_method(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs, int64_t n,
                                 int n_orb, int use_hb, double eps1, int method, double* d_sigma, double* h_energy,
                                 void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(method == DQ_ELOC_STRINGS || method == DQ_ELOC_WALK, DQ_EINVAL, "unknown energy method");
    DQ_REQUIRE(ctx, DQ_EINVAL, "null context");
    DQ_REQUIRE(n >= 0 && (n == 0 || (d_dets && d_coeffs)), DQ_EINVAL, "bad arguments");
    DQ_REQUIRE(n_orb >= 1 && n_orb <= ctx->n_orb, DQ_EINVAL, "n_orb must be in [1, n_orb of the integral context]");
    auto* c = const_cast<dq_ctx*>(ctx);
    if (use_hb)
      DQ_REQUIRE(c->hb, DQ_EINVAL, "compute_variational_energy: Heat-Bath mode requires non-null hb_table");
    hipStream_t s = dq::as_stream(stream);
    if (n == 0) {
      if (h_energy) *h_energy = 0.0;
      return;
    }
    dq::DetHash H;
    H.build(d_dets, n, s);  // tolerant: first occurrence wins (local_conn.cpp:171-172)
    // the walk reads coefficients through the DetMap: number the first
    // occurrences consecutively, as the reference's tolerant DetMap does
    if (method == DQ_ELOC_WALK) dq::hash_to_rank(d_dets, n, H.view(), H.vals.p, s);
    const dq::u32* len = use_hb ? c->hb_lengths(eps1, s) : nullptr;
    const double sthr = use_hb ? (eps1 > dq::kMatThr ? eps1 : dq::kMatThr) : dq::kMatThr;
    DevBuf<double> sig_own(0, s);
    double* sig = d_sigma;
    if (!sig) {
      sig_own.alloc(n);
      sig = sig_own.p;
    }
    const unsigned g = static_cast<unsigned>((n + dq::kWavesPerBlock - 1) / dq::kWavesPerBlock);
    dq::StrIndexBuf X;
    DevBuf<double> ca(0, s), cb(0, s);
    if (method == DQ_ELOC_STRINGS) {
      X.build(d_dets, n, H.view(), s);
      ca.alloc(n);
      cb.alloc(n);
      dq::str_gather_coef(X.view, d_coeffs, ca.p, cb.p, s);
      double st, dmin;
      dq::eloc_thresholds(*c, use_hb, eps1, &st, &dmin);
      dq::energy_str_kernel<<<g, dq::kWave * dq::kWavesPerBlock, 0, s>>>(
          c->view(), d_dets, d_coeffs, n, n_orb, st, dmin, X.view, dq::StrCoef{ca.p, cb.p}, sig, c->d_err.p);
    } else {
      dq::energy_kernel<<<g, dq::kWave * dq::kWavesPerBlock, 0, s>>>(c->view(), d_dets, d_coeffs, n, n_orb,
                                                                      use_hb ? 1 : 0, sthr, len, H.view(), sig,
                                                                      c->d_err.p);
    }
    DQ_HIP(hipGetLastError());
    if (h_energy) {
      const int64_t nb = (n + 255) / 256;
      DevBuf<double> part(nb, s), e(1, s);
      dq::dot_blocks<<<static_cast<unsigned>(nb), 256, 0, s>>>(d_coeffs, sig, n, part.p);
      dq::sum_serial<<<1, 1, 0, s>>>(part.p, nb, e.p);
      DQ_HIP(hipGetLastError());
      DQ_HIP(hipMemcpyAsync(h_energy, e.p, sizeof(double), hipMemcpyDeviceToHost, s));
      c->check_rows(s);
    } else {
      DQ_HIP(hipStreamSynchronize(s));  // the DetMap and string index are freed on return
    }
  });
}

int dq_ham_eff(const uint32_t* d_vv_row, const uint32_t* d_vv_col, const double* d_vv_val, int64_t nnz_vv,
               const uint32_t* d_vp_row, const uint32_t* d_vp_col, const double* d_vp_val, int64_t nnz_vp,
               int64_t n_v, const double* d_h_pp, int64_t n_p, double e_ref, int reg_sigma, double epsilon,
               void* stream, dq_blocks** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(out, DQ_EINVAL, "null output");
    DQ_REQUIRE(nnz_vv >= 0 && nnz_vp >= 0 && n_v >= 0 && n_p >= 0, DQ_EINVAL, "negative size");
    DQ_REQUIRE(n_v < (1LL << 32) && n_p < (1LL << 32), DQ_EINVAL, "dimension exceeds u32 indices");
    hipStream_t s = dq::as_stream(stream);
    auto B = std::make_unique<dq_blocks>();
    DevBuf<int> err(1, s);
    DQ_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    const int vbits = dq::key_bits(n_v);
    // H_VV canonical (coo_to_csr: duplicates summed, zeros dropped)
    DevBuf<u64> vk(nnz_vv > 0 ? nnz_vv : 1, s);
    DevBuf<double> vv(nnz_vv > 0 ? nnz_vv : 1, s);
    if (nnz_vv > 0) {
      dq::make_keys<<<dq::grid_of(nnz_vv), 256, 0, s>>>(d_vv_row, d_vv_col, nnz_vv, n_v, n_v, 0, vk.p, err.p);
      DQ_HIP(hipMemcpyAsync(vv.p, d_vv_val, nnz_vv * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    DevBuf<u64> vk_s(nnz_vv > 0 ? nnz_vv : 1, s), vvk(0, s);
    DevBuf<double> vv_s(nnz_vv > 0 ? nnz_vv : 1, s), vvv(0, s);
    dq::sort_pairs_u64_f64(vk.p, vv.p, vk_s.p, vv_s.p, static_cast<size_t>(nnz_vv), vbits, s);
    const int64_t m_vv = dq::reduce_sorted(dq::kMerge, vk_s.p, vv_s.p, nnz_vv, vvk, vvv, s);
    int herr = 0;
    if (n_p == 0) {  // ham_eff.cpp:71-77
      DQ_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
      DQ_HIP(hipStreamSynchronize(s));
      DQ_REQUIRE(!herr, DQ_ERANGE, "coo_to_csr: row index out of bounds");
      blocks_from_keys(*B, vvk, vvv, m_vv, vbits, false, s);
      DQ_HIP(hipStreamSynchronize(s));
      *out = B.release();
      return;
    }
    DQ_REQUIRE(d_h_pp, DQ_EINVAL, "null h_pp_diag");
    // H_VP as CSR (row-major keys, canonical) and CSC (column-major keys)
    DevBuf<u64> pk(nnz_vp > 0 ? nnz_vp : 1, s), rk(0, s);
    DevBuf<double> pv(nnz_vp > 0 ? nnz_vp : 1, s), rv(0, s);
    DevBuf<int> noncanon(1, s);
    DQ_HIP(hipMemsetAsync(noncanon.p, 0, sizeof(int), s));
    if (nnz_vp > 0) {
      dq::make_keys<<<dq::grid_of(nnz_vp), 256, 0, s>>>(d_vp_row, d_vp_col, nnz_vp, n_v, n_p, 0, pk.p, err.p);
      DQ_HIP(hipMemcpyAsync(pv.p, d_vp_val, nnz_vp * sizeof(double), hipMemcpyDeviceToDevice, s));
      dq::check_canonical<<<dq::grid_of(nnz_vp), 256, 0, s>>>(pk.p, pv.p, nnz_vp, noncanon.p);
    }
    int hnc = 0;
    DQ_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipMemcpyAsync(&hnc, noncanon.p, sizeof(int), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    DQ_REQUIRE(!herr, DQ_ERANGE, "coo_to_csr: index out of bounds");
    int64_t mr = nnz_vp;
    if (hnc) {  // coo_to_csr: sort, sum duplicates, drop zeros
      DevBuf<u64> rk_s(nnz_vp, s);
      DevBuf<double> rv_s(nnz_vp, s);
      dq::sort_pairs_u64_f64(pk.p, pv.p, rk_s.p, rv_s.p, static_cast<size_t>(nnz_vp), vbits, s);
      mr = dq::reduce_sorted(dq::kMerge, rk_s.p, rv_s.p, nnz_vp, rk, rv, s);
    } else {  // already canonical (get_ham_conn output is)
      rk = std::move(pk);
      rv = std::move(pv);
    }
    // The canonical entries, stably sorted by column alone, are the CSC in
    // (column, row) order (coo_to_csc).
    const int64_t mc = mr;
    DevBuf<u64> ck_in(mc > 0 ? mc : 1, s), ck(mc > 0 ? mc : 1, s);
    DevBuf<double> cv(mc > 0 ? mc : 1, s);
    if (mc > 0) dq::swap_halves<<<dq::grid_of(mc), 256, 0, s>>>(rk.p, mc, ck_in.p);
    dq::sort_pairs_u64_f64_bits(ck_in.p, rv.p, ck.p, cv.p, static_cast<size_t>(mc), 32, dq::key_bits(n_p), s);
    DevBuf<int64_t> rptr(0, s), cptr(0, s);
    dq::build_ptr(rk.p, mr, n_v, rptr, s);
    dq::build_ptr(ck.p, mc, n_p, cptr, s);
    DevBuf<u32> rrow(mr > 0 ? mr : 1, s), rcol(mr > 0 ? mr : 1, s), ccol(mc > 0 ? mc : 1, s), crow(mc > 0 ? mc : 1, s);
    if (mr > 0) dq::split_kv<<<dq::grid_of(mr), 256, 0, s>>>(rk.p, mr, rrow.p, rcol.p);
    if (mc > 0) dq::split_kv<<<dq::grid_of(mc), 256, 0, s>>>(ck.p, mc, ccol.p, crow.p);
    // Row accumulation in LDS when V fits a few windows.
    int lds_cols = dq::kEffMaxCols;
    if (const char* e = std::getenv("DQ_HEFF_LDS_COLS")) {  // tuning and tests: narrower windows
      const long v = std::strtol(e, nullptr, 10);
      if (v >= dq::kWave && v <= dq::kEffMaxCols) lds_cols = static_cast<int>(v / dq::kWave * dq::kWave);
    }
    const int64_t win64 = (n_v + dq::kWave - 1) / dq::kWave * dq::kWave;
    const int win = static_cast<int>(win64 < lds_cols ? win64 : lds_cols);
    const bool lds_path = n_v > 0 && (n_v + win - 1) / win <= dq::kEffMaxPasses && mc < (1LL << 32);
    // dinv and skip (and the column records for the LDS path)
    DevBuf<double> dinv(n_p, s);
    DevBuf<unsigned char> skip(n_p, s);
    DevBuf<dq::ColRec> crec(lds_path ? n_p : 0, s);
    dq::dinv_kernel<<<dq::grid_of(n_p), 256, 0, s>>>(d_h_pp, n_p, e_ref, epsilon, reg_sigma, cptr.p, cv.p, dinv.p,
                                                      skip.p, lds_path ? crec.p : nullptr);
    if (lds_path) {
      const size_t m1 = mr > 0 ? static_cast<size_t>(mr) : 1;
      DevBuf<int64_t> nsub(m1, s), mpos(mr + 1, s);
      DevBuf<double> e_d(m1, s), e_as(m1, s), diag(n_v, s);
      DevBuf<u32> e_q0(m1, s), e_len(m1, s);
      if (mr > 0)
        dq::eff_entries<<<dq::grid_of(mr), 256, 0, s>>>(rcol.p, rv.p, crec.p, mr, e_d.p, e_as.p, e_q0.p, e_len.p,
                                                         nsub.p);
      dq::exclusive_scan_i64(nsub.p, mr, mpos.p, s);
      const int64_t ns = dq::read_i64(mpos.p + mr, s);
      DevBuf<int64_t> sptr(n_v + 1, s), vvptr(0, s), cap(n_v, s), soff(n_v + 1, s), cnt(n_v, s), ooff(n_v + 1, s);
      DevBuf<double> st_as(ns > 0 ? ns : 1, s);
      DevBuf<u32> st_q(ns > 0 ? ns : 1, s), st_lm1(ns > 0 ? ns : 1, s);
      if (mr > 0)
        dq::eff_steps<<<dq::grid_of(mr), 256, 0, s>>>(e_as.p, e_q0.p, e_len.p, nsub.p, mpos.p, mr, st_q.p,
                                                       st_lm1.p, st_as.p);
      dq::eff_row_ptr<<<dq::grid_of(n_v + 1), 256, 0, s>>>(rptr.p, mpos.p, n_v, sptr.p);
      dq::eff_diag<<<static_cast<unsigned>((n_v + 63) / 64), 64, 0, s>>>(rptr.p, e_d.p, n_v, diag.p);
      dq::build_ptr(vvk.p, m_vv, n_v, vvptr, s);
      const unsigned rows4 = static_cast<unsigned>((n_v + 3) / 4);
      dq::eff_caps<<<rows4, 4 * dq::kWave, 0, s>>>(sptr.p, st_lm1.p, vvptr.p, n_v, cap.p);
      dq::exclusive_scan_i64(cap.p, n_v, soff.p, s);
      const int64_t nstage = dq::read_i64(soff.p + n_v, s);
      DevBuf<u32> scol(nstage > 0 ? nstage : 1, s);
      DevBuf<double> sval(nstage > 0 ? nstage : 1, s);
      dq::eff_rows<<<static_cast<unsigned>(n_v), dq::kWave, static_cast<size_t>(win + dq::kWave) * sizeof(double),
                     s>>>(sptr.p, st_q.p, st_lm1.p, st_as.p, crow.p, cv.p, diag.p, vvptr.p, vvk.p, vvv.p, n_v, win,
                          soff.p, scol.p, sval.p, cnt.p);
      DQ_HIP(hipGetLastError());
      dq::exclusive_scan_i64(cnt.p, n_v, ooff.p, s);
      const int64_t nf = dq::read_i64(ooff.p + n_v, s);
      B->vv_row.alloc(nf);
      B->vv_col.alloc(nf);
      B->vv_val.alloc(nf);
      if (nf > 0)
        dq::eff_gather<<<rows4, 4 * dq::kWave, 0, s>>>(soff.p, ooff.p, scol.p, sval.p, n_v, B->vv_row.p,
                                                        B->vv_col.p, B->vv_val.p);
      DQ_HIP(hipGetLastError());
      B->nnz_vv = nf;
      DQ_HIP(hipStreamSynchronize(s));
      *out = B.release();
      return;
    }
    // Otherwise: triples, one output range per H_VP entry in row-major order,
    // stable-sorted by (i, j) and folded.
    DevBuf<int64_t> tcnt(mr > 0 ? mr : 1, s), toff(mr + 1, s);
    if (mr > 0)
      dq::triple_count<<<dq::grid_of(mr), 256, 0, s>>>(rcol.p, rv.p, cptr.p, dinv.p, skip.p, mr, tcnt.p);
    dq::exclusive_scan_i64(tcnt.p, mr, toff.p, s);
    const int64_t nt = dq::read_i64(toff.p + mr, s);
    DevBuf<u64> tk(nt > 0 ? nt : 1, s), tk_s(nt > 0 ? nt : 1, s), ak(0, s);
    DevBuf<double> tv(nt > 0 ? nt : 1, s), tv_s(nt > 0 ? nt : 1, s), av(0, s);
    if (nt > 0)
      dq::triple_fill<<<static_cast<unsigned>((mr + dq::kTripleBlock - 1) / dq::kTripleBlock), dq::kTripleBlock,
                        0, s>>>(rrow.p, rcol.p, rv.p, cptr.p, crow.p, cv.p, dinv.p, mr, toff.p, tk.p, tv.p);
    DQ_HIP(hipGetLastError());
    dq::sort_pairs_u64_f64(tk.p, tv.p, tk_s.p, tv_s.p, static_cast<size_t>(nt), vbits, s);
    const int64_t na = dq::reduce_sorted(dq::kFold, tk_s.p, tv_s.p, nt, ak, av, s);
    // H_eff = csr_add(VV, ADD, 1e-15): stable sort of [VV ; ADD]
    const int64_t nc = m_vv + na;
    DevBuf<u64> ek(nc > 0 ? nc : 1, s), ek_s(nc > 0 ? nc : 1, s), fk(0, s);
    DevBuf<double> ev(nc > 0 ? nc : 1, s), ev_s(nc > 0 ? nc : 1, s), fv(0, s);
    if (m_vv > 0) {
      DQ_HIP(hipMemcpyAsync(ek.p, vvk.p, m_vv * sizeof(u64), hipMemcpyDeviceToDevice, s));
      DQ_HIP(hipMemcpyAsync(ev.p, vvv.p, m_vv * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    if (na > 0) {
      DQ_HIP(hipMemcpyAsync(ek.p + m_vv, ak.p, na * sizeof(u64), hipMemcpyDeviceToDevice, s));
      DQ_HIP(hipMemcpyAsync(ev.p + m_vv, av.p, na * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    dq::sort_pairs_u64_f64(ek.p, ev.p, ek_s.p, ev_s.p, static_cast<size_t>(nc), vbits, s);
    const int64_t nf = dq::reduce_sorted(dq::kAdd, ek_s.p, ev_s.p, nc, fk, fv, s);
    blocks_from_keys(*B, fk, fv, nf, vbits, false, s);
    DQ_HIP(hipStreamSynchronize(s));
    *out = B.release();
  });
}

}  // extern "C"

// ----------------------------------------------------------------------------- allocation-free energy
// compute_variational_energy (string method) for callers that cannot allocate
// device memory.  Every temporary (DetMap hash, string index and its build
// scratch, sigma when the caller passes none, amplitudes in member order,
// block sums) is carved from `scratch`, a fixed array in each device's copy of
// this code object.  Every call on a GPU shares that GPU's array, so calls that
// overlap on one GPU race; everything the kernels read is written earlier in
// the same call, so nothing left by a previous call is used.
namespace {

constexpr int64_t kNoAllocMaxRows = DQ_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocBytes = DQ_NOALLOC_SCRATCH_BYTES;
constexpr size_t kNoAllocAlign = 256;

std::mutex g_scratch_mu;
void* g_scratch_addr[64] = {};  // device address of `scratch`, per device

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// The device address of this device's `scratch`, resolved once per device
// with that device current (the host-side name is not a device address).
unsigned char* scratch_on_current_device() {
  int dev = 0;
  DQ_HIP(hipGetDevice(&dev));
  DQ_REQUIRE(dev >= 0 && dev < 64, DQ_ERUNTIME, "device index out of range for the allocation-free path");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* ptr = nullptr;
    DQ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    DQ_REQUIRE(ptr != nullptr, DQ_ERUNTIME, "no device address for the allocation-free scratch array");
    g_scratch_addr[dev] = ptr;
  }
  return static_cast<unsigned char*>(g_scratch_addr[dev]);
}

struct NoAllocLayout {
  size_t hash_vals, hash_keys, xmem, xscr, sigma, ca, cb, part, total;
};

size_t na_align(size_t b) { return (b + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; }

NoAllocLayout noalloc_layout(int64_t n, bool own_sigma) {
  const size_t un = static_cast<size_t>(n);
  const size_t cap = dq::hash_capacity(n);
  NoAllocLayout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += na_align(bytes);
    return at;
  };
  L.hash_vals = take(4 * cap);
  L.hash_keys = take(16 * cap);
  L.xmem = take(dq::str_index_bytes(n));
  L.xscr = take(dq::str_index_scratch_bytes(n));
  L.sigma = take(own_sigma ? 8 * un : 0);
  L.ca = take(8 * un);
  L.cb = take(8 * un);
  L.part = take(8 * ((un + 255) / 256 + 1));
  L.total = o;
  return L;
}

}  // namespace

extern "C" {

int dq_noalloc_limits(int64_t* max_rows, int64_t* scratch_bytes) {
  return dq::guarded([&] {
    if (max_rows) *max_rows = kNoAllocMaxRows;
    if (scratch_bytes) *scratch_bytes = static_cast<int64_t>(kNoAllocBytes);
  });
}

int dq_variational_energy_noalloc(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs, int64_t n,
                                  int n_orb, int use_hb, double eps1, double* d_sigma, double* d_energy,
                                  void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx, DQ_EINVAL, "null context");
    DQ_REQUIRE(d_energy, DQ_EINVAL, "null energy output");
    DQ_REQUIRE(n >= 0 && (n == 0 || (d_dets && d_coeffs)), DQ_EINVAL, "bad arguments");
    DQ_REQUIRE(n <= kNoAllocMaxRows, DQ_ERANGE,
               "dq_variational_energy_noalloc: " + std::to_string(n) + " rows exceed the allocation-free maximum of " +
                   std::to_string(kNoAllocMaxRows));
    DQ_REQUIRE(n_orb >= 1 && n_orb <= ctx->n_orb, DQ_EINVAL, "n_orb must be in [1, n_orb of the integral context]");
    auto* c = const_cast<dq_ctx*>(ctx);
    if (use_hb)
      DQ_REQUIRE(c->hb, DQ_EINVAL, "compute_variational_energy: Heat-Bath mode requires non-null hb_table");
    hipStream_t s = dq::as_stream(stream);
    unsigned char* base = scratch_on_current_device();
    const NoAllocLayout L = noalloc_layout(n, d_sigma == nullptr);
    DQ_REQUIRE(L.total <= kNoAllocBytes, DQ_ERANGE,
               "dq_variational_energy_noalloc: " + std::to_string(L.total) + " scratch bytes needed, " +
                   std::to_string(kNoAllocBytes) + " available");
    double* part = reinterpret_cast<double*>(base + L.part);
    if (n == 0) {
      dq::fill_u32(part, 2, 0u, s);  // +0.0
      DQ_HIP(hipMemcpyAsync(d_energy, part, sizeof(double), hipMemcpyDeviceToDevice, s));
      return;
    }
    const dq::u64 cap = dq::hash_capacity(n);
    auto* vals = reinterpret_cast<dq::u32*>(base + L.hash_vals);
    auto* keys = reinterpret_cast<dq::u64*>(base + L.hash_keys);
    dq::hash_build_async(d_dets, n, keys, vals, cap - 1, s);  // first occurrence wins (local_conn.cpp:171-172)
    const dq::HashView H{keys, vals, cap - 1};
    const dq::StrIndex X = dq::str_index_build(d_dets, n, H, base + L.xmem, base + L.xscr, s);
    double* sig = d_sigma ? d_sigma : reinterpret_cast<double*>(base + L.sigma);
    double* ca = reinterpret_cast<double*>(base + L.ca);
    double* cb = reinterpret_cast<double*>(base + L.cb);
    dq::str_gather_coef(X, d_coeffs, ca, cb, s);
    double st, dmin;
    dq::eloc_thresholds(*c, use_hb, eps1, &st, &dmin);
    const unsigned g = static_cast<unsigned>((n + dq::kWavesPerBlock - 1) / dq::kWavesPerBlock);
    dq::energy_str_kernel<<<g, dq::kWave * dq::kWavesPerBlock, 0, s>>>(c->view(), d_dets, d_coeffs, n, n_orb, st,
                                                                        dmin, X, dq::StrCoef{ca, cb}, sig,
                                                                        c->d_err.p);
    const int64_t nb = (n + 255) / 256;
    dq::dot_blocks<<<static_cast<unsigned>(nb), 256, 0, s>>>(d_coeffs, sig, n, part);
    dq::sum_serial<<<1, 1, 0, s>>>(part, nb, d_energy);
    DQ_HIP(hipGetLastError());
    // Rows with orbitals at or above n_orb are reported here; a caller that is
    // capturing a graph cannot wait, and gets them from the next checked call.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    DQ_HIP(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) c->check_rows(s);
  });
}

}  // extern "C"
