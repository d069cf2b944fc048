// Connected / perturbative determinant spaces (P-space discovery) and the
// Epstein–Nesbet PT2 correction, on gfx950.
//
//  * gen_connected_dets (bridge.cpp:262-271; det_space.cpp:219-227, 94-98):
//    the sorted unique set of every single and double excitation of V.
//  * gen_perturbative_dets (bridge.cpp:278-338): P = C \ V, sorted.
//      none    (det_space.cpp:233-255): C = all singles and doubles;
//      static  (ham_utils.cpp:546-701): singles with |H| > 1e-15, doubles of
//              the heat-bath prefix w >= eps1, both added orbitals virtual;
//      dynamic the same with singles also needing |H psi_i| >= eps1 and the
//              doubles cutoff eps1 / max(|psi_i|, 1e-12).
//  * compute_pt2 (local_conn.cpp:247-418):
//      internal = sum_i r_i^2 / (E - H_ii), r_i = sigma_i - E c_i, sigma_i
//                 over V with every excitation (no screening);
//      external = sum_a v_a^2 / (E - H_aa), v_a = sum_i H_ai c_i over the
//                 excitations a of V outside V with |H_ai| > 1e-15 and
//                 |H_ai c_i| >= eps1, doubles from the heat-bath prefix
//                 w >= eps1 / max(|c_i|, 1e-12) or all doubles;
//      rows with c_i^2 < 1e-15 take no part; |E - H| < 1e-15 is clamped to
//      +-1e-15 by its sign.
//
// Design.  Candidates are generated one wavefront per row (the shared row
// walker of walk.hpp for screened sets and PT2, a closed-form enumerator for
// the unscreened sets, which need no integrals) in a count pass, an int64
// scan, and a fill pass into one HBM buffer per chunk of rows (chunks bound
// the candidate buffer).  Each chunk is sorted by determinant with a stable
// radix sort — on a packed 2 n_orb-bit key (alpha << n_orb | beta) when
// n_orb <= 32, on the 128-bit (alpha, beta) pair otherwise — then reduced:
// run heads (minus members of V) for sets, sequential sums of each run in
// emission order for PT2 amplitudes.  Within one chunk a run's sum therefore
// adds H_ai c_i in row order, as the reference does with one thread; partial
// sums of several chunks are added in chunk order.  H_aa, the terms and
// their fixed-order sums follow.  The internal term uses the V-driven string
// index (eloc_strings.hpp) for sigma_i.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "assemble.hpp"
#include "eloc_strings.hpp"
#include "hash.hpp"
#include "rowrank.hpp"
#include "strings.hpp"
#include "walk.hpp"

struct dq_detset {
  dq::PoolBuf<dq::u64> dets;  // released on the null stream (the caller's stream may be gone)
  dq::PoolBuf<double> vals;
  int64_t n = 0;
  bool has_vals = false;
};

namespace dq {

namespace {

constexpr int kWpb = 4;                          // waves per block of the row kernels
constexpr int64_t kDefaultChunk = int64_t(1) << 28;  // candidates per pass

unsigned grid_for(int64_t n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

// ------------------------------------------------------------------ sorting
struct DetKey {
  u64 a, b;
};
struct DetDecomp {
  __host__ __device__ hipcub::tuple<u64&, u64&> operator()(DetKey& k) const {
    return hipcub::tuple<u64&, u64&>(k.a, k.b);
  }
};

__global__ void pack_keys(const u64* __restrict__ d, int64_t n, int n_orb, u64* __restrict__ key) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) key[i] = (d[2 * i] << n_orb) | d[2 * i + 1];
}

__global__ void unpack_keys(const u64* __restrict__ key, int64_t n, int n_orb, u64* __restrict__ d) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    reinterpret_cast<ulonglong2*>(d)[i] = make_ulonglong2(key[i] >> n_orb, key[i] & low_mask(n_orb));
}

// Stable ascending sort of n determinants by (alpha, beta), carrying vin
// (may be NULL); asynchronous on s.
void sort_dets(const u64* din, const double* vin, u64* dout, double* vout, int64_t n, int n_orb, hipStream_t s) {
  if (n == 0) return;
  const size_t N = static_cast<size_t>(n);
  size_t tmp = 0;
  if (n_orb <= 32) {
    PoolBuf<u64> k1(N, s), k2(N, s);
    pack_keys<<<grid_for(n, 256), 256, 0, s>>>(din, n, n_orb, k1.p);
    const int bits = 2 * n_orb;
    if (vin) {
      DQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k1.p, k2.p, vin, vout, N, 0, bits, s));
      PoolBuf<unsigned char> t(tmp + 1, s);  // never NULL: a NULL buffer only queries the size
      DQ_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, k1.p, k2.p, vin, vout, N, 0, bits, s));
    } else {
      DQ_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, k1.p, k2.p, N, 0, bits, s));
      PoolBuf<unsigned char> t(tmp + 1, s);  // never NULL: a NULL buffer only queries the size
      DQ_HIP(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, k1.p, k2.p, N, 0, bits, s));
    }
    unpack_keys<<<grid_for(n, 256), 256, 0, s>>>(k2.p, n, n_orb, dout);
  } else {
    const DetKey* ki = reinterpret_cast<const DetKey*>(din);
    DetKey* ko = reinterpret_cast<DetKey*>(dout);
    if (vin) {
      DQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ki, ko, vin, vout, N, DetDecomp{}, s));
      PoolBuf<unsigned char> t(tmp + 1, s);  // never NULL: a NULL buffer only queries the size
      DQ_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, ki, ko, vin, vout, N, DetDecomp{}, s));
    } else {
      DQ_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, ki, ko, N, DetDecomp{}, s));
      PoolBuf<unsigned char> t(tmp + 1, s);  // never NULL: a NULL buffer only queries the size
      DQ_HIP(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, ki, ko, N, DetDecomp{}, s));
    }
  }
  DQ_HIP(hipGetLastError());
}

// flag[i] = 1 if sorted det i starts a run and is not in H (H may be empty).
__global__ void head_flags(const u64* __restrict__ d, int64_t n, HashView H, int64_t* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 a = d[2 * i], b = d[2 * i + 1];
  bool head = i == 0 || a != d[2 * i - 2] || b != d[2 * i - 1];
  if (head && hash_find(H, a, b) >= 0) head = false;
  flag[i] = head ? 1 : 0;
}

// pos = exclusive scan of the flags (n + 1 entries): det i is kept iff
// pos[i + 1] > pos[i].
__global__ void compact_dets(const u64* __restrict__ d, const int64_t* __restrict__ pos, int64_t n,
                             u64* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || pos[i + 1] == pos[i]) return;
  reinterpret_cast<ulonglong2*>(out)[pos[i]] = reinterpret_cast<const ulonglong2*>(d)[i];
}

// Run heads (flags from head_flags with an empty H): the run's values summed
// in their stored order, starting from 0.0 as an accumulator map does.
// Run r of the sorted 128-bit dets spans [hpos[r], hpos[r + 1]); pos is the
// exclusive scan of the head flags (pos[i + 1] != pos[i] at a head).
__global__ void run_starts(const int64_t* __restrict__ pos, int64_t n, int64_t* __restrict__ hpos) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (pos[i + 1] != pos[i]) hpos[pos[i]] = i;
  if (i == 0) hpos[pos[n]] = n;
}

// One thread per run, its bounds known, values summed in order from 0.0.
__global__ void fold_runs(const u64* __restrict__ d, const double* __restrict__ v, const int64_t* __restrict__ hpos,
                          int64_t m, u64* __restrict__ od, double* __restrict__ ov) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int64_t b = hpos[r], e = hpos[r + 1];
  double acc = 0.0;
#pragma unroll 8
  for (int64_t j = b; j < e; ++j) acc += v[j];
  reinterpret_cast<ulonglong2*>(od)[r] = reinterpret_cast<const ulonglong2*>(d)[b];
  ov[r] = acc;
}

int64_t read_last(const int64_t* d, int64_t n, hipStream_t s) {
  int64_t h = 0;
  DQ_HIP(hipMemcpyAsync(&h, d + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  return h;
}

// Sorted unique dets of d[0..n) minus H into a new buffer; returns the count.
int64_t unique_minus(const u64* d, int64_t n, int n_orb, const HashView& H, PoolBuf<u64>& out, hipStream_t s) {
  if (n == 0) {
    out.alloc(0, s);
    return 0;
  }
  PoolBuf<u64> sorted(2 * static_cast<size_t>(n), s);
  sort_dets(d, nullptr, sorted.p, nullptr, n, n_orb, s);
  PoolBuf<int64_t> pos(static_cast<size_t>(n) + 1, s);
  head_flags<<<grid_for(n, 256), 256, 0, s>>>(sorted.p, n, H, pos.p);
  exclusive_scan_i64(pos.p, n, pos.p, s);
  const int64_t m = read_last(pos.p, n, s);
  out.alloc(2 * static_cast<size_t>(m), s);
  compact_dets<<<grid_for(n, 256), 256, 0, s>>>(sorted.p, pos.p, n, out.p);
  DQ_HIP(hipGetLastError());
  return m;
}

// Sorted unique dets of d[0..n) with the values of each run summed in order.
int64_t reduce_runs(const u64* d, const double* v, int64_t n, int n_orb, PoolBuf<u64>& od, PoolBuf<double>& ov,
                    hipStream_t s) {
  if (n == 0) {
    od.alloc(0, s);
    ov.alloc(0, s);
    return 0;
  }
  PoolBuf<u64> sd(2 * static_cast<size_t>(n), s);
  PoolBuf<double> sv(static_cast<size_t>(n), s);
  sort_dets(d, v, sd.p, sv.p, n, n_orb, s);
  PoolBuf<int64_t> pos(static_cast<size_t>(n) + 1, s);
  head_flags<<<grid_for(n, 256), 256, 0, s>>>(sd.p, n, HashView{}, pos.p);
  exclusive_scan_i64(pos.p, n, pos.p, s);
  const int64_t m = read_last(pos.p, n, s);
  od.alloc(2 * static_cast<size_t>(m), s);
  ov.alloc(static_cast<size_t>(m), s);
  PoolBuf<int64_t> hpos(static_cast<size_t>(m) + 1, s);
  run_starts<<<grid_for(n, 256), 256, 0, s>>>(pos.p, n, hpos.p);
  fold_runs<<<grid_for(m, 256), 256, 0, s>>>(sd.p, sv.p, hpos.p, m, od.p, ov.p);
  DQ_HIP(hipGetLastError());
  return m;
}

// ------------------------------------------------------------------ fixed-order sums
__global__ void block_sums(const double* __restrict__ x, int64_t n, double* __restrict__ part) {
  __shared__ double buf[256];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  buf[threadIdx.x] = i < n ? x[i] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) buf[threadIdx.x] += buf[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = buf[0];
}

__global__ void serial_sum(const double* __restrict__ part, int64_t m, double* __restrict__ out) {
  double t = 0.0;
  for (int64_t i = 0; i < m; ++i) t += part[i];
  *out = t;
}

// Sum of x[0..n): 256-element trees, level by level, until at most 256
// partial sums remain, then those in order; the order depends on n only.
// Synchronises.
double fixed_sum(const double* x, int64_t n, hipStream_t s) {
  if (n == 0) return 0.0;
  PoolBuf<double> a, b;
  const double* cur = x;
  int64_t m = n;
  while (m > 256) {
    const int64_t nb = (m + 255) / 256;
    PoolBuf<double>& dst = cur == a.p ? b : a;
    dst.alloc(static_cast<size_t>(nb), s);
    block_sums<<<static_cast<unsigned>(nb), 256, 0, s>>>(cur, m, dst.p);
    cur = dst.p;
    m = nb;
  }
  PoolBuf<double> out(1, s);
  serial_sum<<<1, 1, 0, s>>>(cur, m, out.p);
  DQ_HIP(hipGetLastError());
  double h = 0.0;
  DQ_HIP(hipMemcpyAsync(&h, out.p, sizeof(double), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  return h;
}

__device__ __forceinline__ double clamp_denom(double d) {
  if (fabs(d) < 1e-15) d = d >= 0.0 ? 1e-15 : -1e-15;  // local_conn.cpp:327-330, 402-405
  return d;
}

// ------------------------------------------------------------------ candidate layout
// Candidates leave the generators either as packed keys (n_orb <= 32: one u64
// whose order is the (alpha, beta) order) or as (alpha, beta) pairs.  A packed
// key is alpha << n_orb | beta, or, when every row of V lies in one
// (n_alpha, n_beta) sector and the colex ranks need fewer radix passes,
// rank(alpha) << shift | rank(beta): colex ranks (the combinatorial number
// system) order strings of one popcount like their integer values, so the
// key order is again the (alpha, beta) order.  The generator computes the
// ranks from the row's rank tables (rowrank.hpp), a few lookups per ket.
struct CandOut {
  u64* key = nullptr;         // packed layout
  ulonglong2* det = nullptr;  // pair layout
  int shift = 0;
  bool colex = false;         // packed keys are colex ranks (the generator computes them)
  __device__ __forceinline__ void put(int64_t at, u64 a, u64 b) const {
    if (key)
      key[at] = (a << shift) | b;
    else
      det[at] = make_ulonglong2(a, b);
  }
};

// How packed keys decode (and how many bits the radix sort needs).
struct KeyCodec {
  int shift = 0;
  int bits = 0;
  const u32* sa = nullptr;  // colex: rank -> alpha string (null: packed words)
  const u32* sb = nullptr;  // colex: rank -> beta string
  bool ranked_hash = false;  // the membership map holds (rank, rank) pairs
  __device__ __forceinline__ void decode(u64 k, u64& a, u64& b) const {
    const u64 hi = k >> shift, lo = k & low_mask(shift);
    a = sa ? sa[hi] : hi;
    b = sb ? sb[lo] : lo;
  }
  // The pair to look up in the membership map.
  __device__ __forceinline__ void probe_pair(u64 k, u64& a, u64& b) const {
    if (ranked_hash) {
      a = k >> shift;
      b = k & low_mask(shift);
    } else {
      decode(k, a, b);
    }
  }
};

KeyCodec packed_codec(int n_orb) {
  KeyCodec c;
  c.shift = n_orb;
  c.bits = 2 * n_orb;
  return c;
}

bool packed_layout(int n_orb) { return n_orb <= 32; }

CandOut cand_out(void* buf, int n_orb, const KeyCodec& kc) {
  CandOut o;
  if (packed_layout(n_orb))
    o.key = static_cast<u64*>(buf);
  else
    o.det = static_cast<ulonglong2*>(buf);
  o.shift = packed_layout(n_orb) ? kc.shift : n_orb;
  o.colex = packed_layout(n_orb) && kc.sa != nullptr;
  return o;
}

// (rank(alpha), rank(beta)) of each row (one sector, n_orb <= 32).
__global__ void colex_pairs(const u64* __restrict__ dets, int64_t n, u64* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = colex_rank(kBinomDev, static_cast<u32>(dets[2 * i]));
  out[2 * i + 1] = colex_rank(kBinomDev, static_cast<u32>(dets[2 * i + 1]));
}

// The head kernel's prefilter: 2^18 bits (32 KB), staged in LDS; about 7%
// of the non-members pass at 20,000 keys.
constexpr int kFiltLg = 18;
constexpr int kFiltWords = (1 << kFiltLg) / 32;

// A membership test for sorted candidate keys: the map and its prefilter.
struct Members {
  HashView H;
  const u32* filt = nullptr;
};

void build_filter(const u64* pairs, int64_t n, PoolBuf<u32>& bits, hipStream_t s) {
  bits.alloc(kFiltWords, s);
  filter_build(pairs, n, kFiltLg, bits.p, s);
}

// Membership map of the rows keyed by their colex rank pairs, so that sorted
// colex keys are looked up without decoding them to strings; with its filter.
void rank_map(const u64* dets, int64_t n, DetHash& M, PoolBuf<u32>& filt, hipStream_t s) {
  PoolBuf<u64> pr(2 * static_cast<size_t>(n), s);
  colex_pairs<<<grid_for(n, 256), 256, 0, s>>>(dets, n, pr.p);
  DQ_HIP(hipGetLastError());
  M.build(pr.p, n, s);
  build_filter(pr.p, n, filt, s);
}

int bits_of(u64 count) {  // bits for the values 0 .. count - 1 (at least 1)
  int b = 1;
  while (b < 63 && (1ull << b) < count) ++b;
  return b;
}

// Electron counts of the rows dets[0..n): mm = {min na, max na, min nb, max nb}.
void electron_range(const u64* dets, int64_t n, int mm[4], hipStream_t s) {
  mm[0] = mm[2] = 64;
  mm[1] = mm[3] = -1;
  if (n == 0) return;
  PoolBuf<int> dmm(4, s);
  DQ_HIP(hipMemcpyAsync(dmm.p, mm, 4 * sizeof(int), hipMemcpyHostToDevice, s));
  pop_range<<<grid_for(n, 256), 256, 0, s>>>(dets, n, dmm.p);
  DQ_HIP(hipGetLastError());
  DQ_HIP(hipMemcpyAsync(mm, dmm.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
}

// Colex codec for rows with electron counts mm (electron_range) when they
// share one sector and the ranks save a radix pass; the packed codec
// otherwise.  sa / sb hold the decode tables while the codec is in use.
KeyCodec choose_codec(const int mm[4], int n_orb, PoolBuf<u32>& sa, PoolBuf<u32>& sb, hipStream_t s) {
  KeyCodec kc = packed_codec(n_orb);
  if (!packed_layout(n_orb) || mm[1] < 0) return kc;
  if (mm[0] != mm[1] || mm[2] != mm[3] || mm[1] > n_orb || mm[3] > n_orb) return kc;
  const u64 ca = binom_host(n_orb, mm[0]), cb = binom_host(n_orb, mm[2]);
  constexpr u64 kMaxTable = u64(1) << 26;
  if (ca == 0 || cb == 0 || ca > kMaxTable || cb > kMaxTable) return kc;
  const int ba = bits_of(ca), bb = bits_of(cb);
  if ((ba + bb + 7) / 8 >= (2 * n_orb + 7) / 8) return kc;  // no radix pass saved
  sa.alloc(ca, s);
  sb.alloc(cb, s);
  colex_strings<<<static_cast<unsigned>((ca + 255) / 256), 256, 0, s>>>(sa.p, static_cast<u32>(ca), mm[0], n_orb);
  colex_strings<<<static_cast<unsigned>((cb + 255) / 256), 256, 0, s>>>(sb.p, static_cast<u32>(cb), mm[2], n_orb);
  DQ_HIP(hipGetLastError());
  kc.shift = bb;
  kc.bits = ba + bb;
  kc.sa = sa.p;
  kc.sb = sb.p;
  return kc;
}

// Exclusive sum of n (< 2^32) u32 flags into pos, one decoupled look-back
// pass; returns the total.  Synchronises.
int64_t scan_flags(const u32* flag, u32* pos, int64_t n, hipStream_t s) {
  size_t tmp = 0;
  DQ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flag, pos, n, s));
  {
    PoolBuf<unsigned char> t(tmp + 1, s);
    DQ_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, flag, pos, n, s));
  }
  u32 last[2] = {0, 0};
  DQ_HIP(hipMemcpyAsync(&last[0], pos + n - 1, sizeof(u32), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipMemcpyAsync(&last[1], flag + n - 1, sizeof(u32), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  return static_cast<int64_t>(last[0]) + last[1];
}

// Packed keys: run heads (minus members of H) and their unpacked pairs.
__global__ void head_flags_keys(const u64* __restrict__ k, int64_t n, KeyCodec kc, HashView H,
                                u32* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool head = i == 0 || k[i] != k[i - 1];
  if (head && H.vals) {
    u64 a, b;
    kc.probe_pair(k[i], a, b);
    if (hash_find(H, a, b) >= 0) head = false;
  }
  flag[i] = head ? 1u : 0u;
}

// head_flags_keys against a map with its prefilter: the filter is staged in
// LDS once per block and the blocks stride over the keys.
constexpr int kHeadThreads = 512;

__global__ __launch_bounds__(kHeadThreads) void head_flags_keys_filt(const u64* __restrict__ k, int64_t n,
                                                                      KeyCodec kc, HashView H,
                                                                      const u32* __restrict__ filt,
                                                                      u32* __restrict__ flag) {
  __shared__ u32 fs[kFiltWords];
  for (int t = threadIdx.x; t < kFiltWords; t += blockDim.x) fs[t] = filt[t];
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const u64 key = k[i];
    bool head = i == 0 || key != k[i - 1];
    if (head) {
      u64 a, b;
      kc.probe_pair(key, a, b);
      if (filt_test(fs, filt_slot(a, b, kFiltLg)) && hash_find(H, a, b) >= 0) head = false;
    }
    flag[i] = head ? 1u : 0u;
  }
}

void flag_heads(const u64* k, int64_t n, const KeyCodec& kc, const Members& M, u32* flag, hipStream_t s) {
  if (M.H.vals && M.filt) {
    int dev = 0, cus = 0;
    DQ_HIP(hipGetDevice(&dev));
    DQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t want = (n + kHeadThreads - 1) / kHeadThreads, cap = 4 * static_cast<int64_t>(cus);
    head_flags_keys_filt<<<static_cast<unsigned>(want < cap ? want : cap), kHeadThreads, 0, s>>>(k, n, kc, M.H,
                                                                                                 M.filt, flag);
  } else {
    head_flags_keys<<<grid_for(n, 256), 256, 0, s>>>(k, n, kc, M.H, flag);
  }
  DQ_HIP(hipGetLastError());
}

__global__ void compact_keys(const u64* __restrict__ k, const u32* __restrict__ flag, const u32* __restrict__ pos,
                             int64_t n, KeyCodec kc, u64* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  u64 a, b;
  kc.decode(k[i], a, b);
  reinterpret_cast<ulonglong2*>(out)[pos[i]] = make_ulonglong2(a, b);
}

// Heads sum their run in stored order from 0.0.  Packed keys: run r spans
// [hpos[r], hpos[r + 1]) (n < 2^32); with `stop` (heads flagged against a
// non-empty H, so the runs of keys in H have no head) it ends at the first
// other key.
__global__ void run_starts_keys(const u32* __restrict__ flag, const u32* __restrict__ pos, int64_t n, u32 m,
                                u32* __restrict__ hpos) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) hpos[pos[i]] = static_cast<u32>(i);
  if (i == 0) hpos[m] = static_cast<u32>(n);
}

__global__ void fold_runs_keys(const u64* __restrict__ k, const double* __restrict__ v, const u32* __restrict__ hpos,
                               int64_t m, KeyCodec kc, int stop, u64* __restrict__ od, double* __restrict__ ov) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const u32 b = hpos[r];
  u32 e = hpos[r + 1];
  const u64 key = k[b];
  if (stop && k[e - 1] != key) {  // runs of keys in H follow: the run ends at the first other key
    u32 lo = b, hi = e - 1;         // k[lo] == key, k[hi] != key (sorted)
    while (hi - lo > 1) {
      const u32 mid = lo + (hi - lo) / 2;
      if (k[mid] == key)
        lo = mid;
      else
        hi = mid;
    }
    e = hi;
  }
  double acc = 0.0;
#pragma unroll 8
  for (u32 j = b; j < e; ++j) acc += v[j];
  u64 wa, wb;
  kc.decode(key, wa, wb);
  reinterpret_cast<ulonglong2*>(od)[r] = make_ulonglong2(wa, wb);
  ov[r] = acc;
}

// Sorted unique pairs of the candidates cand[0..n) (either layout) minus H.
int64_t unique_cands(void* cand, int64_t n, int n_orb, const KeyCodec& kc, const Members& M, PoolBuf<u64>& out,
                     hipStream_t s) {
  if (!packed_layout(n_orb)) return unique_minus(static_cast<const u64*>(cand), n, n_orb, M.H, out, s);
  if (n == 0) {
    out.alloc(0, s);
    return 0;
  }
  DQ_REQUIRE(n < (int64_t(1) << 32), DQ_EINVAL, "too many candidates in one pass");
  const size_t N = static_cast<size_t>(n);
  PoolBuf<u64> sk(N, s);
  size_t tmp = 0;
  const u64* k = static_cast<const u64*>(cand);
  DQ_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, k, sk.p, N, 0, kc.bits, s));
  {
    PoolBuf<unsigned char> t(tmp + 1, s);
    DQ_HIP(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, k, sk.p, N, 0, kc.bits, s));
  }
  PoolBuf<u32> flag(N, s), pos(N, s);
  flag_heads(sk.p, n, kc, M, flag.p, s);
  const int64_t m = scan_flags(flag.p, pos.p, n, s);
  out.alloc(2 * static_cast<size_t>(m), s);
  compact_keys<<<grid_for(n, 256), 256, 0, s>>>(sk.p, flag.p, pos.p, n, kc, out.p);
  DQ_HIP(hipGetLastError());
  return m;
}

// Sorted unique pairs of the candidates with each run's values summed in their
// (stable) order; packed keys found in H are dropped (the pair layout takes
// candidates already free of H).
int64_t reduce_cands(void* cand, const double* v, int64_t n, int n_orb, const KeyCodec& kc, const Members& M,
                     PoolBuf<u64>& od, PoolBuf<double>& ov, hipStream_t s) {
  if (!packed_layout(n_orb)) return reduce_runs(static_cast<const u64*>(cand), v, n, n_orb, od, ov, s);
  if (n == 0) {
    od.alloc(0, s);
    ov.alloc(0, s);
    return 0;
  }
  DQ_REQUIRE(n < (int64_t(1) << 32), DQ_EINVAL, "too many candidates in one pass");
  const size_t N = static_cast<size_t>(n);
  PoolBuf<u64> sk(N, s);
  PoolBuf<double> sv(N, s);
  size_t tmp = 0;
  const u64* k = static_cast<const u64*>(cand);
  DQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k, sk.p, v, sv.p, N, 0, kc.bits, s));
  {
    PoolBuf<unsigned char> t(tmp + 1, s);
    DQ_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, k, sk.p, v, sv.p, N, 0, kc.bits, s));
  }
  PoolBuf<u32> flag(N, s), pos(N, s);
  flag_heads(sk.p, n, kc, M, flag.p, s);
  const int64_t m = scan_flags(flag.p, pos.p, n, s);
  od.alloc(2 * static_cast<size_t>(m), s);
  ov.alloc(static_cast<size_t>(m), s);
  PoolBuf<u32> hpos(static_cast<size_t>(m) + 1, s);
  run_starts_keys<<<grid_for(n, 256), 256, 0, s>>>(flag.p, pos.p, n, static_cast<u32>(m), hpos.p);
  if (m)  // every candidate may lie in H
    fold_runs_keys<<<grid_for(m, 256), 256, 0, s>>>(sk.p, sv.p, hpos.p, m, kc, M.H.vals != nullptr, od.p, ov.p);
  DQ_HIP(hipGetLastError());
  return m;
}

// Bytes of device memory one candidate costs while its pass is in flight
// (buffer, sorted copy, scan, sort scratch; values for PT2).
size_t cand_bytes(int n_orb, bool vals) {
  const bool pk = packed_layout(n_orb);
  const size_t key = pk ? 8 : 16;
  return 3 * key + (pk ? 8 : 8) + (vals ? 24 : 0);
}

// ------------------------------------------------------------------ unscreened candidates
__device__ __forceinline__ int64_t c2(int x) { return static_cast<int64_t>(x) * (x - 1) / 2; }

// Singles alpha, singles beta, doubles aa, bb, ab (det_ops.hpp:123-268).
__device__ __forceinline__ int64_t conn_total(u64 xa, u64 xb, int n_orb) {
  const u64 full = low_mask(n_orb);
  const int na = popc(xa), nb = popc(xb), va = popc(~xa & full), vb = popc(~xb & full);
  return static_cast<int64_t>(na) * va + static_cast<int64_t>(nb) * vb + c2(na) * c2(va) + c2(nb) * c2(vb) +
         static_cast<int64_t>(na) * va * nb * vb;
}

__global__ void conn_count_kernel(const u64* __restrict__ dets, int64_t n, int n_orb, int64_t* __restrict__ counts,
                                  int* __restrict__ err) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 xa = dets[2 * i], xb = dets[2 * i + 1];
  if ((xa | xb) & ~low_mask(n_orb)) {
    counts[i] = 0;
    atomicOr(err, 1);
    return;
  }
  counts[i] = conn_total(xa, xb, n_orb);
}

// Pair index t -> (p, q), p > q, t = p (p - 1) / 2 + q.
__device__ __forceinline__ void tri_pair(int t, int& p, int& q) {
  p = tri_row(t);
  q = t - p * (p - 1) / 2;
}

// One wavefront per row of [row0, row1): every excitation, written at
// offsets[row] - obase.
__global__ __launch_bounds__(kWave* kWpb) void conn_fill_kernel(const u64* __restrict__ dets, int64_t row0,
                                                                 int64_t row1, int n_orb,
                                                                 const int64_t* __restrict__ offsets, int64_t obase,
                                                                 CandOut out) {
  __shared__ RowLists lists[kWpb];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = row0 + static_cast<int64_t>(blockIdx.x) * kWpb + wid;
  RowLists& L = lists[wid];
  const bool active = row < row1;
  u64 xa = 0, xb = 0;
  if (active) {
    xa = dets[2 * row];
    xb = dets[2 * row + 1];
  }
  build_lists(L, xa, xb, n_orb);
  __syncthreads();
  if (!active || ((xa | xb) & ~low_mask(n_orb))) return;
  const u64 full = low_mask(n_orb);
  const int na = popc(xa), nb = popc(xb), va = popc(~xa & full), vb = popc(~xb & full);
  const int64_t s0 = static_cast<int64_t>(na) * va, s1 = s0 + static_cast<int64_t>(nb) * vb;
  const int64_t pva = c2(va), pvb = c2(vb);
  const int64_t s2 = s1 + c2(na) * pva, s3 = s2 + c2(nb) * pvb;
  const int64_t total = s3 + static_cast<int64_t>(na) * va * nb * vb;
  const int64_t o0 = offsets[row] - obase;
  for (int64_t t = lane; t < total; t += kWave) {
    u64 ka = xa, kb = xb;
    if (t < s0) {
      const int u = static_cast<int>(t);
      ka ^= (1ULL << L.occA[u / va]) | (1ULL << L.virA[u % va]);
    } else if (t < s1) {
      const int u = static_cast<int>(t - s0);
      kb ^= (1ULL << L.occB[u / vb]) | (1ULL << L.virB[u % vb]);
    } else if (t < s2) {
      const int64_t u = t - s1;
      int p, q, r, w;
      tri_pair(static_cast<int>(u / pva), p, q);
      tri_pair(static_cast<int>(u % pva), r, w);
      ka ^= (1ULL << L.occA[p]) | (1ULL << L.occA[q]) | (1ULL << L.virA[r]) | (1ULL << L.virA[w]);
    } else if (t < s3) {
      const int64_t u = t - s2;
      int p, q, r, w;
      tri_pair(static_cast<int>(u / pvb), p, q);
      tri_pair(static_cast<int>(u % pvb), r, w);
      kb ^= (1ULL << L.occB[p]) | (1ULL << L.occB[q]) | (1ULL << L.virB[r]) | (1ULL << L.virB[w]);
    } else {
      int64_t u = t - s3;
      const int pb = static_cast<int>(u % vb);
      u /= vb;
      const int hb = static_cast<int>(u % nb);
      u /= nb;
      const int pa = static_cast<int>(u % va);
      const int ha = static_cast<int>(u / va);
      ka ^= (1ULL << L.occA[ha]) | (1ULL << L.virA[pa]);
      kb ^= (1ULL << L.occB[hb]) | (1ULL << L.virB[pb]);
    }
    out.put(o0 + t, ka, kb);
  }
}

// ------------------------------------------------------------------ screened candidates (walker)
enum SinkMode { kScreen = 0, kPt2 = 1 };

// kScreen: singles with |v| > 1e-15 (and |v psi_i| >= eps when dyn), every
//          double the walker hands over (heat-bath prefix, both added
//          orbitals virtual); no values, no membership test.
// kPt2:    singles and doubles with |v| > 1e-15, |v c_i| >= eps, not in H
//          (an empty H keeps the members of V for the caller to drop);
//          value v c_i.
template <int MODE, bool FILL>
struct CandSink {
  static constexpr bool kNeedKet = true;
  static constexpr bool kCountOnly = false;
  __device__ __forceinline__ void add_doubles(int) {}
  HashView H;
  double ci, eps;
  int dyn;
  int64_t pos;
  CandOut out;
  double* val;
  const ColexRank* T;  // the row's colex tables (out.colex)
  u32 xa, xb;

  __device__ __forceinline__ void put(bool keep, u64 ka, u64 kb, double v) {
    if (MODE == kPt2 && keep) keep = hash_find(H, ka, kb) < 0;  // local_conn.cpp:341,360
    const u64 m = __ballot(keep);
    if (FILL && keep) {
      const int64_t at = pos + __popcll(m & lanemask_lt());
      if (out.colex)
        out.put(at, crank_near(*T, 0, xa, static_cast<u32>(ka)), crank_near(*T, 1, xb, static_cast<u32>(kb)));
      else
        out.put(at, ka, kb);
      if (MODE == kPt2) val[at] = v * ci;
    }
    pos += __popcll(m);
  }
  __device__ __forceinline__ void diag(double) {}
  __device__ __forceinline__ void single(bool ok, u64 ka, u64 kb, double v) {
    bool keep = ok && fabs(v) > kMatThr;
    if (MODE == kPt2 || dyn) keep = keep && fabs(v * ci) >= eps;
    put(keep, ka, kb, v);
  }
  __device__ __forceinline__ void dbl(bool ok, u64 ka, u64 kb, double v) {
    if (MODE == kPt2)
      put(ok && fabs(v) > kMatThr && fabs(v * ci) >= eps, ka, kb, v);
    else
      put(ok, ka, kb, v);
  }
};

// One wavefront per row of [row0, row1).  COUNT (FILL false) writes
// counts[row]; FILL writes at offsets[row] - obase.
//   kScreen: dbl_mode 3, cut = eps (static) or eps / max(|psi_i|, 1e-12)
//            (dynamic, ham_utils.cpp:616-620).
//   kPt2:    rows with c_i^2 < 1e-15 give nothing (local_conn.cpp:303);
//            dbl_mode 2 with cut = eps / max(|c_i|, 1e-12)
//            (local_conn.cpp:366-371) or dbl_mode 0.
template <int MODE, bool FILL>
__global__ __launch_bounds__(kWave* kWpb) void cand_kernel(DevView V, const u64* __restrict__ dets,
                                                            const double* __restrict__ coef, int64_t row0,
                                                            int64_t row1, int n_orb, int use_hb, int dyn, double eps,
                                                            HashView H, int64_t* __restrict__ counts,
                                                            const int64_t* __restrict__ offsets, int64_t obase,
                                                            CandOut out, double* __restrict__ out_val,
                                                            int* __restrict__ err) {
  __shared__ RowLists lists[kWpb];
  __shared__ ColexRank ranks[FILL ? kWpb : 1];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = row0 + static_cast<int64_t>(blockIdx.x) * kWpb + wid;
  RowLists& L = lists[wid];
  const bool active = row < row1;
  u64 xa = 0, xb = 0;
  if (active) {
    xa = dets[2 * row];
    xb = dets[2 * row + 1];
  }
  build_lists(L, xa, xb, n_orb);
  if (FILL && out.colex)
    build_rank_tables<false>(ranks[wid], kBinomDev, static_cast<u32>(xa), static_cast<u32>(xb), n_orb);
  __syncthreads();
  if (!active) return;
  if ((xa | xb) & ~low_mask(n_orb)) {
    if (!FILL && lane == 0) {
      counts[row] = 0;
      atomicOr(err, 1);
    }
    return;
  }
  const double ci = coef ? coef[row] : 0.0;
  if (MODE == kPt2 && ci * ci < 1e-15) {
    if (!FILL && lane == 0) counts[row] = 0;
    return;
  }
  CandSink<MODE, FILL> S;
  S.H = H;
  S.ci = ci;
  S.eps = eps;
  S.dyn = dyn;
  S.pos = FILL ? offsets[row] - obase : 0;
  S.out = out;
  S.val = out_val;
  S.T = &ranks[FILL ? wid : 0];
  S.xa = static_cast<u32>(xa);
  S.xb = static_cast<u32>(xb);
  int dbl_mode;
  double cut;
  if (MODE == kScreen) {
    dbl_mode = 3;
    cut = dyn ? eps / fmax(fabs(ci), 1e-12) : eps;
  } else {
    dbl_mode = use_hb ? 2 : 0;
    cut = eps / fmax(fabs(ci), 1e-12);
  }
  const RowEnv R = make_env(V, xa, xb, n_orb, L);
  walk_row(R, S, dbl_mode, nullptr, cut);
  if (!FILL && lane == 0) counts[row] = S.pos;
}

// ------------------------------------------------------------------ PT2 terms
// Internal term of each row: sigma_i over V (string index, every excitation,
// |H| > 1e-15), r_i = sigma_i - E c_i (local_conn.cpp:296-331).
__global__ __launch_bounds__(kWave* kWpb) void pt2_internal_kernel(DevView V, const u64* __restrict__ dets,
                                                                    const double* __restrict__ coef, int64_t n,
                                                                    int n_orb, double e_ref, StrIndex X, StrCoef C,
                                                                    double* __restrict__ term) {
  __shared__ StrWave lds[kWpb];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kWpb + wid;
  if (row >= n) return;  // whole waves leave; the rest sync per wave only
  StrWave& W = lds[wid];
  const u64 xa = dets[2 * row], xb = dets[2 * row + 1];
  build_lists(W.L, xa, xb, n_orb);
  wave_sync();
  const double ci = coef[row];
  if (((xa | xb) & ~low_mask(n_orb)) || ci * ci < 1e-15) {  // bad rows are reported by the count pass
    if (lane == 0) term[row] = 0.0;
    return;
  }
  // off-diagonal part (the diagonal is weighted by 0 here and added below)
  double a = str_energy_row(V, X, C, xa, xb, n_orb, 0.0, kMatThr, 0.0, W);
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, kWave);
  if (lane == 0) {
    const double h = diag_value(V, xa, xb);
    const double sigma = h * ci + a;
    const double r = sigma - e_ref * ci;
    const double den = clamp_denom(e_ref - h);
    term[row] = (r * r) / den;
  }
}

// External term of each unique a: v_a^2 / (E - H_aa) (local_conn.cpp:397-408).
// LDS form: h_pp, (pp|qq) and (pq|qp) staged once per block (8 (2 n^2 + n)
// bytes), grid-stride over a.
__global__ void pt2_external_terms(DevView V, const u64* __restrict__ d, const double* __restrict__ v, int64_t n,
                                   double e_ref, double* __restrict__ term) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = diag_value(V, d[2 * i], d[2 * i + 1]);
  const double den = clamp_denom(e_ref - h);
  term[i] = (v[i] * v[i]) / den;
}

__global__ __launch_bounds__(256) void pt2_external_terms_lds(DevView V, const u64* __restrict__ d,
                                                               const double* __restrict__ v, int64_t n, double e_ref,
                                                               double* __restrict__ term) {
  extern __shared__ double tab[];
  const int no = V.n, nn = no * no;
  double* hd = tab;
  double* jd = tab + no;
  double* dd = jd + nn;  // J - K
  for (int t = threadIdx.x; t < nn; t += blockDim.x) {
    jd[t] = V.Jd[t];
    dd[t] = V.Jd[t] - V.Kd[t];
  }
  for (int t = threadIdx.x; t < no; t += blockDim.x) hd[t] = V.h1[t * no + t];
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double h = diag_value_jdd(
        d[2 * i], d[2 * i + 1], [&](int p) { return hd[p]; }, [&](int p, int q) { return jd[p * no + q]; },
        [&](int p, int q) { return dd[p * no + q]; });
    const double den = clamp_denom(e_ref - h);
    term[i] = (v[i] * v[i]) / den;
  }
}

// List form for n_orb <= 32 and at most K electrons per spin.  Each
// determinant's occupied orbitals become lists of LDS byte offsets once; every
// sum of diag_value_jdd then runs straight-line over the lists, in the same
// order.  The lists are padded to K with orbital 32, whose table row and
// column hold zeros, so a padded term adds +0.0 and leaves the running sum
// unchanged (the sums start at +0.0 and so are never -0.0).  Tables: J and
// J - K with an odd row stride of 33 (the rows of distinct orbitals fall in
// distinct banks), then the diagonal of h.
constexpr int kTabStride = 33;
constexpr int kTabD = kTabStride * kTabStride, kTabH = 2 * kTabStride * kTabStride;

template <int K>
__global__ __launch_bounds__(256) void pt2_external_terms_list(DevView V, int n_orb, const u64* __restrict__ d,
                                                                const double* __restrict__ v, int64_t n,
                                                                double e_ref, double* __restrict__ term) {
  constexpr int LS = kTabStride;
  __shared__ double tab[kTabH + LS];
  const int no = V.n;
  for (int t = threadIdx.x; t < LS * LS; t += blockDim.x) {
    const int p = t / LS, q = t % LS;
    const bool in = p < n_orb && q < n_orb;
    tab[t] = in ? V.Jd[p * no + q] : 0.0;
    tab[kTabD + t] = in ? V.Jd[p * no + q] - V.Kd[p * no + q] : 0.0;
  }
  for (int t = threadIdx.x; t < LS; t += blockDim.x) tab[kTabH + t] = t < n_orb ? V.h1[t * no + t] : 0.0;
  __syncthreads();
  const char* base = reinterpret_cast<const char*>(tab);
  auto at = [&](u32 off) { return *reinterpret_cast<const double*>(base + off); };
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    u32 xa = static_cast<u32>(d[2 * i]), xb = static_cast<u32>(d[2 * i + 1]);
    u32 ca[K], cb[K], ra[K], rb[K];  // column and row byte offsets (orbital 32 once a string is used up)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const u32 pa = static_cast<u32>(__builtin_ctzll(static_cast<u64>(xa) | (1ull << 32)));
      const u32 pb = static_cast<u32>(__builtin_ctzll(static_cast<u64>(xb) | (1ull << 32)));
      xa &= xa - 1;
      xb &= xb - 1;
      ca[k] = pa * 8;
      cb[k] = pb * 8;
      ra[k] = pa * (LS * 8);
      rb[k] = pb * (LS * 8);
    }
    double e1 = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) e1 += at(kTabH * 8 + ca[k]);
#pragma unroll
    for (int k = 0; k < K; ++k) e1 += at(kTabH * 8 + cb[k]);
    double e2 = 0.0;
#pragma unroll
    for (int a = 1; a < K; ++a)
#pragma unroll
      for (int b = 0; b < a; ++b) e2 += at(kTabD * 8 + ra[a] + ca[b]);
#pragma unroll
    for (int a = 0; a < K; ++a) {
#pragma unroll
      for (int b = 0; b < K; ++b) e2 += at(rb[a] + ca[b]);
#pragma unroll
      for (int b = 0; b < a; ++b) e2 += at(kTabD * 8 + rb[a] + cb[b]);
    }
    const double h = e1 + e2;
    const double den = clamp_denom(e_ref - h);
    term[i] = (v[i] * v[i]) / den;
  }
}

// ------------------------------------------------------------------ driver
// Row chunks [r_k, r_{k+1}) whose candidate counts sum to at most `budget`
// (one row at least), from the host copy of the offsets.
std::vector<int64_t> chunk_rows(const std::vector<int64_t>& off, int64_t budget) {
  const int64_t n = static_cast<int64_t>(off.size()) - 1;
  std::vector<int64_t> cut{0};
  while (cut.back() < n) {
    const int64_t r0 = cut.back();
    const int64_t lim = off[r0] + budget;
    int64_t r1 = static_cast<int64_t>(std::upper_bound(off.begin() + r0 + 1, off.end(), lim) - off.begin()) - 1;
    if (r1 <= r0) r1 = r0 + 1;
    cut.push_back(r1);
  }
  return cut;
}

std::vector<int64_t> host_offsets(const int64_t* d_off, int64_t n, hipStream_t s) {
  std::vector<int64_t> h(static_cast<size_t>(n) + 1);
  DQ_HIP(hipMemcpyAsync(h.data(), d_off, h.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  return h;
}

void check_err(int* d_err, hipStream_t s) {
  int flag = 0;
  DQ_HIP(hipMemcpyAsync(&flag, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  if (flag) throw Error(DQ_EINVAL, "determinant occupies orbitals at or above n_orb");
}

struct Part {
  PoolBuf<u64> d;
  PoolBuf<double> v;
  int64_t n = 0;
};

// Concatenation of the parts (in order) into one buffer.
void concat(std::vector<Part>& parts, bool vals, PoolBuf<u64>& d, PoolBuf<double>& v, int64_t& n, hipStream_t s) {
  n = 0;
  for (auto& p : parts) n += p.n;
  d.alloc(2 * static_cast<size_t>(n), s);
  if (vals) v.alloc(static_cast<size_t>(n), s);
  int64_t at = 0;
  for (auto& p : parts) {
    if (p.n) {
      DQ_HIP(hipMemcpyAsync(d.p + 2 * at, p.d.p, 16 * static_cast<size_t>(p.n), hipMemcpyDeviceToDevice, s));
      if (vals)
        DQ_HIP(hipMemcpyAsync(v.p + at, p.v.p, 8 * static_cast<size_t>(p.n), hipMemcpyDeviceToDevice, s));
    }
    at += p.n;
    p.d.release();
    p.v.release();
  }
}

// A set outlives the stream of the call that made it: after the call's work is
// done its buffers are released on the null stream.
void hand_over(dq_detset& r, hipStream_t s) {
  DQ_HIP(hipStreamSynchronize(s));
  r.dets.s = nullptr;
  r.vals.s = nullptr;
}

// Candidate sets -> sorted unique set minus H.
//   count(counts): per-row candidate counts (all rows);
//   fill(r0, r1, offsets, obase, out): candidates of rows [r0, r1).
//   kc: the key codec the fill writes (colex keys only from cand_kernel).
template <class CountF, class FillF>
void build_set(int64_t n, int n_orb, int64_t budget, const KeyCodec& kc, const Members& M, CountF&& count,
               FillF&& fill, dq_detset& out, hipStream_t s) {
  PoolBuf<int64_t> off(static_cast<size_t>(n) + 1, s);
  count(off.p);
  exclusive_scan_i64(off.p, n, off.p, s);
  const std::vector<int64_t> ho = host_offsets(off.p, n, s);
  const std::vector<int64_t> cut = chunk_rows(ho, budget);
  std::vector<Part> parts;
  for (size_t k = 0; k + 1 < cut.size(); ++k) {
    const int64_t r0 = cut[k], r1 = cut[k + 1], m = ho[r1] - ho[r0];
    PoolBuf<u64> cand((packed_layout(n_orb) ? 1 : 2) * static_cast<size_t>(m), s);
    if (m) fill(r0, r1, off.p, ho[r0], cand_out(cand.p, n_orb, kc));
    DQ_HIP(hipGetLastError());
    Part p;
    p.n = unique_cands(cand.p, m, n_orb, kc, M, p.d, s);
    parts.push_back(std::move(p));
  }
  if (parts.size() == 1) {
    out.dets = std::move(parts[0].d);
    out.n = parts[0].n;
    return;
  }
  PoolBuf<u64> all;
  PoolBuf<double> none;
  int64_t total = 0;
  concat(parts, false, all, none, total, s);
  out.n = unique_minus(all.p, total, n_orb, HashView{}, out.dets, s);
}

// Candidates per pass: as asked, else what a third of the free device memory
// holds, within [2^22, 2^30].
int64_t chunk_budget(int64_t chunk_items, int n_orb, bool vals) {
  if (chunk_items > 0) return chunk_items;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return kDefaultChunk;
  }
  const int64_t fit = static_cast<int64_t>(free_b / 3 / cand_bytes(n_orb, vals));
  return std::min<int64_t>(std::max<int64_t>(fit, int64_t(1) << 22), int64_t(1) << 30);
}

void require_dets(const u64* d_dets, int64_t n, int n_orb) {
  DQ_REQUIRE(n >= 0 && (n == 0 || d_dets), DQ_EINVAL, "bad arguments");
  DQ_REQUIRE(n_orb >= 1 && n_orb <= 64, DQ_EINVAL, "n_orb out of range [1,64]");
  DQ_REQUIRE(n < static_cast<int64_t>(kEmpty), DQ_ERUNTIME, "DetMap size exceeds u32 index capacity");
}

}  // namespace
}  // namespace dq

extern "C" {

int dq_gen_connected(const uint64_t* d_dets, int64_t n, int n_orb, int64_t chunk_items, void* stream,
                     dq_detset** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(out, DQ_EINVAL, "null output");
    dq::require_dets(d_dets, n, n_orb);
    hipStream_t s = dq::as_stream(stream);
    auto r = std::make_unique<dq_detset>();
    if (n > 0) {
      dq::PoolBuf<int> err(1, s);
      DQ_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
      dq::build_set(
          n, n_orb, dq::chunk_budget(chunk_items, n_orb, false), dq::packed_codec(n_orb), dq::Members{},
          [&](int64_t* counts) {
            dq::conn_count_kernel<<<dq::grid_for(n, 256), 256, 0, s>>>(d_dets, n, n_orb, counts, err.p);
            dq::check_err(err.p, s);
          },
          [&](int64_t r0, int64_t r1, const int64_t* off, int64_t obase, dq::CandOut cand) {
            dq::conn_fill_kernel<<<dq::grid_for(r1 - r0, dq::kWpb), dq::kWave * dq::kWpb, 0, s>>>(
                d_dets, r0, r1, n_orb, off, obase, cand);
          },
          *r, s);
    }
    dq::hand_over(*r, s);
    *out = r.release();
  });
}

int dq_gen_perturbative(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, int n_orb, const double* d_psi,
                        int mode, double eps1, int64_t chunk_items, void* stream, dq_detset** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(out, DQ_EINVAL, "null output");
    dq::require_dets(d_dets, n, n_orb);
    hipStream_t s = dq::as_stream(stream);
    // bridge.cpp:285-290: V's DetMap is strict
    dq::DetHash HV;
    if (n > 0) {
      const bool dup = HV.build(d_dets, n, s);
      DQ_REQUIRE(!dup, DQ_EINVAL, "DetMap::from_ordered: duplicate determinant found");
    }
    DQ_REQUIRE(mode == DQ_SCREEN_NONE || mode == DQ_SCREEN_STATIC || mode == DQ_SCREEN_DYNAMIC, DQ_EINVAL,
               "invalid screen mode");
    dq_ctx* c = const_cast<dq_ctx*>(ctx);
    if (mode != DQ_SCREEN_NONE) {
      DQ_REQUIRE(c, DQ_EINVAL, "an integral context is required for screened modes");
      DQ_REQUIRE(c->hb, DQ_EINVAL, "Heat-Bath table not initialized");
      DQ_REQUIRE(n_orb <= c->n_orb, DQ_EINVAL, "n_orb must be in [1, n_orb of the integral context]");
      if (mode == DQ_SCREEN_DYNAMIC) DQ_REQUIRE(n == 0 || d_psi, DQ_EINVAL, "psi_v size mismatch");
    }
    auto r = std::make_unique<dq_detset>();
    if (n > 0) {
      dq::PoolBuf<int> err(1, s);
      DQ_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
      const int64_t budget = dq::chunk_budget(chunk_items, n_orb, false);
      dq::PoolBuf<dq::u32> fv;
      dq::build_filter(d_dets, n, fv, s);
      if (mode == DQ_SCREEN_NONE) {
        dq::build_set(
            n, n_orb, budget, dq::packed_codec(n_orb), dq::Members{HV.view(), fv.p},
            [&](int64_t* counts) {
              dq::conn_count_kernel<<<dq::grid_for(n, 256), 256, 0, s>>>(d_dets, n, n_orb, counts, err.p);
              dq::check_err(err.p, s);
            },
            [&](int64_t r0, int64_t r1, const int64_t* off, int64_t obase, dq::CandOut cand) {
              dq::conn_fill_kernel<<<dq::grid_for(r1 - r0, dq::kWpb), dq::kWave * dq::kWpb, 0, s>>>(
                  d_dets, r0, r1, n_orb, off, obase, cand);
            },
            *r, s);
      } else {
        const dq::DevView V = c->view();
        const int dyn = mode == DQ_SCREEN_DYNAMIC;
        const double* psi = dyn ? d_psi : nullptr;
        constexpr int tb = dq::kWave * dq::kWpb;
        dq::PoolBuf<dq::u32> sa, sb;
        int mm[4];
        dq::electron_range(d_dets, n, mm, s);
        dq::KeyCodec kc = dq::choose_codec(mm, n_orb, sa, sb, s);
        dq::DetHash HR;
        dq::PoolBuf<dq::u32> fr;
        if (kc.sa) {
          dq::rank_map(d_dets, n, HR, fr, s);
          kc.ranked_hash = true;
        }
        dq::build_set(
            n, n_orb, budget, kc, kc.sa ? dq::Members{HR.view(), fr.p} : dq::Members{HV.view(), fv.p},
            [&](int64_t* counts) {
              dq::cand_kernel<dq::kScreen, false><<<dq::grid_for(n, dq::kWpb), tb, 0, s>>>(
                  V, d_dets, psi, 0, n, n_orb, 1, dyn, eps1, dq::HashView{}, counts, nullptr, 0, dq::CandOut{}, nullptr,
                  err.p);
              dq::check_err(err.p, s);
            },
            [&](int64_t r0, int64_t r1, const int64_t* off, int64_t obase, dq::CandOut cand) {
              dq::cand_kernel<dq::kScreen, true><<<dq::grid_for(r1 - r0, dq::kWpb), tb, 0, s>>>(
                  V, d_dets, psi, r0, r1, n_orb, 1, dyn, eps1, dq::HashView{}, nullptr, off, obase, cand, nullptr,
                  err.p);
            },
            *r, s);
      }
    }
    dq::hand_over(*r, s);
    *out = r.release();
  });
}

int dq_compute_pt2(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs, int64_t n, int n_orb,
                   double e_ref, int use_hb, double eps1, int64_t chunk_items, double* h_internal,
                   double* h_external, int64_t* h_n_ext, void* stream, dq_detset** ext_out) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx, DQ_EINVAL, "null context");
    dq::require_dets(d_dets, n, n_orb);
    DQ_REQUIRE(n == 0 || d_coeffs, DQ_EINVAL, "coefficient size mismatch");
    DQ_REQUIRE(n_orb <= ctx->n_orb, DQ_EINVAL, "n_orb must be in [1, n_orb of the integral context]");
    dq_ctx* c = const_cast<dq_ctx*>(ctx);
    if (use_hb) DQ_REQUIRE(c->hb, DQ_EINVAL, "Heat-Bath table not initialized");  // bridge.cpp:735-737
    hipStream_t s = dq::as_stream(stream);
    double e_int = 0.0, e_ext = 0.0;
    auto ext = std::make_unique<dq_detset>();
    ext->has_vals = true;
    if (n > 0) {
      const dq::DevView V = c->view();
      dq::PoolBuf<int> err(1, s);
      DQ_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
      dq::DetHash H;
      H.build(d_dets, n, s);  // tolerant: first occurrence (local_conn.cpp:272-273)
      constexpr int tb = dq::kWave * dq::kWpb;
      // Packed keys leave the members of V to the sorted heads (one probe per
      // unique ket instead of one per candidate in both walks); the pair
      // layout probes in the walks.
      const dq::HashView Hw = dq::packed_layout(n_orb) ? dq::HashView{} : H.view();

      // external: candidates with values, by chunks of rows
      dq::PoolBuf<int64_t> off(static_cast<size_t>(n) + 1, s);
      dq::cand_kernel<dq::kPt2, false><<<dq::grid_for(n, dq::kWpb), tb, 0, s>>>(
          V, d_dets, d_coeffs, 0, n, n_orb, use_hb, 0, eps1, Hw, off.p, nullptr, 0, dq::CandOut{}, nullptr,
          err.p);
      dq::check_err(err.p, s);
      dq::exclusive_scan_i64(off.p, n, off.p, s);
      const std::vector<int64_t> ho = dq::host_offsets(off.p, n, s);
      const std::vector<int64_t> cut = dq::chunk_rows(ho, dq::chunk_budget(chunk_items, n_orb, true));
      dq::PoolBuf<dq::u32> sa, sb;
      int mm[4];
      dq::electron_range(d_dets, n, mm, s);
      dq::KeyCodec kc = dq::choose_codec(mm, n_orb, sa, sb, s);
      dq::DetHash HR;
      dq::PoolBuf<dq::u32> fr, fh;
      if (kc.sa) {
        dq::rank_map(d_dets, n, HR, fr, s);
        kc.ranked_hash = true;
      } else {
        dq::build_filter(d_dets, n, fh, s);
      }
      const dq::Members mem = kc.sa ? dq::Members{HR.view(), fr.p} : dq::Members{H.view(), fh.p};
      std::vector<dq::Part> parts;
      for (size_t k = 0; k + 1 < cut.size(); ++k) {
        const int64_t r0 = cut[k], r1 = cut[k + 1], m = ho[r1] - ho[r0];
        dq::PoolBuf<dq::u64> cd((dq::packed_layout(n_orb) ? 1 : 2) * static_cast<size_t>(m), s);
        dq::PoolBuf<double> cv(static_cast<size_t>(m), s);
        if (m)
          dq::cand_kernel<dq::kPt2, true><<<dq::grid_for(r1 - r0, dq::kWpb), tb, 0, s>>>(
              V, d_dets, d_coeffs, r0, r1, n_orb, use_hb, 0, eps1, Hw, nullptr, off.p, ho[r0],
              dq::cand_out(cd.p, n_orb, kc), cv.p, err.p);
        DQ_HIP(hipGetLastError());
        dq::Part p;
        p.n = dq::reduce_cands(cd.p, cv.p, m, n_orb, kc, mem, p.d, p.v, s);
        parts.push_back(std::move(p));
      }
      if (parts.size() == 1) {
        ext->dets = std::move(parts[0].d);
        ext->vals = std::move(parts[0].v);
        ext->n = parts[0].n;
      } else {
        dq::PoolBuf<dq::u64> all;
        dq::PoolBuf<double> allv;
        int64_t total = 0;
        dq::concat(parts, true, all, allv, total, s);
        ext->n = dq::reduce_runs(all.p, allv.p, total, n_orb, ext->dets, ext->vals, s);
      }
      if (ext->n) {
        dq::PoolBuf<double> term(static_cast<size_t>(ext->n), s);
        const size_t tab = sizeof(double) * (2 * static_cast<size_t>(V.n) * V.n + V.n);
        const int kmax = std::max(mm[1], mm[3]);  // the external dets keep their rows' electron counts
        const unsigned lg = std::min<unsigned>(dq::grid_for(ext->n, 256), 2048);
        if (n_orb <= 32 && kmax <= 8)
          dq::pt2_external_terms_list<8><<<lg, 256, 0, s>>>(V, n_orb, ext->dets.p, ext->vals.p, ext->n, e_ref,
                                                            term.p);
        else if (n_orb <= 32 && kmax <= 16)
          dq::pt2_external_terms_list<16><<<lg, 256, 0, s>>>(V, n_orb, ext->dets.p, ext->vals.p, ext->n, e_ref,
                                                             term.p);
        else if (tab <= 60 * 1024)
          dq::pt2_external_terms_lds<<<std::min<unsigned>(dq::grid_for(ext->n, 256), 2048), 256, tab, s>>>(
              V, ext->dets.p, ext->vals.p, ext->n, e_ref, term.p);
        else
          dq::pt2_external_terms<<<dq::grid_for(ext->n, 256), 256, 0, s>>>(V, ext->dets.p, ext->vals.p, ext->n,
                                                                           e_ref, term.p);
        DQ_HIP(hipGetLastError());
        e_ext = dq::fixed_sum(term.p, ext->n, s);
      }

      // internal: sigma_i through the string index of V
      dq::StrIndexBuf X;
      X.build(d_dets, n, H.view(), s);
      dq::PoolBuf<double> ca(static_cast<size_t>(n), s), cb(static_cast<size_t>(n), s), term(static_cast<size_t>(n), s);
      dq::str_gather_coef(X.view, d_coeffs, ca.p, cb.p, s);
      dq::pt2_internal_kernel<<<dq::grid_for(n, dq::kWpb), tb, 0, s>>>(V, d_dets, d_coeffs, n, n_orb, e_ref, X.view,
                                                                         dq::StrCoef{ca.p, cb.p}, term.p);
      DQ_HIP(hipGetLastError());
      e_int = dq::fixed_sum(term.p, n, s);
    }
    if (h_internal) *h_internal = e_int;
    if (h_external) *h_external = e_ext;
    if (h_n_ext) *h_n_ext = ext->n;
    dq::hand_over(*ext, s);
    if (ext_out) *ext_out = ext.release();
  });
}

int dq_detset_size(const dq_detset* set, int64_t* n, int* has_values) {
  return dq::guarded([&] {
    DQ_REQUIRE(set, DQ_EINVAL, "null set");
    if (n) *n = set->n;
    if (has_values) *has_values = set->has_vals ? 1 : 0;
  });
}

int dq_detset_copy(const dq_detset* set, uint64_t* dets, double* vals, int dst_on_device, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(set, DQ_EINVAL, "null set");
    DQ_REQUIRE(!vals || set->has_vals, DQ_EINVAL, "the set carries no values");
    hipStream_t s = dq::as_stream(stream);
    const hipMemcpyKind k = dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (set->n && dets) DQ_HIP(hipMemcpyAsync(dets, set->dets.p, 16 * static_cast<size_t>(set->n), k, s));
    if (set->n && vals) DQ_HIP(hipMemcpyAsync(vals, set->vals.p, 8 * static_cast<size_t>(set->n), k, s));
    DQ_HIP(hipStreamSynchronize(s));
  });
}

int dq_detset_free(dq_detset* set) {
  delete set;
  return DQ_OK;
}

}  // extern "C"
