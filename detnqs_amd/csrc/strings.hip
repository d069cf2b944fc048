// Build of the string index (alpha and beta groups of V) on the device: a
// stable radix sort of (word, row) pairs, group heads by adjacent difference,
// group ids by an inclusive scan, a word -> group hash.  Deterministic: members
// of a group are in row order whatever the launch.
#include <hipcub/hipcub.hpp>

#include "eloc_strings.hpp"
#include "hash.hpp"
#include "strings.hpp"

namespace dq {

namespace {

constexpr size_t kAl = 256;
size_t al(size_t x) { return (x + kAl - 1) / kAl * kAl; }

struct Carve {
  unsigned char* p;
  template <class T>
  T* take(size_t count) {
    T* r = reinterpret_cast<T*>(p);
    p += al(count * sizeof(T));
    return r;
  }
};

size_t radix_temp(int64_t n) {
  size_t t = 0;
  DQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t, static_cast<const u64*>(nullptr),
                                            static_cast<u64*>(nullptr), static_cast<const u32*>(nullptr),
                                            static_cast<u32*>(nullptr), static_cast<int>(n), 0, 64));
  return t;
}

size_t scan_temp(int64_t n) {
  size_t t = 0, e = 0;
  DQ_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, t, static_cast<const u32*>(nullptr), static_cast<u32*>(nullptr),
                                          static_cast<int>(n)));
  DQ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, e, static_cast<const u32*>(nullptr), static_cast<u32*>(nullptr),
                                          static_cast<int>(n)));
  return std::max(t, e);
}

// Later duplicates get the reserved word, so they sort after every group.
__global__ void split_word(const u64* __restrict__ dets, const unsigned char* __restrict__ first, int64_t n, int w,
                           u64* __restrict__ key, u32* __restrict__ val) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = first[i] ? dets[2 * i + w] : ~0ULL;
  val[i] = static_cast<u32>(i);
}

__global__ void mark_heads(const u64* __restrict__ key, int64_t n, u32* __restrict__ head) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  head[k] = (k == 0 || key[k] != key[k - 1]) ? 1u : 0u;
}

// flag[i] = first[i] (u32, for the rank scan)
__global__ void first_u32(const unsigned char* __restrict__ first, int64_t n, u32* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = first[i];
}

// Every occupied DetMap slot: row of the first occurrence -> its rank.
__global__ void vals_to_rank(u32* __restrict__ vals, u64 cap, const u32* __restrict__ rank) {
  const u64 s = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= cap) return;
  const u32 v = vals[s];
  if (v != kEmpty) vals[s] = rank[v];
}

__global__ void fill_groups(const u64* __restrict__ dets, const u64* __restrict__ key, const u32* __restrict__ rows,
                            const u32* __restrict__ rank,
                            const u32* __restrict__ gid1, int64_t n, int w, u32* __restrict__ off,
                            u32* __restrict__ mrow, u64* __restrict__ mword, u64* __restrict__ hkey,
                            u32* __restrict__ hval, u64 mask) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 g = gid1[k] - 1;
  const u32 row = rows[k];
  mrow[k] = rank[row];
  mword[k] = dets[2 * static_cast<size_t>(row) + (1 - w)];
  if (k == n - 1) off[g + 1] = static_cast<u32>(n);
  if (k == 0 || key[k] != key[k - 1]) {
    off[g] = static_cast<u32>(k);
    const u64 wd = key[k];
    if (wd == ~0ULL) return;  // the duplicates' group is never looked up
    u64 s = word_hash(wd) & mask;
    while (true) {
      if (atomicCAS(reinterpret_cast<unsigned long long*>(&hkey[s]), ~0ULL, static_cast<unsigned long long>(wd)) ==
          ~0ULL) {
        hval[s] = g;
        break;
      }
      s = (s + 1) & mask;
    }
  }
}

__global__ void mark_first(const u64* __restrict__ dets, int64_t n, HashView H, unsigned char* __restrict__ first) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  first[i] = hash_find(H, dets[2 * i], dets[2 * i + 1]) == static_cast<long>(i) ? 1 : 0;
}

__global__ void gather_coef(const u32* __restrict__ a_row, const u32* __restrict__ b_row, const double* __restrict__ c,
                            int64_t n, double* __restrict__ ca, double* __restrict__ cb) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  ca[k] = c[a_row[k]];
  cb[k] = c[b_row[k]];
}

}  // namespace

size_t str_index_bytes(int64_t n) {
  const size_t cap = hash_capacity(n);
  const size_t un = static_cast<size_t>(n);
  const size_t side = al(4 * (un + 1)) + al(4 * un) + al(8 * un) + al(8 * cap) + al(4 * cap);
  return 2 * side + al(un);
}

size_t str_index_scratch_bytes(int64_t n) {
  const size_t un = static_cast<size_t>(std::max<int64_t>(n, 1));
  return 2 * al(8 * un) + 4 * al(4 * un) + al(std::max(radix_temp(n), scan_temp(n)));
}

StrIndex str_index_build(const u64* dets, int64_t n, const HashView& H, void* mem, void* scratch, hipStream_t s,
                         u32* rank_vals) {
  DQ_REQUIRE(n < (int64_t{1} << 31), DQ_ERUNTIME, "string index supports fewer than 2^31 rows");
  const size_t cap = hash_capacity(n);
  const size_t un = static_cast<size_t>(n);
  Carve m{static_cast<unsigned char*>(mem)};
  StrIndex X;
  u32* a_off = m.take<u32>(un + 1);
  u32* a_row = m.take<u32>(un);
  u64* a_word = m.take<u64>(un);
  u64* a_hkey = m.take<u64>(cap);
  u32* a_hval = m.take<u32>(cap);
  u32* b_off = m.take<u32>(un + 1);
  u32* b_row = m.take<u32>(un);
  u64* b_word = m.take<u64>(un);
  u64* b_hkey = m.take<u64>(cap);
  u32* b_hval = m.take<u32>(cap);
  unsigned char* first = m.take<unsigned char>(un);
  X.n = n;
  X.a_off = a_off;
  X.a_row = a_row;
  X.a_word = a_word;
  X.a_hkey = a_hkey;
  X.a_hval = a_hval;
  X.b_off = b_off;
  X.b_row = b_row;
  X.b_word = b_word;
  X.b_hkey = b_hkey;
  X.b_hval = b_hval;
  X.hmask = cap - 1;
  fill_u32(a_hkey, 2 * cap, 0xFFFFFFFFu, s);
  fill_u32(b_hkey, 2 * cap, 0xFFFFFFFFu, s);
  fill_u32(a_off, 1, 0u, s);
  fill_u32(b_off, 1, 0u, s);
  if (n == 0) return X;
  const size_t sn = std::max<size_t>(un, 1);
  Carve w{static_cast<unsigned char*>(scratch)};
  u64* kin = w.take<u64>(sn);
  u64* kout = w.take<u64>(sn);
  u32* vin = w.take<u32>(sn);
  u32* vout = w.take<u32>(sn);
  u32* head = w.take<u32>(sn);
  u32* rank = w.take<u32>(sn);
  void* tmp = w.p;
  const size_t rt = radix_temp(n), st = scan_temp(n);
  const unsigned g = static_cast<unsigned>((n + 255) / 256);
  mark_first<<<g, 256, 0, s>>>(dets, n, H, first);
  first_u32<<<g, 256, 0, s>>>(first, n, head);
  {
    size_t t = st;
    DQ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, t, head, rank, static_cast<int>(n), s));
  }
  for (int wd = 0; wd < 2; ++wd) {
    split_word<<<g, 256, 0, s>>>(dets, first, n, wd, kin, vin);
    size_t t = rt;
    DQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, t, kin, kout, vin, vout, static_cast<int>(n), 0, 64, s));
    mark_heads<<<g, 256, 0, s>>>(kout, n, head);
    t = st;
    DQ_HIP(hipcub::DeviceScan::InclusiveSum(tmp, t, head, vin, static_cast<int>(n), s));
    if (wd == 0)
      fill_groups<<<g, 256, 0, s>>>(dets, kout, vout, rank, vin, n, 0, a_off, a_row, a_word, a_hkey, a_hval,
                                    cap - 1);
    else
      fill_groups<<<g, 256, 0, s>>>(dets, kout, vout, rank, vin, n, 1, b_off, b_row, b_word, b_hkey, b_hval,
                                    cap - 1);
  }
  if (rank_vals) {
    const u64 hcap = H.mask + 1;
    vals_to_rank<<<static_cast<unsigned>((hcap + 255) / 256), 256, 0, s>>>(rank_vals, hcap, rank);
  }
  DQ_HIP(hipGetLastError());
  return X;
}

void hash_to_rank(const u64* dets, int64_t n, const HashView& H, u32* vals, hipStream_t s) {
  if (n == 0) return;
  const size_t un = static_cast<size_t>(n);
  DevBuf<unsigned char> first(un, s);
  DevBuf<u32> flag(un, s), rank(un, s);
  const unsigned g = static_cast<unsigned>((n + 255) / 256);
  mark_first<<<g, 256, 0, s>>>(dets, n, H, first.p);
  first_u32<<<g, 256, 0, s>>>(first.p, n, flag.p);
  size_t t = 0;
  DQ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t, flag.p, rank.p, static_cast<int>(n), s));
  DevBuf<unsigned char> tmp(t + 1, s);
  DQ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, t, flag.p, rank.p, static_cast<int>(n), s));
  const u64 cap = H.mask + 1;
  vals_to_rank<<<static_cast<unsigned>((cap + 255) / 256), 256, 0, s>>>(vals, cap, rank.p);
  DQ_HIP(hipGetLastError());
}

void str_gather_coef(const StrIndex& X, const double* c, double* ca, double* cb, hipStream_t s) {
  if (X.n == 0) return;
  gather_coef<<<static_cast<unsigned>((X.n + 255) / 256), 256, 0, s>>>(X.a_row, X.b_row, c, X.n, ca, cb);
  DQ_HIP(hipGetLastError());
}

void StrIndexBuf::build(const u64* dets, int64_t n, const HashView& H, hipStream_t s, bool ordered, u32* rank_vals) {
  if (ordered)
    mem.alloc(str_index_bytes(n), s);
  else
    mem.alloc(str_index_bytes(n));
  DevBuf<unsigned char> scratch(str_index_scratch_bytes(n), s);
  view = str_index_build(dets, n, H, mem.p, scratch.p, s, rank_vals);
}

}  // namespace dq
