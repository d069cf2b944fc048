// DetMap on the device (det_space.cpp:29-65): open-addressing hash in HBM
// with 128-bit keys (alpha word, beta word), u32 row indices, load factor
// <= 1/2 and linear probing.  Built by DetHash::build (project.hip): slots are
// claimed by a CAS on the value word; a second pass lowers each key's value to
// the smallest row index holding it (first occurrence, as from_ordered keeps)
// and reports whether duplicates were seen.
#pragma once
#include "dq_common.hpp"

namespace dq {

constexpr u32 kEmpty = 0xFFFFFFFFu;

__host__ __device__ __forceinline__ u64 mix64(u64 x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBULL;
  x ^= x >> 31;
  return x;
}
__device__ __forceinline__ u64 det_hash(u64 a, u64 b) { return mix64(a ^ mix64(b + 0x9E3779B97F4A7C15ULL)); }

struct HashView {
  const u64* keys = nullptr;
  const u32* vals = nullptr;
  u64 mask = 0;
};

// Row index of (a, b) in V, or -1.  An empty map (vals == nullptr) finds nothing.
__device__ __forceinline__ long hash_find(const HashView& H, u64 a, u64 b) {
  if (!H.vals) return -1;
  u64 s = det_hash(a, b) & H.mask;
  while (true) {
    const u32 v = H.vals[s];
    if (v == kEmpty) return -1;
    if (H.keys[2 * s] == a && H.keys[2 * s + 1] == b) return static_cast<long>(v);
    s = (s + 1) & H.mask;
  }
}

struct DetHash {
  DevBuf<u64> keys;
  DevBuf<u32> vals;
  u64 mask = 0;
  HashView view() const { return HashView{keys.p, vals.p, mask}; }
  // Builds the map of dets[0..n) on stream s and synchronises it; returns true
  // if some determinant occurs more than once.  `ordered`: the map is a
  // temporary of a call on s (stream-ordered release); otherwise it outlives
  // the call.
  bool build(const u64* dets, int64_t n, hipStream_t s, bool ordered = true);
};

// Slots for n keys: the smallest power of two >= max(16, 2n).
u64 hash_capacity(int64_t n);
// Builds the map of dets[0..n) into caller buffers (mask + 1 slots, keys of
// 2 (mask + 1) words) on stream s, without the duplicate report; asynchronous.
void hash_build_async(const u64* dets, int64_t n, u64* keys, u32* vals, u64 mask, hipStream_t s);

// Membership prefilter of a map's keys: bit det_hash(a, b) >> (64 - lg) of a
// 2^lg-bit set.  A clear bit proves (a, b) absent, so a kernel tests it (in
// LDS or a cache) before the probe of the map in HBM; a set bit decides
// nothing.
__device__ __forceinline__ u32 filt_slot(u64 a, u64 b, int lg) { return static_cast<u32>(det_hash(a, b) >> (64 - lg)); }
__device__ __forceinline__ bool filt_test(const u32* bits, u32 h) { return (bits[h >> 5] >> (h & 31)) & 1u; }
// Sets the bits of the pairs[0..n) into bits[2^lg / 32] (cleared first) on s.
void filter_build(const u64* pairs, int64_t n, int lg, u32* bits, hipStream_t s);

}  // namespace dq
