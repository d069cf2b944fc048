#include <hipcub/hipcub.hpp>

#include "sort.hpp"

namespace dq {

namespace {

template <class V>
void sort_pairs(const u64* ki, const V* vi, u64* ko, V* vo, size_t n, int end_bit, hipStream_t s,
                int begin_bit = 0) {
  if (n == 0) return;
  size_t tmp = 0;
  DQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ki, ko, vi, vo, n, begin_bit, end_bit, s));
  DevBuf<unsigned char> t(tmp + 1, s);  // never NULL: a NULL buffer only queries the size
  DQ_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, ki, ko, vi, vo, n, begin_bit, end_bit, s));
}

__global__ void split_dets(const u64* __restrict__ d, size_t n, u64* __restrict__ a, u64* __restrict__ b,
                           u64* __restrict__ idx) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] = d[2 * i];
  b[i] = d[2 * i + 1];
  idx[i] = i;
}

__global__ void gather_u64(const u64* __restrict__ src, const u64* __restrict__ perm, size_t n,
                           u64* __restrict__ dst) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// flag[i] = 1 if sorted det i differs from det i-1
__global__ void mark_heads(const u64* __restrict__ a, const u64* __restrict__ b, const u64* __restrict__ perm,
                           size_t n, u32* __restrict__ flag) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 bi = b[perm[i]];
  flag[i] = (i == 0 || a[i] != a[i - 1] || bi != b[perm[i - 1]]) ? 1u : 0u;
}

__global__ void scatter_unique(const u64* __restrict__ a, const u64* __restrict__ b, const u64* __restrict__ perm,
                               const u32* __restrict__ flag, const u32* __restrict__ pos, size_t n,
                               u64* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const size_t o = pos[i];
  out[2 * o] = a[i];
  out[2 * o + 1] = b[perm[i]];
}

}  // namespace

void sort_pairs_u64_f64(const u64* ki, const double* vi, u64* ko, double* vo, size_t n, int end_bit,
                        hipStream_t s) {
  sort_pairs(ki, vi, ko, vo, n, end_bit, s);
}

void sort_pairs_u64_f64_bits(const u64* ki, const double* vi, u64* ko, double* vo, size_t n, int begin_bit,
                             int end_bit, hipStream_t s) {
  sort_pairs(ki, vi, ko, vo, n, end_bit, s, begin_bit);
}

void sort_pairs_u64_u64(const u64* ki, const u64* vi, u64* ko, u64* vo, size_t n, int end_bit, hipStream_t s) {
  sort_pairs(ki, vi, ko, vo, n, end_bit, s);
}

size_t sort_unique_dets(const u64* dets, size_t n, u64* out, hipStream_t s) {
  if (n == 0) return 0;
  DevBuf<u64> a(n, s), b(n, s), idx(n, s), bs(n, s), p1(n, s), as(n, s), p2(n, s);
  const unsigned g = static_cast<unsigned>((n + 255) / 256);
  split_dets<<<g, 256, 0, s>>>(dets, n, a.p, b.p, idx.p);
  // LSD: stable sort by beta, then by alpha
  sort_pairs(b.p, idx.p, bs.p, p1.p, n, 64, s);
  gather_u64<<<g, 256, 0, s>>>(a.p, p1.p, n, idx.p);  // alpha in beta order
  sort_pairs(idx.p, p1.p, as.p, p2.p, n, 64, s);
  DevBuf<u32> flag(n, s), pos(n, s);
  mark_heads<<<g, 256, 0, s>>>(as.p, b.p, p2.p, n, flag.p);
  size_t tmp = 0;
  DQ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flag.p, pos.p, n, s));
  DevBuf<unsigned char> t(tmp + 1, s);
  DQ_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, flag.p, pos.p, n, s));
  scatter_unique<<<g, 256, 0, s>>>(as.p, b.p, p2.p, flag.p, pos.p, n, out);
  u32 last_pos = 0, last_flag = 0;
  DQ_HIP(hipMemcpyAsync(&last_pos, pos.p + n - 1, sizeof(u32), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipMemcpyAsync(&last_flag, flag.p + n - 1, sizeof(u32), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  return static_cast<size_t>(last_pos) + last_flag;
}

}  // namespace dq
