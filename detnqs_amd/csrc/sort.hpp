// Device radix sorts (hipCUB) used to canonicalise COO output and P spaces.
#pragma once
#include "dq_common.hpp"

namespace dq {
// Stable sort of (key, value) by key over [0, end_bit) bits.
void sort_pairs_u64_f64(const u64* keys_in, const double* vals_in, u64* keys_out, double* vals_out, size_t n,
                        int end_bit, hipStream_t s);
// Stable sort by bits [begin_bit, end_bit) of the key only.
void sort_pairs_u64_f64_bits(const u64* keys_in, const double* vals_in, u64* keys_out, double* vals_out, size_t n,
                             int begin_bit, int end_bit, hipStream_t s);
void sort_pairs_u64_u64(const u64* keys_in, const u64* vals_in, u64* keys_out, u64* vals_out, size_t n,
                        int end_bit, hipStream_t s);
// Sorted unique 128-bit keys (alpha, beta) ordered by alpha then beta
// (det.hpp:30-36).  Returns the number of unique keys written to out.
size_t sort_unique_dets(const u64* dets, size_t n, u64* out, hipStream_t s);
}  // namespace dq
