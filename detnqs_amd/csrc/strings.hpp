// String index of a selected space V: its rows grouped by alpha word and by
// beta word, for the V-driven local energy.
//
// The reference evaluates sigma_i = sum_{j in V} H_ij c_j by generating every
// connected determinant of row i and looking it up in V (local_conn.cpp:
// 148-229).  Only determinants of V contribute, and a determinant y of V is
// connected to x iff
//   * y_beta == x_beta and y_alpha differs from x_alpha by one or two
//     orbitals (alpha single, alpha-alpha double),
//   * y_alpha == x_alpha and y_beta differs by one or two orbitals, or
//   * y_alpha and y_beta each differ by one orbital (alpha-beta double).
// With V grouped by alpha word and by beta word, the first two kinds are the
// members of x's beta group and alpha group; the third kind are members of the
// alpha groups of x's alpha single excitations.  Each such y gets the matrix
// element the reference computes for that excitation (same operands, same
// order, same screening), so the set of (y, H_xy) pairs is the reference's;
// only the order of the floating-point sum over them differs.
#pragma once
#include "dq_common.hpp"

namespace dq {

// Device view of one V's string index (all arrays in HBM).  Groups hold only
// the first occurrence of each determinant (the row the DetMap maps it to), so
// every member is a valid partner; later duplicates are collected after the
// last group under a word no lookup can match.  A word with all 64 bits set is
// reserved for that (it needs 64 electrons of one spin).
// A member's coefficient index is not its row but the rank of its row among
// first occurrences: the reference's tolerant DetMap drops later duplicates and
// numbers the rest consecutively (det_space.cpp:40-54), and the energy then
// reads coeffs[that number] (local_conn.cpp:192-194, 322-333), so with a
// duplicate before y the coefficient read for y is another row's.  Without
// duplicates the rank is the row.
struct StrIndex {
  // alpha groups: rows sharing an alpha word; members sorted by row
  const u32* a_off = nullptr;   // group g: [a_off[g], a_off[g+1])
  const u32* a_row = nullptr;   // member's coefficient index (rank, see above)
  const u64* a_word = nullptr;  // member's beta word
  const u64* a_hkey = nullptr;  // hash: alpha word -> group
  const u32* a_hval = nullptr;
  // beta groups: rows sharing a beta word; members sorted by row
  const u32* b_off = nullptr;
  const u32* b_row = nullptr;   // member's coefficient index
  const u64* b_word = nullptr;  // member's alpha word
  const u64* b_hkey = nullptr;
  const u32* b_hval = nullptr;
  u64 hmask = 0;  // both hashes have hmask + 1 slots
  int64_t n = 0;  // members per side (rows of V)
};

// Amplitudes in member order, refreshed whenever psi changes: ca[k] =
// c[a_row[k]], cb[k] = c[b_row[k]], so group scans read them contiguously.
struct StrCoef {
  const double* ca = nullptr;
  const double* cb = nullptr;
};
void str_gather_coef(const StrIndex& X, const double* c, double* ca, double* cb, hipStream_t s);

// Byte sizes of the index of n rows (persistent part) and of the scratch its
// build needs (sort buffers and temporary storage).
size_t str_index_bytes(int64_t n);
size_t str_index_scratch_bytes(int64_t n);

// Builds the index of dets[0..n) into `mem` (str_index_bytes(n)) using
// `scratch` (str_index_scratch_bytes(n)) on stream s; asynchronous.  First
// occurrences come from the DetMap H of the same dets (H maps the row's
// determinant to the row itself).
struct HashView;
// With rank_vals (the DetMap's value array, writable), the DetMap's values are
// replaced by the same ranks afterwards, so lookups through it give the
// coefficient index too.
StrIndex str_index_build(const u64* dets, int64_t n, const HashView& H, void* mem, void* scratch, hipStream_t s,
                         u32* rank_vals = nullptr);
// The same replacement for a DetMap without a string index (temporaries on s).
void hash_to_rank(const u64* dets, int64_t n, const HashView& H, u32* vals, hipStream_t s);

// Owned index (single-instance path and DetMap objects).
// `ordered`: the index is a temporary of a call on stream s (stream-ordered
// release); otherwise it outlives the call.  The build scratch is released in
// stream order either way.
struct StrIndexBuf {
  DevBuf<unsigned char> mem;
  StrIndex view;
  void build(const u64* dets, int64_t n, const HashView& H, hipStream_t s, bool ordered = true,
             u32* rank_vals = nullptr);
};

}  // namespace dq
