// Lane-ordered H_VV of a selected space (see eloc_strings.hpp): built once per
// V by a count pass and a fill pass of the string visitor, then used for
// energies as a product over the recorded layout, bit-identical to the string
// kernel.  The serving loop builds it in its arena; dq_vvcache owns its own.
#pragma once
#include "batch.hpp"
#include "eloc_strings.hpp"

namespace dq {

// Pass 1: rounds[row] = most pairs any lane of row's wave keeps (rows
// [0, n)); *nnz += pairs (nnz may be NULL).  A row with orbitals at or above
// n_orb sets *err.
void vv_count(const dq_ctx& c, const u64* dets, int64_t n, int n_orb, double sthr, double dmin, const StrIndex& X,
              int64_t* rounds, unsigned long long* nnz, int* err, hipStream_t s);
// Pass 2 (off = exclusive scan of rounds, off[n] = total rounds): fills
// col/val (64 * off[n] slots each), unused slots kVvHole.
void vv_fill(const dq_ctx& c, const u64* dets, int64_t n, int n_orb, double sthr, double dmin, const StrIndex& X,
             const int64_t* off, u32* col, double* val, hipStream_t s);
// sigma_i from the layout (rows with c_i^2 < 1e-12 give 0).
void vv_sigma(const int64_t* off, const u32* col, const double* val, const double* c, int64_t n, double* sigma,
              hipStream_t s);
// sum_i c_i sigma_i in the single-instance order (256-row blocks by a tree,
// block sums in order); part holds (n + 255) / 256 doubles.
void energy_reduce(const double* c, const double* sigma, int64_t n, double* part, double* out, hipStream_t s);

}  // namespace dq

struct dq_vvcache {
  const dq_ctx* ctx = nullptr;
  const dq_detmap* vmap = nullptr;
  int64_t n = 0, rounds = 0, nnz = 0;
  int n_orb = 0, use_hb = 0;
  double eps1 = 0.0;
  dq::DevBuf<int64_t> off;
  dq::DevBuf<dq::u32> col;
  dq::DevBuf<double> val;
};
