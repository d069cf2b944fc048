#pragma once
#include "ctx.hpp"

namespace dq {
void exclusive_scan_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s);
void launch_local_conn(const dq_ctx& c, const DevView& V, const uint64_t* dets, int64_t n, int n_orb, bool use_hb,
                       double eps1, const u32* hb_len, int64_t* counts, const int64_t* offsets,
                       uint64_t* out_dets, double* out_vals, bool fill, hipStream_t s,
                       double sthr_set = -1.0);  // >= 0: single-excitation threshold instead of eps1's
void launch_diag(const DevView& V, const uint64_t* dets, int64_t n, double* out, hipStream_t s);
}  // namespace dq
