// Row-distributed slogdet kernels for 32 < n <= 64 (column buckets 48 and 64); one translation unit per
// column bucket so the instantiations build in parallel.
#include "slogdet_rows.hpp"

namespace dq {

bool slog_rows_64(int lanes, int mode, bool f32, const SlogArgs& x) {
  if (lanes != 64) return false;
  if (x.n <= 48)
    slog::launch_rows_mode<48, 64>(mode, f32, x);
  else
    slog::launch_rows_mode<64, 64>(mode, f32, x);
  return true;
}

}  // namespace dq
