// Row-distributed slogdet kernels for 24 < n <= 32; one translation unit per
// column bucket so the instantiations build in parallel.
#include "slogdet_rows.hpp"

namespace dq {

bool slog_rows_32(int lanes, int mode, bool f32, const SlogArgs& x) {
  switch (lanes) {
    case 16: slog::launch_rows_mode<32, 16>(mode, f32, x); return true;
    case 32: slog::launch_rows_mode<32, 32>(mode, f32, x); return true;
    default: return false;
  }
}

}  // namespace dq
