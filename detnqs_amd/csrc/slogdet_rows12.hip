// Row-distributed slogdet kernels for n <= 12; one translation unit per
// column bucket so the instantiations build in parallel.
#include "slogdet_rows.hpp"

namespace dq {

bool slog_rows_12(int lanes, int mode, bool f32, const SlogArgs& x) {
  switch (lanes) {
    case 4: slog::launch_rows_mode<12, 4>(mode, f32, x); return true;
    case 8: slog::launch_rows_mode<12, 8>(mode, f32, x); return true;
    case 16: slog::launch_rows_mode<12, 16>(mode, f32, x); return true;
    default: return false;
  }
}

}  // namespace dq
