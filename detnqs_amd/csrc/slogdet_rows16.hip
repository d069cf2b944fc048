// Row-distributed slogdet kernels for 12 < n <= 16; one translation unit per
// column bucket so the instantiations build in parallel.
#include "slogdet_rows.hpp"

namespace dq {

bool slog_rows_16(int lanes, int mode, bool f32, const SlogArgs& x) {
  switch (lanes) {
    case 4: slog::launch_rows_mode<16, 4>(mode, f32, x); return true;
    case 8: slog::launch_rows_mode<16, 8>(mode, f32, x); return true;
    case 16: slog::launch_rows_mode<16, 16>(mode, f32, x); return true;
    default: return false;
  }
}

}  // namespace dq
