// Launch interface shared by slogdet.hip (entry points, n <= 8 kernels, tile
// table) and the slogdet_rows*.hip files (row-distributed kernels, 8 < n <= 64).
#pragma once

#include "dq_common.hpp"

namespace dq {

// One batched slogdet launch.  mode: 0 forward, 1 forward + LU/perm out,
// 2 backward refactorising A, 3 backward from LU/perm.  The element type of A,
// sign, logabs, cot and out_mat is float when f32, double otherwise (modes 1
// and 3 are FP64 only).
struct SlogArgs {
  const void* A;
  const int32_t* perm_in;
  int64_t B;
  int n;
  void* sign;
  void* logabs;
  const void* cot;
  void* out_mat;
  int32_t* perm_out;
  hipStream_t s;
  bool fixed_n = true;  // n equal to the bucket's NC takes the kernel with the bounds fixed
};

// Tile codes: lanes per matrix, plus this flag for the kernel with the bounds
// read at run time also when n equals its bucket's column count.
constexpr int kTileRuntimeN = 0x100;

// Row-distributed kernels by column bucket (n <= NC): `lanes` lanes per
// matrix, each holding ceil(NC / lanes) rows.  Return false when the bucket
// has no kernel for `lanes`.
bool slog_rows_12(int lanes, int mode, bool f32, const SlogArgs& x);
bool slog_rows_16(int lanes, int mode, bool f32, const SlogArgs& x);
bool slog_rows_24(int lanes, int mode, bool f32, const SlogArgs& x);
bool slog_rows_32(int lanes, int mode, bool f32, const SlogArgs& x);
bool slog_rows_64(int lanes, int mode, bool f32, const SlogArgs& x);  // buckets 48 and 64

}  // namespace dq
