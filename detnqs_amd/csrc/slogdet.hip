// Batched slogdet forward / backward on gfx950, computed in FP64 (vector FP64,
// no MFMA: n <= 64); FP64 and FP32 inputs.
//
// Semantics follow the reference FFI (csrc/ffi/logdet_cuda.cu):
//   forward : sign = (-1)^{#row swaps} * prod sign(u_kk), logabs = sum log|u_kk|;
//             a zero pivot or a non-finite |u_kk| gives (0, -inf)   (:249-289)
//   backward: grad = cot_logabs * A^{-T}, zero when singular; cot_sign ignored
//             (:293-386, :528)
// The factorisation is LAPACK getf2 (Golub & Van Loan Alg. 3.4.1): at step k the
// pivot is the first row, in the current row order, of maximal |a_rk|;
// multipliers use the reciprocal of the pivot when |pivot| >= DBL_MIN; the
// update is a_rj - l * u_kj without FMA contraction.  A^{-T} solves
// U^T y = e_c, L^T w = y and scatters w through the permutation (getrs 'T').
//
// Two kernel families:
//   small (n <= 8)     : one matrix per lane, fully register resident; the wave
//                        stages its 64 contiguous matrices through LDS so global
//                        loads/stores are coalesced.
//   rows  (8 < n <= 64): a group of G lanes per matrix, each lane holding every
//                        G-th row (slogdet_rows.hpp); G per n comes from a tile
//                        table the autotuner may set (dq_slogdet_tile_set).
#include <atomic>
#include <cfloat>

#include "slogdet.hpp"

namespace dq {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ void slog_acc(double piv, double& s, double& la, bool& bad) {
  const double ad = fabs(piv);
  if (!(ad > 0.0 && isfinite(ad))) bad = true;
  la += log(ad);
  if (piv < 0.0) s = -s;
}

// ------------------------------------------------------------------ small n
template <int N>
__device__ __forceinline__ void lu_small(double (&a)[N][N], int (&perm)[N], int& swaps, bool& sing) {
  for (int i = 0; i < N; ++i) perm[i] = i;
  swaps = 0;
  sing = false;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int p = k;
    double m = fabs(a[k][k]);
#pragma unroll
    for (int r = k + 1; r < N; ++r) {
      const double t = fabs(a[r][k]);
      if (t > m) {
        m = t;
        p = r;
      }
    }
#pragma unroll
    for (int r = k + 1; r < N; ++r) {
      const bool sw = (p == r);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const double t0 = a[k][j], t1 = a[r][j];
        a[k][j] = sw ? t1 : t0;
        a[r][j] = sw ? t0 : t1;
      }
      const int q0 = perm[k], q1 = perm[r];
      perm[k] = sw ? q1 : q0;
      perm[r] = sw ? q0 : q1;
    }
    swaps += (p != k);
    const double piv = a[k][k];
    if (piv == 0.0) {
      sing = true;
      continue;
    }
    const double rp = 1.0 / piv;
    double l[N];
#pragma unroll
    for (int r = k + 1; r < N; ++r) l[r] = a[r][k] * rp;
    if (!(fabs(piv) >= DBL_MIN)) {
#pragma unroll
      for (int r = k + 1; r < N; ++r) l[r] = a[r][k] / piv;
    }
#pragma unroll
    for (int r = k + 1; r < N; ++r) {
      a[r][k] = l[r];
#pragma unroll
      for (int j = k + 1; j < N; ++j) a[r][j] = a[r][j] - l[r] * a[k][j];
    }
  }
}

// mode: 0 = forward, 1 = forward + LU/perm out, 2 = backward from A, 3 = backward from LU/perm
// T: element type of A, the results and cot (double, or float computed in FP64).
template <int N, int MODE, class T>
__global__ __launch_bounds__(kWave) void slogdet_small(const T* __restrict__ A, const int32_t* __restrict__ perm_in,
                                                       int64_t B, T* __restrict__ sign, T* __restrict__ logabs,
                                                       const T* __restrict__ cot, T* __restrict__ out_mat,
                                                       int32_t* __restrict__ perm_out) {
  constexpr int NN = N * N;
  // The forward stages the wave's 64 matrices through LDS in ROUNDS passes of
  // 64 / ROUNDS matrices (coalesced loads, then each lane copies its own
  // matrix), so a wave holds 1/ROUNDS of the LDS and more waves fit on a CU.
  // Two passes measured best (7x7: 0.158 ms per 2M matrices against 0.169
  // with four and 0.196 with eight).  The backward keeps one pass: its output
  // matrix is built in LDS.
  constexpr int ROUNDS = (MODE <= 1 && NN >= 16) ? 2 : 1;
  constexpr int PER = kWave / ROUNDS;
  __shared__ double stage[PER * NN];
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kWave;
  const int nm = static_cast<int>((B - b0) < kWave ? (B - b0) : kWave);
  const int t = threadIdx.x;
  const T* src = A + b0 * NN;
  const bool valid = t < nm;
  double a[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) a[i][j] = (i == j ? 1.0 : 0.0);
  auto copy_own = [&](int m0) {
    if (valid && t >= m0 && t < m0 + PER) {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) a[i][j] = stage[(t - m0) * NN + i * N + j];
    }
  };
  if (nm == kWave) {
    // A full wave: a fixed number of loads per lane and round, and round r + 1
    // loaded before round r is stored and copied, so loads stay in flight.
    constexpr int ITEMS = (PER * NN + kWave - 1) / kWave;
    T buf0[ITEMS], buf1[ITEMS];
    auto load = [&](int r, T (&buf)[ITEMS]) {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const int e = t + j * kWave;
        buf[j] = (e < PER * NN) ? src[static_cast<int64_t>(r * PER) * NN + e] : T(0);
      }
    };
    auto store = [&](const T (&buf)[ITEMS]) {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const int e = t + j * kWave;
        if (e < PER * NN) stage[e] = static_cast<double>(buf[j]);
      }
    };
    load(0, buf0);
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
      if (r + 1 < ROUNDS) load(r + 1, (r & 1) ? buf0 : buf1);
      store((r & 1) ? buf1 : buf0);
      __syncthreads();
      copy_own(r * PER);
      __syncthreads();
    }
  } else {
    for (int r = 0; r < ROUNDS; ++r) {
      const int m0 = r * PER;
      const int cnt = nm - m0 < 0 ? 0 : (nm - m0 < PER ? nm - m0 : PER);
      for (int e = t; e < cnt * NN; e += kWave) stage[e] = src[static_cast<int64_t>(m0) * NN + e];
      __syncthreads();
      copy_own(m0);
      __syncthreads();
    }
  }

  int perm[N];
  int swaps = 0;
  bool sing = false;
  if (MODE == 3) {
#pragma unroll
    for (int i = 0; i < N; ++i) perm[i] = valid ? perm_in[(b0 + t) * N + i] : i;
#pragma unroll
    for (int i = 0; i < N; ++i) sing |= (a[i][i] == 0.0);
  } else {
    lu_small<N>(a, perm, swaps, sing);
  }

  if (MODE == 0 || MODE == 1) {
    double s = (swaps & 1) ? -1.0 : 1.0, la = 0.0;
    bool bad = sing;
#pragma unroll
    for (int k = 0; k < N; ++k) slog_acc(a[k][k], s, la, bad);
    if (valid) {
      sign[b0 + t] = static_cast<T>(bad ? 0.0 : s);
      logabs[b0 + t] = static_cast<T>(bad ? -INFINITY : la);
    }
    if (MODE == 1) {
      if (valid)
#pragma unroll
        for (int i = 0; i < N; ++i) perm_out[(b0 + t) * N + i] = perm[i];
      T* dst = out_mat + b0 * NN;
      for (int r = 0; r < ROUNDS; ++r) {
        const int m0 = r * PER;
        const int cnt = nm - m0 < 0 ? 0 : (nm - m0 < PER ? nm - m0 : PER);
        if (t >= m0 && t < m0 + PER) {
#pragma unroll
          for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) stage[(t - m0) * NN + i * N + j] = a[i][j];
        }
        __syncthreads();
        for (int e = t; e < cnt * NN; e += kWave) dst[static_cast<int64_t>(m0) * NN + e] = static_cast<T>(stage[e]);
        __syncthreads();
      }
    }
    return;
  }

  // backward: column c of X = A^{-T}: U^T y = e_c, L^T w = y, X[perm[k]][c] = w[k]
  const double c0 = valid ? static_cast<double>(cot[b0 + t]) : 0.0;
#pragma unroll
  for (int c = 0; c < N; ++c) {
    double z[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) s -= a[k][i] * z[k];
      z[i] = s / a[i][i];
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
      double s = z[i];
#pragma unroll
      for (int k = i + 1; k < N; ++k) s -= a[k][i] * z[k];
      z[i] = s;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) stage[t * NN + perm[k] * N + c] = sing ? 0.0 : c0 * z[k];
  }
  __syncthreads();
  T* dst = out_mat + b0 * NN;
  for (int e = t; e < nm * NN; e += kWave) dst[e] = static_cast<T>(stage[e]);
}

template <int MODE, class T>
void launch_small(int n, const T* A, const int32_t* perm_in, int64_t B, T* sign, T* logabs, const T* cot, T* out_mat,
                  int32_t* perm_out, hipStream_t s) {
  const int64_t blocks = (B + kWave - 1) / kWave;
  DQ_REQUIRE(blocks < (1LL << 31), DQ_EINVAL, "batch too large");
  const dim3 grid(static_cast<unsigned>(blocks));
#define DQ_SMALL(NV) \
  case NV:           \
    slogdet_small<NV, MODE, T><<<grid, kWave, 0, s>>>(A, perm_in, B, sign, logabs, cot, out_mat, perm_out); \
    break;
  switch (n) {
    DQ_SMALL(2)
    DQ_SMALL(3)
    DQ_SMALL(4)
    DQ_SMALL(5)
    DQ_SMALL(6)
    DQ_SMALL(7)
    DQ_SMALL(8)
    default:
      throw Error(DQ_EINVAL, "internal: bad small n");
  }
#undef DQ_SMALL
}

// Tile code of the row-distributed kernels (lanes per matrix, | kTileRuntimeN
// for the kernel with run-time bounds at n == NC), per kernel kind (the mode:
// 0 forward, 1 forward + LU, 2 backward refactorising, 3 backward from the LU)
// and n; 0 means the default.  Set by dq_slogdet_tile_set (the autotuner);
// read at every launch.
std::atomic<int> g_lanes[4][65];

// Column bucket of the row-distributed kernels for n.
int bucket_of(int n) { return n <= 12 ? 12 : n <= 16 ? 16 : n <= 24 ? 24 : n <= 32 ? 32 : 64; }

// Defaults measured on MI355X (262,144 FP64 matrices per n up to 32, 65,536
// above): per kind {forward, forward + LU, backward from A, backward from
// the LU}.  The forward alone is fastest with few lanes per matrix, the
// backward from the LU with one column per lane; at 25..32 groups of 16
// (two rows per lane) overflow the instruction cache.
int default_lanes(int kind, int n) {
  static const int t12[4] = {4, 4, 4, 16}, t16[4] = {4, 8, 16, 16}, t24[4] = {8, 8, 16, 32};
  int c;
  if (n <= 12)
    c = t12[kind];
  else if (n <= 16)
    c = t16[kind];
  else if (n <= 24)
    c = t24[kind];
  else if (n <= 32)
    c = 32;
  else
    c = 64;
  return n == bucket_of(n) ? c : (c & ~kTileRuntimeN);
}

// Tile codes each column bucket has kernels for (slogdet_rows*.hip): the
// lanes per matrix, and at n equal to an unrolled bucket's column count each
// also with run-time bounds.
int lane_choices(int n, int* out) {
  static const int c12[] = {4, 8, 16}, c24[] = {8, 16, 32}, c32[] = {16, 32}, c64[] = {64};
  const int* c = n <= 16 ? c12 : n <= 24 ? c24 : n <= 32 ? c32 : c64;
  const int k = n <= 16 ? 3 : n <= 24 ? 3 : n <= 32 ? 2 : 1;
  int m = 0;
  for (int i = 0; i < k; ++i) out[m++] = c[i];
  if (n <= 32 && n == bucket_of(n))
    for (int i = 0; i < k; ++i) out[m++] = c[i] | kTileRuntimeN;
  return m;
}

bool launch_rows(int n, int code, int mode, bool f32, SlogArgs x) {
  const int lanes = code & ~kTileRuntimeN;
  x.fixed_n = !(code & kTileRuntimeN);
  if (n <= 12) return slog_rows_12(lanes, mode, f32, x);
  if (n <= 16) return slog_rows_16(lanes, mode, f32, x);
  if (n <= 24) return slog_rows_24(lanes, mode, f32, x);
  if (n <= 32) return slog_rows_32(lanes, mode, f32, x);
  return slog_rows_64(lanes, mode, f32, x);
}

template <int MODE, class T>
void dispatch(int n, const T* A, const int32_t* perm_in, int64_t B, T* sign, T* logabs, const T* cot, T* out_mat,
              int32_t* perm_out, hipStream_t s) {
  if (B == 0) return;
  if (n <= 8) {
    launch_small<MODE, T>(n, A, perm_in, B, sign, logabs, cot, out_mat, perm_out, s);
  } else {
    const int set = g_lanes[MODE][n].load(std::memory_order_relaxed);
    const SlogArgs x{A, perm_in, B, n, sign, logabs, cot, out_mat, perm_out, s, true};
    if (!launch_rows(n, set ? set : default_lanes(MODE, n), MODE, sizeof(T) == 4, x))
      throw Error(DQ_EINVAL, "internal: no slogdet kernel for this tile");
  }
  DQ_HIP(hipGetLastError());
}

void check_args(const void* A, int64_t B, int n) {
  DQ_REQUIRE(n >= 2 && n <= 64, DQ_EINVAL, "N must be in [2, 64]");
  DQ_REQUIRE(B >= 0, DQ_EINVAL, "Batch size must be non-negative");
  DQ_REQUIRE(B == 0 || A, DQ_EINVAL, "null input");
}

}  // namespace
}  // namespace dq

extern "C" {

int dq_slogdet_fwd_f64(const double* d_A, int64_t batch, int n, double* d_sign, double* d_logabs, void* stream) {
  return dq::guarded([&] {
    dq::check_args(d_A, batch, n);
    DQ_REQUIRE(batch == 0 || (d_sign && d_logabs), DQ_EINVAL, "null output");
    dq::dispatch<0, double>(n, d_A, nullptr, batch, d_sign, d_logabs, nullptr, nullptr, nullptr, dq::as_stream(stream));
  });
}

int dq_slogdet_fwd_lu_f64(const double* d_A, int64_t batch, int n, double* d_sign, double* d_logabs, double* d_lu,
                          int32_t* d_perm, void* stream) {
  return dq::guarded([&] {
    dq::check_args(d_A, batch, n);
    DQ_REQUIRE(batch == 0 || (d_sign && d_logabs && d_lu && d_perm), DQ_EINVAL, "null output");
    dq::dispatch<1, double>(n, d_A, nullptr, batch, d_sign, d_logabs, nullptr, d_lu, d_perm, dq::as_stream(stream));
  });
}

int dq_slogdet_bwd_f64(const double* d_A, const double* d_cot_sign, const double* d_cot_logabs, int64_t batch, int n,
                       double* d_grad, void* stream) {
  (void)d_cot_sign;  // sign is not differentiable (logdet_cuda.cu:528)
  return dq::guarded([&] {
    dq::check_args(d_A, batch, n);
    DQ_REQUIRE(batch == 0 || (d_cot_logabs && d_grad), DQ_EINVAL, "null argument");
    dq::dispatch<2, double>(n, d_A, nullptr, batch, nullptr, nullptr, d_cot_logabs, d_grad, nullptr, dq::as_stream(stream));
  });
}

int dq_slogdet_bwd_lu_f64(const double* d_lu, const int32_t* d_perm, const double* d_cot_logabs, int64_t batch, int n,
                          double* d_grad, void* stream) {
  return dq::guarded([&] {
    dq::check_args(d_lu, batch, n);
    DQ_REQUIRE(batch == 0 || (d_perm && d_cot_logabs && d_grad), DQ_EINVAL, "null argument");
    dq::dispatch<3, double>(n, d_lu, d_perm, batch, nullptr, nullptr, d_cot_logabs, d_grad, nullptr, dq::as_stream(stream));
  });
}

// The FP32 targets (logdet_cuda.cu:554-586 instantiate them for float): the
// same kernels read FP32 and compute in FP64; results are rounded to FP32.
int dq_slogdet_fwd_f32(const float* d_A, int64_t batch, int n, float* d_sign, float* d_logabs, void* stream) {
  return dq::guarded([&] {
    dq::check_args(d_A, batch, n);
    DQ_REQUIRE(batch == 0 || (d_sign && d_logabs), DQ_EINVAL, "null output");
    dq::dispatch<0, float>(n, d_A, nullptr, batch, d_sign, d_logabs, nullptr, nullptr, nullptr,
                           dq::as_stream(stream));
  });
}

int dq_slogdet_bwd_f32(const float* d_A, const float* d_cot_sign, const float* d_cot_logabs, int64_t batch, int n,
                       float* d_grad, void* stream) {
  (void)d_cot_sign;
  return dq::guarded([&] {
    dq::check_args(d_A, batch, n);
    DQ_REQUIRE(batch == 0 || (d_cot_logabs && d_grad), DQ_EINVAL, "null argument");
    dq::dispatch<2, float>(n, d_A, nullptr, batch, nullptr, nullptr, d_cot_logabs, d_grad, nullptr,
                           dq::as_stream(stream));
  });
}

// Tile of the n > 8 kernels: lanes per matrix (each lane holds every lanes-th
// row), | 0x100 for the kernel with run-time bounds at n equal to a bucket's
// column count, per kernel kind and n.  Not part of the reference interface;
// the autotuner sets it.
int dq_slogdet_tile_choices(int n, int* lanes, int cap, int* count) {
  return dq::guarded([&] {
    DQ_REQUIRE(n >= 9 && n <= 64, DQ_EINVAL, "tiles exist for 8 < n <= 64");
    DQ_REQUIRE(count && (cap <= 0 || lanes), DQ_EINVAL, "null argument");
    int c[8];
    const int k = dq::lane_choices(n, c);
    for (int i = 0; i < k && i < cap; ++i) lanes[i] = c[i];
    *count = k;
  });
}

int dq_slogdet_tile_set(int kind, int n, int lanes) {
  return dq::guarded([&] {
    DQ_REQUIRE(kind >= 0 && kind <= 3, DQ_EINVAL, "kind must be 0 (forward), 1 (forward + LU), 2 or 3 (backward)");
    DQ_REQUIRE(n >= 9 && n <= 64, DQ_EINVAL, "tiles exist for 8 < n <= 64");
    if (lanes != 0) {
      int c[8];
      const int k = dq::lane_choices(n, c);
      bool ok = false;
      for (int i = 0; i < k; ++i) ok |= (c[i] == lanes);
      DQ_REQUIRE(ok, DQ_EINVAL, "no slogdet kernel with tile code " + std::to_string(lanes) + " at n = " +
                                    std::to_string(n));
    }
    dq::g_lanes[kind][n].store(lanes, std::memory_order_relaxed);
  });
}

int dq_slogdet_tile_get(int kind, int n, int* lanes) {
  return dq::guarded([&] {
    DQ_REQUIRE(kind >= 0 && kind <= 3, DQ_EINVAL, "kind must be 0 (forward), 1 (forward + LU), 2 or 3 (backward)");
    DQ_REQUIRE(n >= 9 && n <= 64, DQ_EINVAL, "tiles exist for 8 < n <= 64");
    DQ_REQUIRE(lanes, DQ_EINVAL, "null argument");
    const int set = dq::g_lanes[kind][n].load(std::memory_order_relaxed);
    *lanes = set ? set : dq::default_lanes(kind, n);
  });
}

}  // extern "C"
