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
//   small  (n <= 8) : one matrix per lane, fully register resident; the wave
//                     stages its 64 contiguous matrices through LDS so global
//                     loads/stores are coalesced.
//   group  (n > 8)  : one matrix per group of G lanes (G = 16/32/64), lane r
//                     holds row r; the pivot search is a butterfly over the
//                     group and the pivot row is broadcast with ds_bpermute.
//                     The backward stages L\U in LDS and lets lane c solve for
//                     column c of A^{-T}.
#include <cfloat>

#include "dq_common.hpp"

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

// ------------------------------------------------------------------ groups of lanes
// One round of the pivot search across a 16-lane DPP row: the lane's
// candidate (|a|, position << 8 | lane) against its partner's, larger |a|
// first, then the smaller position.  Candidates are totally ordered
// (positions are distinct), so any all-reduce pattern picks the same row as
// the shuffle butterfly.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ void pivot_round16(double& bc, int& key) {
  const long long b = __double_as_longlong(bc);
  const int lo = dpp_i32<CTRL>(static_cast<int>(b)), hi = dpp_i32<CTRL>(static_cast<int>(b >> 32));
  const double oc = __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
  const int ok = dpp_i32<CTRL>(key);
  if (oc > bc || (oc == bc && (ok >> 8) < (key >> 8))) {
    bc = oc;
    key = ok;
  }
}

// NB: register columns per lane (bucket >= n), G: lanes per matrix (power of two >= NB).
template <int NB, int G, int MODE, class T>
__global__ __launch_bounds__(kWave) void slogdet_group(const T* __restrict__ A, const int32_t* __restrict__ perm_in,
                                                       int64_t B, int n, T* __restrict__ sign, T* __restrict__ logabs,
                                                       const T* __restrict__ cot, T* __restrict__ out_mat,
                                                       int32_t* __restrict__ perm_out) {
  constexpr int GPW = kWave / G;  // matrices per wave
  __shared__ double lu_s[(MODE >= 2) ? GPW * NB * NB : 1];
  __shared__ int perm_s[(MODE >= 2) ? GPW * NB : 1];
  const int lane = threadIdx.x;
  const int g = lane / G, r = lane % G;
  const int64_t m = static_cast<int64_t>(blockIdx.x) * GPW + g;
  const bool valid = m < B;
  const bool row_ok = r < n;
  const size_t nn = static_cast<size_t>(n) * n;
  const T* Am = A + (valid ? m : 0) * nn;

  double a[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) a[j] = (j == r) ? 1.0 : 0.0;
  if (valid && row_ok) {
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < n) a[j] = Am[static_cast<size_t>(r) * n + j];
  }

  int step = -1;
  bool sing = false;
  if (MODE == 3) {
    // rows of L\U already in position order; lane r holds position r
    step = r;
    if (valid && row_ok) sing = (a[r] == 0.0);
    // reduce singular flag over the group
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) sing |= __shfl_xor(static_cast<int>(sing), off, G) != 0;
  } else {
    int pos = r;
    int swaps = 0;
    double s = 1.0, la = 0.0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      if (k < n) {
        double bc = (step < 0 && row_ok) ? fabs(a[k]) : -1.0;
        if (bc != bc) bc = -0.5;  // NaN never wins over a number
        int bp = pos, bl = r;
        if constexpr (G == 16) {
          // xor 1, xor 2 (quad_perm), then mirror within 8 and within 16 lanes
          int key = (bp << 8) | bl;
          pivot_round16<0xB1>(bc, key);
          pivot_round16<0x4E>(bc, key);
          pivot_round16<0x141>(bc, key);
          pivot_round16<0x140>(bc, key);
          bp = key >> 8;
          bl = key & 0xFF;
        } else {
#pragma unroll
          for (int off = G / 2; off > 0; off >>= 1) {
            const double oc = __shfl_xor(bc, off, G);
            const int op = __shfl_xor(bp, off, G);
            const int ol = __shfl_xor(bl, off, G);
            if (oc > bc || (oc == bc && op < bp)) {
              bc = oc;
              bp = op;
              bl = ol;
            }
          }
        }
        if (r == bl) {
          pos = k;
          step = k;
        } else if (pos == k) {
          pos = bp;
        }
        swaps += (bp != k);
        const int src = g * G + bl;
        double prow[NB];
#pragma unroll
        for (int j = k; j < NB; ++j)
          if (j < n) prow[j] = __shfl(a[j], src, kWave);
        const double piv = prow[k];
        slog_acc(piv, s, la, bad);
        if (piv == 0.0) {
          sing = true;
        } else if (step < 0 && row_ok) {
          double l = a[k] * (1.0 / piv);
          if (!(fabs(piv) >= DBL_MIN)) l = a[k] / piv;
          a[k] = l;
#pragma unroll
          for (int j = k + 1; j < NB; ++j)
            if (j < n) a[j] = a[j] - l * prow[j];
        }
      }
    }
    if (MODE == 0 || MODE == 1) {
      if (valid && r == 0) {
        const bool b = bad || sing;
        sign[m] = static_cast<T>(b ? 0.0 : ((swaps & 1) ? -s : s));
        logabs[m] = static_cast<T>(b ? -INFINITY : la);
      }
      if (MODE == 1 && valid && row_ok) {
        T* dst = out_mat + m * nn + static_cast<size_t>(step) * n;
#pragma unroll
        for (int j = 0; j < NB; ++j)
          if (j < n) dst[j] = static_cast<T>(a[j]);
        perm_out[m * n + step] = r;
      }
      return;
    }
  }

  // backward: stage L\U (position order) and the permutation in LDS
  if (MODE == 3) {
    if (row_ok) perm_s[g * NB + r] = valid ? perm_in[m * n + r] : r;
  } else if (row_ok) {
    perm_s[g * NB + step] = r;
  }
  if (row_ok) {
    const int at = (MODE == 3) ? r : step;
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < n) lu_s[(g * NB + at) * NB + j] = a[j];
  }
  __syncthreads();
  if (!valid || !row_ok) return;
  const double* LU = lu_s + g * NB * NB;
  const int c = r;  // this lane solves column c
  const double cc = static_cast<double>(cot[m]);
  double z[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (i < n) {
      double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) s -= LU[k * NB + i] * z[k];
      z[i] = s / LU[i * NB + i];
    }
  }
#pragma unroll
  for (int i = NB - 1; i >= 0; --i) {
    if (i < n) {
      double s = z[i];
#pragma unroll
      for (int k = i + 1; k < NB; ++k)
        if (k < n) s -= LU[k * NB + i] * z[k];
      z[i] = s;
    }
  }
  T* Gm = out_mat + m * nn;
#pragma unroll
  for (int k = 0; k < NB; ++k)
    if (k < n) Gm[static_cast<size_t>(perm_s[g * NB + k]) * n + c] = static_cast<T>(sing ? 0.0 : cc * z[k]);
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

template <int MODE, class T>
void launch_group(int n, const T* A, const int32_t* perm_in, int64_t B, T* sign, T* logabs, const T* cot, T* out_mat,
                  int32_t* perm_out, hipStream_t s) {
  auto go = [&](auto kern, int gpw) {
    const int64_t blocks = (B + gpw - 1) / gpw;
    DQ_REQUIRE(blocks < (1LL << 31), DQ_EINVAL, "batch too large");
    kern<<<dim3(static_cast<unsigned>(blocks)), kWave, 0, s>>>(A, perm_in, B, n, sign, logabs, cot, out_mat, perm_out);
  };
  if (n <= 12)
    go(slogdet_group<12, 16, MODE, T>, 4);
  else if (n <= 16)
    go(slogdet_group<16, 16, MODE, T>, 4);
  else if (n <= 24)
    go(slogdet_group<24, 32, MODE, T>, 2);
  else if (n <= 32)
    go(slogdet_group<32, 32, MODE, T>, 2);
  else if (n <= 48)
    go(slogdet_group<48, 64, MODE, T>, 1);
  else
    go(slogdet_group<64, 64, MODE, T>, 1);
}

template <int MODE, class T>
void dispatch(int n, const T* A, const int32_t* perm_in, int64_t B, T* sign, T* logabs, const T* cot, T* out_mat,
              int32_t* perm_out, hipStream_t s) {
  if (B == 0) return;
  if (n <= 8)
    launch_small<MODE, T>(n, A, perm_in, B, sign, logabs, cot, out_mat, perm_out, s);
  else
    launch_group<MODE, T>(n, A, perm_in, B, sign, logabs, cot, out_mat, perm_out, s);
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

}  // extern "C"
