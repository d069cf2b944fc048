// Lane-ordered H_VV (count pass, fill pass, product) and its C-ABI object.
//
// The reference's driver builds H_VV once per outer step and multiplies it in
// every inner step (operator/kernel.py:48-145 over get_ham_vv); the fused
// energy (local_conn.cpp:148-229) instead regenerates every element per call.
// This layout keeps the elements the energy would generate, in the order the
// string kernel adds them, so a repeated energy on the same V costs one pass
// over 12 bytes per slot instead of the whole string walk, with the same bits.
#include <algorithm>

#include "assemble.hpp"
#include "vvcache.hpp"

namespace dq {

namespace {

constexpr int kVvWaves = 4;

__global__ __launch_bounds__(kWave* kVvWaves) void vv_count_kernel(DevView V, const u64* __restrict__ dets, int64_t n,
                                                                   int n_orb, double sthr, double dmin, StrIndex X,
                                                                   int64_t* __restrict__ rounds,
                                                                   unsigned long long* __restrict__ nnz,
                                                                   int* __restrict__ err) {
  __shared__ StrWave lds[kVvWaves];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kVvWaves + wid;
  if (row >= n) return;  // whole waves leave; the rest sync per wave only
  StrWave& W = lds[wid];
  const u64 xa = dets[2 * row], xb = dets[2 * row + 1];
  build_lists(W.L, xa, xb, n_orb);
  wave_sync();
  if ((xa | xb) & ~low_mask(n_orb)) {
    if (lane == 0) {
      rounds[row] = 0;
      atomicOr(err, 1);
    }
    return;
  }
  VvCountSink S;
  str_visit_row(V, X, xa, xb, n_orb, sthr, dmin, W, S);
  u32 mx = S.n, tot = S.n;
  for (int o = 32; o > 0; o >>= 1) {
    mx = max(mx, static_cast<u32>(__shfl_xor(static_cast<int>(mx), o, kWave)));
    tot += static_cast<u32>(__shfl_xor(static_cast<int>(tot), o, kWave));
  }
  if (lane == 0) {
    rounds[row] = static_cast<int64_t>(mx);
    if (nnz) atomicAdd(nnz, static_cast<unsigned long long>(tot));
  }
}

__global__ __launch_bounds__(kWave* kVvWaves) void vv_fill_kernel(DevView V, const u64* __restrict__ dets, int64_t n,
                                                                  int n_orb, double sthr, double dmin, StrIndex X,
                                                                  const int64_t* __restrict__ off,
                                                                  u32* __restrict__ col, double* __restrict__ val) {
  __shared__ StrWave lds[kVvWaves];
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kVvWaves + wid;
  if (row >= n) return;
  StrWave& W = lds[wid];
  const u64 xa = dets[2 * row], xb = dets[2 * row + 1];
  build_lists(W.L, xa, xb, n_orb);
  wave_sync();
  const int64_t r0 = off[row], r1 = off[row + 1];
  const size_t first = static_cast<size_t>(r0) * kWave + lane;
  VvEmitSink S{X.a_row, X.b_row, col + first, val + first, static_cast<u32>(row), 0};
  if (!((xa | xb) & ~low_mask(n_orb))) str_visit_row(V, X, xa, xb, n_orb, sthr, dmin, W, S);
  // the lane's unused slots (the count pass sized the row for its busiest lane)
  for (int64_t t = r0 + S.n; t < r1; ++t) {
    col[static_cast<size_t>(t) * kWave + lane] = kVvHole;
    val[static_cast<size_t>(t) * kWave + lane] = 0.0;
  }
}

__global__ __launch_bounds__(kWave* kVvWaves) void vv_sigma_kernel(const int64_t* __restrict__ off,
                                                                   const u32* __restrict__ col,
                                                                   const double* __restrict__ val,
                                                                   const double* __restrict__ c, int64_t n,
                                                                   double* __restrict__ sigma) {
  const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kVvWaves + wid;
  if (row >= n) return;
  const double ci = c[row];
  if (ci * ci < 1e-12) {  // local_conn.cpp:190
    if (lane == 0) sigma[row] = 0.0;
    return;
  }
  double a = vv_row_sum(col, val, off[row], off[row + 1], c);
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, kWave);
  if (lane == 0) sigma[row] = a;
}

unsigned grid(int64_t n) { return static_cast<unsigned>((n + kVvWaves - 1) / kVvWaves); }

}  // namespace

void vv_count(const dq_ctx& c, const u64* dets, int64_t n, int n_orb, double sthr, double dmin, const StrIndex& X,
              int64_t* rounds, unsigned long long* nnz, int* err, hipStream_t s) {
  if (n == 0) return;
  vv_count_kernel<<<grid(n), kWave * kVvWaves, 0, s>>>(c.view(), dets, n, n_orb, sthr, dmin, X, rounds, nnz, err);
  DQ_HIP(hipGetLastError());
}

void vv_fill(const dq_ctx& c, const u64* dets, int64_t n, int n_orb, double sthr, double dmin, const StrIndex& X,
             const int64_t* off, u32* col, double* val, hipStream_t s) {
  if (n == 0) return;
  vv_fill_kernel<<<grid(n), kWave * kVvWaves, 0, s>>>(c.view(), dets, n, n_orb, sthr, dmin, X, off, col, val);
  DQ_HIP(hipGetLastError());
}

void vv_sigma(const int64_t* off, const u32* col, const double* val, const double* c, int64_t n, double* sigma,
              hipStream_t s) {
  if (n == 0) return;
  vv_sigma_kernel<<<grid(n), kWave * kVvWaves, 0, s>>>(off, col, val, c, n, sigma);
  DQ_HIP(hipGetLastError());
}

}  // namespace dq

extern "C" {

int dq_vvcache_build(const dq_ctx* ctx, const dq_detmap* vmap, const uint64_t* d_dets, int64_t n, int n_orb,
                     int use_hb, double eps1, void* stream, dq_vvcache** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx && vmap && out, DQ_EINVAL, "null argument");
    DQ_REQUIRE(n >= 0 && (n == 0 || d_dets), DQ_EINVAL, "bad dets");
    DQ_REQUIRE(vmap->n == n, DQ_EINVAL, "DetMap size does not match n_rows");
    DQ_REQUIRE(n < (int64_t{1} << 31), DQ_EINVAL, "the lane-ordered H_VV takes fewer than 2^31 rows");
    DQ_REQUIRE(n_orb >= 1 && n_orb <= ctx->n_orb, DQ_EINVAL, "n_orb must be in [1, n_orb of the integral context]");
    auto* c = const_cast<dq_ctx*>(ctx);
    if (use_hb)
      DQ_REQUIRE(c->hb, DQ_EINVAL, "compute_variational_energy: Heat-Bath mode requires non-null hb_table");
    hipStream_t s = dq::as_stream(stream);
    auto V = std::make_unique<dq_vvcache>();
    V->ctx = ctx;
    V->vmap = vmap;
    V->n = n;
    V->n_orb = n_orb;
    V->use_hb = use_hb;
    V->eps1 = eps1;
    V->off.alloc(static_cast<size_t>(n) + 1);
    double sthr, dmin;
    dq::eloc_thresholds(*c, use_hb, eps1, &sthr, &dmin);
    {
      dq::DevBuf<unsigned long long> cnt(1, s);
      dq::fill_u32(cnt.p, 2, 0u, s);
      dq::vv_count(*c, d_dets, n, n_orb, sthr, dmin, vmap->x.view, V->off.p, cnt.p, c->d_err.p, s);
      dq::exclusive_scan_i64(V->off.p, n, V->off.p, s);
      unsigned long long h_nnz = 0;
      DQ_HIP(hipMemcpyAsync(&h_nnz, cnt.p, sizeof(h_nnz), hipMemcpyDeviceToHost, s));
      DQ_HIP(hipMemcpyAsync(&V->rounds, V->off.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      c->check_rows(s);  // synchronises
      V->nnz = static_cast<int64_t>(h_nnz);
    }
    const size_t slots = static_cast<size_t>(V->rounds) * dq::kWave;
    V->col.alloc(std::max<size_t>(slots, 1));
    V->val.alloc(std::max<size_t>(slots, 1));
    dq::vv_fill(*c, d_dets, n, n_orb, sthr, dmin, vmap->x.view, V->off.p, V->col.p, V->val.p, s);
    DQ_HIP(hipStreamSynchronize(s));
    *out = V.release();
  });
}

int dq_vvcache_info(const dq_vvcache* v, int64_t* n_rows, int64_t* rounds, int64_t* nnz) {
  return dq::guarded([&] {
    DQ_REQUIRE(v, DQ_EINVAL, "null cache");
    if (n_rows) *n_rows = v->n;
    if (rounds) *rounds = v->rounds;
    if (nnz) *nnz = v->nnz;
  });
}

int dq_vvcache_export(const dq_vvcache* v, int64_t* d_off, uint32_t* d_col, double* d_val, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(v, DQ_EINVAL, "null cache");
    hipStream_t s = dq::as_stream(stream);
    const size_t slots = static_cast<size_t>(v->rounds) * dq::kWave;
    if (d_off) DQ_HIP(hipMemcpyAsync(d_off, v->off.p, (v->n + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    if (d_col && slots) DQ_HIP(hipMemcpyAsync(d_col, v->col.p, slots * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    if (d_val && slots) DQ_HIP(hipMemcpyAsync(d_val, v->val.p, slots * sizeof(double), hipMemcpyDeviceToDevice, s));
  });
}

int dq_vvcache_energy(const dq_vvcache* v, const double* d_coeffs, double* d_sigma, double* h_energy,
                      void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(v, DQ_EINVAL, "null cache");
    DQ_REQUIRE(v->n == 0 || d_coeffs, DQ_EINVAL, "null coefficients");
    hipStream_t s = dq::as_stream(stream);
    if (v->n == 0) {
      if (h_energy) *h_energy = 0.0;
      return;
    }
    dq::DevBuf<double> sig_own(0, s);
    double* sig = d_sigma;
    if (!sig) {
      sig_own.alloc(static_cast<size_t>(v->n));
      sig = sig_own.p;
    }
    dq::vv_sigma(v->off.p, v->col.p, v->val.p, d_coeffs, v->n, sig, s);
    if (h_energy) {
      dq::DevBuf<double> part((v->n + 255) / 256 + 1, s), e(1, s);
      dq::energy_reduce(d_coeffs, sig, v->n, part.p, e.p, s);
      DQ_HIP(hipMemcpyAsync(h_energy, e.p, sizeof(double), hipMemcpyDeviceToHost, s));
      DQ_HIP(hipStreamSynchronize(s));
    }
  });
}

int dq_vvcache_free(dq_vvcache* v) {
  return dq::guarded([&] { delete v; });
}

}  // extern "C"
