// Device Hamiltonian operator and energy step: the consumers of the projected
// blocks (operator/kernel.py:48-145, operator/functional.py:31-268), which the
// reference runs as SciPy CSR/CSC products inside a host callback.
//
//  * Blocks arrive as COO (H_VV n_v x n_v, H_VP n_v x n_p); they are stably
//    sorted into CSR for H_VV and H_VP and into CSR of H_PV = H_VP^T (the
//    reference's CSC of H_PV, column order = V order).
//  * Products: one thread per row, a sequential sum in column order starting
//    from 0, which is SciPy's csr_matvec / csc_matvec order, so every y entry
//    matches SciPy bit for bit; y_V = (H_VV x_V) + (H_VP x_P) and
//    y_P = (H_PV x_V) + diag(H_PP) * x_P as two sums then one add, as the
//    reference writes them.
//  * Energy step (real wavefunctions): psi = sign * exp(logabs - max logabs);
//    variational E = <psi_V|H_VV psi_V>/max(<psi_V|psi_V>, eps); proxy on
//    T = V + P; asymmetric E = <psi_V|(H psi)_V>/max(<psi_V|psi_V>, eps);
//    g = (H psi - E psi)/den and cot_logabs = 2 psi g (zero on P for the
//    asymmetric mode).  Dot products use a fixed order (256-element trees,
//    block sums in order), so results are reproducible run to run.
#include <hipcub/hipcub.hpp>

#include <memory>
#include <vector>

#include "assemble.hpp"
#include "dq_common.hpp"
#include "sort.hpp"

struct dq_operator {
  int64_t n_v = 0, n_p = 0, nnz_vv = 0, nnz_vp = 0;
  dq::DevBuf<int64_t> vv_ptr, vp_ptr, pv_ptr;
  dq::DevBuf<dq::u32> vv_col, vp_col, pv_col;
  dq::DevBuf<double> vv_val, vp_val, pv_val, h_pp;
  dq::DevBuf<double> work;  // n_v + n_p products and partial sums
};

namespace dq {
namespace {

constexpr int kT = 256;

// key = major << cb | minor, cb = bits of the minor extent, so the sort needs
// only bits(major extent) + cb bits.
__global__ void coo_keys(const u32* __restrict__ r, const u32* __restrict__ c, int64_t n, int swap, u64 nr, u64 nc,
                         int cb, u64* __restrict__ key, u64* __restrict__ idx, int* __restrict__ err) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (r[i] >= nr || c[i] >= nc) atomicOr(err, 1);
  key[i] = swap ? ((static_cast<u64>(c[i]) << cb) | r[i]) : ((static_cast<u64>(r[i]) << cb) | c[i]);
  idx[i] = static_cast<u64>(i);
}

__global__ void csr_fill(const u64* __restrict__ key, const u64* __restrict__ idx, const double* __restrict__ val_in,
                         int64_t n, int cb, u32* __restrict__ col, double* __restrict__ val,
                         unsigned long long* __restrict__ cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  col[i] = static_cast<u32>(key[i] & ((1ull << cb) - 1));
  val[i] = val_in[idx[i]];
  atomicAdd(&cnt[key[i] >> cb], 1ULL);
}

// Bits needed to hold values 0 .. extent - 1 (at least 1).
int bits_for(u64 extent) {
  int b = 1;
  while (b < 64 && (extent > (1ull << b))) ++b;
  return b;
}

// y[i] = sum_k val[k] x[col[k]] over row i, in column order, from 0.
__global__ void csr_matvec(const int64_t* __restrict__ ptr, const u32* __restrict__ col, const double* __restrict__ val,
                           const double* __restrict__ x, int64_t nrows, double* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  double s = 0.0;
  for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
  y[i] = s;
}

__global__ void add_inplace(double* __restrict__ y, const double* __restrict__ z, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + z[i];
}

__global__ void add_diag(double* __restrict__ y, const double* __restrict__ d, const double* __restrict__ x,
                         int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + d[i] * x[i];
}

// Block maxima of logabs, then a serial max over them (max is order-free).
__global__ void block_max(const double* __restrict__ x, int64_t n, double* __restrict__ part) {
  __shared__ double buf[kT];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
  buf[threadIdx.x] = i < n ? x[i] : -INFINITY;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) buf[threadIdx.x] = fmax(buf[threadIdx.x], buf[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = buf[0];
}

__global__ void serial_max(const double* __restrict__ part, int64_t m, double* __restrict__ out) {
  double t = -INFINITY;
  for (int64_t i = 0; i < m; ++i) t = fmax(t, part[i]);
  *out = m > 0 ? t : 0.0;
}

__global__ void decode_psi(const double* __restrict__ sign, const double* __restrict__ logabs, int64_t n,
                           const double* __restrict__ shift, double* __restrict__ psi) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) psi[i] = sign[i] * exp(logabs[i] - *shift);
}

// Block sums of a[i] * b[i] (tree of 256), for the fixed-order dots.
__global__ void block_dot(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                          double* __restrict__ part) {
  __shared__ double buf[kT];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
  buf[threadIdx.x] = i < n ? a[i] * b[i] : 0.0;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) buf[threadIdx.x] += buf[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = buf[0];
}

// E and den from the four block-sum lists; out = {E, den_safe, num, den}.
__global__ void finish_energy(const double* __restrict__ p_num_v, int64_t m_v, const double* __restrict__ p_num_p,
                              int64_t m_p, const double* __restrict__ p_den_v, const double* __restrict__ p_den_p,
                              int use_p_num, int use_p_den, double eps, double* __restrict__ out) {
  double nv = 0.0, np = 0.0, dv = 0.0, dp = 0.0;
  for (int64_t i = 0; i < m_v; ++i) nv += p_num_v[i];
  for (int64_t i = 0; i < m_v; ++i) dv += p_den_v[i];
  if (use_p_num)
    for (int64_t i = 0; i < m_p; ++i) np += p_num_p[i];
  if (use_p_den)
    for (int64_t i = 0; i < m_p; ++i) dp += p_den_p[i];
  const double num = use_p_num ? nv + np : nv;
  const double den = use_p_den ? dv + dp : dv;
  const double den_safe = fmax(den, eps);
  out[0] = num / den_safe;
  out[1] = den_safe;
  out[2] = num;
  out[3] = den;
}

// cot_logabs = 2 psi g, g = (y - E psi) / den_safe; zero where `zero` is set.
__global__ void cotangent(const double* __restrict__ y, const double* __restrict__ psi, int64_t n,
                          const double* __restrict__ e, int zero, double* __restrict__ cot) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (zero) {
    cot[i] = 0.0;
    return;
  }
  const double g = (y[i] - e[0] * psi[i]) / e[1];
  cot[i] = 2.0 * psi[i] * g;
}

unsigned nb(int64_t n) { return static_cast<unsigned>((n + kT - 1) / kT); }

void build_csr(const u32* r, const u32* c, const double* v, int64_t nnz, int swap, int64_t nrows, int64_t ncols,
               DevBuf<int64_t>& ptr, DevBuf<u32>& col, DevBuf<double>& val, hipStream_t s) {
  ptr.alloc(nrows + 1);
  col.alloc(std::max<int64_t>(nnz, 1));
  val.alloc(std::max<int64_t>(nnz, 1));
  DevBuf<unsigned long long> cnt(nrows + 1, s);
  DQ_HIP(hipMemsetAsync(cnt.p, 0, (nrows + 1) * sizeof(unsigned long long), s));
  if (nnz > 0) {
    DevBuf<u64> k(nnz, s), ko(nnz, s), ix(nnz, s), ixo(nnz, s);
    DevBuf<int> err(1, s);
    DQ_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    // the key's major part is the CSR row (< nrows), its minor part the CSR
    // column (< ncols), whichever COO array each comes from
    const int cb = bits_for(static_cast<u64>(ncols)), kb = cb + bits_for(static_cast<u64>(nrows));
    coo_keys<<<nb(nnz), kT, 0, s>>>(r, c, nnz, swap, static_cast<u64>(swap ? ncols : nrows),
                                    static_cast<u64>(swap ? nrows : ncols), cb, k.p, ix.p, err.p);
    DQ_HIP(hipGetLastError());
    int h = 0;
    DQ_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    DQ_REQUIRE(!h, DQ_ERANGE, "COO index out of range for the block shape");
    sort_pairs_u64_u64(k.p, ix.p, ko.p, ixo.p, static_cast<size_t>(nnz), kb, s);
    csr_fill<<<nb(nnz), kT, 0, s>>>(ko.p, ixo.p, v, nnz, cb, col.p, val.p, cnt.p);
    DQ_HIP(hipGetLastError());
  }
  // exclusive scan of the row counts into ptr (64-bit item count)
  exclusive_scan_i64(reinterpret_cast<const int64_t*>(cnt.p), nrows, ptr.p, s);
}

void apply(const dq_operator& op, const double* xv, const double* xp, double* yv, double* yp, hipStream_t s) {
  const int64_t nv = op.n_v, np = op.n_p;
  if (nv > 0) csr_matvec<<<nb(nv), kT, 0, s>>>(op.vv_ptr.p, op.vv_col.p, op.vv_val.p, xv, nv, yv);
  if (np > 0 && xp) {
    double* tmp = op.work.p;  // H_VP x_P, then added
    if (nv > 0) {
      csr_matvec<<<nb(nv), kT, 0, s>>>(op.vp_ptr.p, op.vp_col.p, op.vp_val.p, xp, nv, tmp);
      add_inplace<<<nb(nv), kT, 0, s>>>(yv, tmp, nv);
    }
    if (yp) {
      csr_matvec<<<nb(np), kT, 0, s>>>(op.pv_ptr.p, op.pv_col.p, op.pv_val.p, xv, np, yp);
      add_diag<<<nb(np), kT, 0, s>>>(yp, op.h_pp.p, xp, np);
    }
  }
  DQ_HIP(hipGetLastError());
}

}  // namespace
}  // namespace dq

extern "C" {

int dq_operator_create(int64_t n_v, int64_t n_p, const uint32_t* d_vv_row, const uint32_t* d_vv_col,
                       const double* d_vv_val, int64_t nnz_vv, const uint32_t* d_vp_row, const uint32_t* d_vp_col,
                       const double* d_vp_val, int64_t nnz_vp, const double* d_h_pp, void* stream,
                       dq_operator** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(out && n_v >= 0 && n_p >= 0 && nnz_vv >= 0 && nnz_vp >= 0, DQ_EINVAL, "bad arguments");
    DQ_REQUIRE(n_v < (int64_t{1} << 32) && n_p < (int64_t{1} << 32), DQ_EINVAL, "block dimension exceeds u32");
    DQ_REQUIRE(nnz_vv == 0 || (d_vv_row && d_vv_col && d_vv_val), DQ_EINVAL, "null H_VV arrays");
    DQ_REQUIRE(nnz_vp == 0 || (d_vp_row && d_vp_col && d_vp_val), DQ_EINVAL, "null H_VP arrays");
    DQ_REQUIRE(n_p == 0 || d_h_pp, DQ_EINVAL, "diag(H_PP) is required when n_p > 0");
    hipStream_t s = dq::as_stream(stream);
    auto op = std::make_unique<dq_operator>();
    op->n_v = n_v;
    op->n_p = n_p;
    op->nnz_vv = nnz_vv;
    op->nnz_vp = nnz_vp;
    dq::build_csr(d_vv_row, d_vv_col, d_vv_val, nnz_vv, 0, n_v, n_v, op->vv_ptr, op->vv_col, op->vv_val, s);
    dq::build_csr(d_vp_row, d_vp_col, d_vp_val, nnz_vp, 0, n_v, n_p, op->vp_ptr, op->vp_col, op->vp_val, s);
    dq::build_csr(d_vp_row, d_vp_col, d_vp_val, nnz_vp, 1, n_p, n_v, op->pv_ptr, op->pv_col, op->pv_val, s);
    op->h_pp.alloc(std::max<int64_t>(n_p, 1));
    if (n_p) DQ_HIP(hipMemcpyAsync(op->h_pp.p, d_h_pp, n_p * sizeof(double), hipMemcpyDeviceToDevice, s));
    op->work.alloc(std::max<int64_t>(n_v + n_p, 1));
    DQ_HIP(hipStreamSynchronize(s));
    *out = op.release();
  });
}

int dq_operator_apply(const dq_operator* op, const double* d_psi_v, const double* d_psi_p, double* d_y_v,
                      double* d_y_p, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(op, DQ_EINVAL, "null operator");
    DQ_REQUIRE(op->n_v == 0 || (d_psi_v && d_y_v), DQ_EINVAL, "null V vectors");
    DQ_REQUIRE(op->n_p == 0 || d_psi_p, DQ_EINVAL, "psi_P is required when n_p > 0");
    dq::apply(*op, d_psi_v, d_psi_p, d_y_v, d_y_p, dq::as_stream(stream));
  });
}

int dq_energy_step(const dq_operator* op, int mode, const double* d_sign, const double* d_logabs, double eps,
                   double* d_psi, double* d_hpsi, double* d_cot_logabs, double* h_energy, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(op, DQ_EINVAL, "null operator");
    DQ_REQUIRE(mode >= 0 && mode <= 2, DQ_EINVAL, "mode must be 0 (variational), 1 (proxy) or 2 (asymmetric)");
    const int64_t nv = op->n_v, np = mode == 0 ? 0 : op->n_p, n = nv + np;
    DQ_REQUIRE(n == 0 || (d_sign && d_logabs && d_psi && d_hpsi), DQ_EINVAL, "null vectors");
    hipStream_t s = dq::as_stream(stream);
    const unsigned gt = dq::nb(n);
    dq::DevBuf<double> part(std::max<int64_t>(4 * gt + 8, 8), s);
    double* pm = part.p;
    double* scal = part.p + 4 * gt;  // shift, then E / den_safe / num / den
    if (n > 0) {
      dq::block_max<<<gt, dq::kT, 0, s>>>(d_logabs, n, pm);
      dq::serial_max<<<1, 1, 0, s>>>(pm, gt, scal);
      dq::decode_psi<<<gt, dq::kT, 0, s>>>(d_sign, d_logabs, n, scal, d_psi);
      dq::apply(*op, d_psi, np ? d_psi + nv : nullptr, d_hpsi, np ? d_hpsi + nv : nullptr, s);
    }
    const unsigned gv = dq::nb(nv), gp = dq::nb(np);
    double* pnv = pm;
    double* pdv = pm + gt;
    double* pnp = pm + 2 * gt;
    double* pdp = pm + 3 * gt;
    if (nv) {
      dq::block_dot<<<gv, dq::kT, 0, s>>>(d_psi, d_hpsi, nv, pnv);
      dq::block_dot<<<gv, dq::kT, 0, s>>>(d_psi, d_psi, nv, pdv);
    }
    if (np) {
      dq::block_dot<<<gp, dq::kT, 0, s>>>(d_psi + nv, d_hpsi + nv, np, pnp);
      dq::block_dot<<<gp, dq::kT, 0, s>>>(d_psi + nv, d_psi + nv, np, pdp);
    }
    dq::finish_energy<<<1, 1, 0, s>>>(pnv, nv ? gv : 0, pnp, np ? gp : 0, pdv, pdp, mode == 1, mode == 1, eps,
                                      scal + 1);
    if (nv && d_cot_logabs) dq::cotangent<<<gv, dq::kT, 0, s>>>(d_hpsi, d_psi, nv, scal + 1, 0, d_cot_logabs);
    if (np && d_cot_logabs)
      dq::cotangent<<<gp, dq::kT, 0, s>>>(d_hpsi + nv, d_psi + nv, np, scal + 1, mode == 2, d_cot_logabs + nv);
    DQ_HIP(hipGetLastError());
    double e = 0.0;
    DQ_HIP(hipMemcpyAsync(&e, scal + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    if (h_energy) *h_energy = e;
  });
}

int dq_operator_free(dq_operator* op) {
  return dq::guarded([&] { delete op; });
}

}  // extern "C"
