// Slater–Condon row assembly (get_local_connections) on gfx950.
//
// One wavefront (64 lanes) owns one determinant row.  The row is emitted in
// the reference's order (local_conn.cpp:36-94):
//   [diagonal] [singles alpha] [singles beta] [doubles]
// where doubles follow det_ops.hpp:161-246 (aa, bb, ab nested loops) or, in
// heat-bath mode, ham_utils.hpp:280-313 (occupied SO pairs, alpha list then
// beta list, each row walked in heat-bath order).  Within a segment the lanes
// evaluate 64 consecutive candidates; a ballot + prefix popcount compacts the
// kept ones so their relative order is preserved exactly.
//
// Two passes over identical control flow: COUNT writes the kept count of each
// row, an exclusive int64 scan turns counts into offsets, FILL writes dets and
// values.  All FP64 arithmetic uses the reference's operands and order and is
// compiled without FMA contraction, so values are bit-identical.
#include "assemble.hpp"

namespace dq {

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

struct RowLists {
  int occA[64], occB[64], virA[64], virB[64];
  int occSO[128];
};

__device__ __forceinline__ u64 lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
}

// ham_eval.cpp:53-86: e1 = sum_p h_pp over the occupied SOs (alpha list then
// beta list), e2 = sum_{i>j} <pq||pq> over that list, returned as e1 + e2.
__device__ double diag_value(const DevView& V, u64 xa, u64 xb) {
  const int n = V.n;
  double e1 = 0.0;
  for (u64 m = xa; m; m &= m - 1) {
    const int p = ctz(m);
    e1 += V.h1[p * n + p];
  }
  for (u64 m = xb; m; m &= m - 1) {
    const int p = ctz(m);
    e1 += V.h1[p * n + p];
  }
  double e2 = 0.0;
  for (u64 m = xa; m; m &= m - 1) {  // p alpha: q alpha below p
    const int p = ctz(m);
    for (u64 q = xa & low_mask(p); q; q &= q - 1) {
      const int k = ctz(q);
      e2 += V.Jd[p * n + k] - V.Kd[p * n + k];
    }
  }
  for (u64 m = xb; m; m &= m - 1) {  // p beta: all alpha, then beta below p
    const int p = ctz(m);
    for (u64 q = xa; q; q &= q - 1) {
      const int k = ctz(q);
      e2 += V.Jd[p * n + k] - 0.0;
    }
    for (u64 q = xb & low_mask(p); q; q &= q - 1) {
      const int k = ctz(q);
      e2 += V.Jd[p * n + k] - V.Kd[p * n + k];
    }
  }
  return e1 + e2;
}

// ham_eval.cpp:88-115 for removing r and adding a (spin `spin`): h_ra plus, over
// the occupied orbitals of the excited det except a (i.e. those of x except r),
// alpha then beta ascending, <r k||a k> = (ra|kk) - [(rk|ka) if same spin].
__device__ double single_value(const DevView& V, u64 xa, u64 xb, int spin, int r, int a) {
  const int n = V.n;
  double v = V.h1[r * n + a];
  const double* J = V.Js + static_cast<size_t>(r) * n * n + a;
  const double* K = V.Ks + static_cast<size_t>(r) * n * n + a;
  u64 ma = spin == 0 ? (xa & ~(1ULL << r)) : xa;
  u64 mb = spin == 1 ? (xb & ~(1ULL << r)) : xb;
  for (u64 m = ma; m; m &= m - 1) {
    const int k = ctz(m);
    v += J[k * n] - (spin == 0 ? K[k * n] : 0.0);
  }
  for (u64 m = mb; m; m &= m - 1) {
    const int k = ctz(m);
    v += J[k * n] - (spin == 1 ? K[k * n] : 0.0);
  }
  return v;
}

__device__ __forceinline__ double g2(const DevView& V, int p, int q, int r, int s) {
  return V.h2[key2(p, q, r, s)];
}

__device__ __forceinline__ bool so_occupied(u64 xa, u64 xb, int so, int n_orb) {
  const int mo = so >> 1;
  if (mo < 0 || mo >= n_orb) return true;  // ham_utils.cpp:470-479
  return (((so & 1) ? xb : xa) >> mo) & 1ULL;
}

struct Emitter {
  int64_t pos;
  uint64_t* out_dets;
  double* out_vals;
  template <bool FILL>
  __device__ __forceinline__ void emit(bool keep, u64 ka, u64 kb, double v) {
    const u64 mask = __ballot(keep);
    if (FILL && keep) {
      const int64_t at = pos + __popcll(mask & lanemask_lt());
      reinterpret_cast<ulonglong2*>(out_dets)[at] = make_ulonglong2(ka, kb);
      out_vals[at] = v;
    }
    pos += __popcll(mask);
  }
};

// Pair index t -> (u, w), u < w < v, lexicographic.
__device__ __forceinline__ void pair_from_index(int t, int v, int& u, int& w) {
  int start = 0;
  u = 0;
  while (t >= start + (v - 1 - u)) {
    start += v - 1 - u;
    ++u;
  }
  w = u + 1 + (t - start);
}

template <bool FILL>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void local_conn_kernel(
    DevView V, const uint64_t* __restrict__ dets, int64_t n_rows, int n_orb, int use_hb, double sthr,
    const u32* __restrict__ hb_len, int64_t* __restrict__ counts, const int64_t* __restrict__ offsets,
    uint64_t* __restrict__ out_dets, double* __restrict__ out_vals, int* __restrict__ err) {
  __shared__ RowLists lists[kWavesPerBlock];
  const int wid = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wid;
  RowLists& L = lists[wid];
  const bool active = row < n_rows;
  u64 xa = 0, xb = 0;
  if (active) {
    xa = dets[2 * row];
    xb = dets[2 * row + 1];
  }
  const u64 full = low_mask(n_orb);
  const u64 va = ~xa & full, vb = ~xb & full;
  const int na = __popcll(xa), nb = __popcll(xb), nva = __popcll(va), nvb = __popcll(vb);
  // scatter the orbital lists (one bit position per lane)
  {
    const u64 bit = 1ULL << lane;
    const u64 below = low_mask(lane);
    if (xa & bit) L.occA[__popcll(xa & below)] = lane;
    if (xb & bit) L.occB[__popcll(xb & below)] = lane;
    if (va & bit) L.virA[__popcll(va & below)] = lane;
    if (vb & bit) L.virB[__popcll(vb & below)] = lane;
    if (xa & bit) L.occSO[__popcll(xa & below)] = 2 * lane;
    if (xb & bit) L.occSO[na + __popcll(xb & below)] = 2 * lane + 1;
  }
  __syncthreads();
  if (!active) return;
  if ((xa | xb) & ~full) {  // orbitals beyond n_orb: no output, flag the call
    if (!FILL && lane == 0) {
      counts[row] = 0;
      atomicOr(err, 1);
    }
    return;
  }

  // FILL writes row r at offsets[r] - offsets[0], so a caller can stream row
  // chunks (offsets + start) through one output buffer.
  Emitter E{FILL ? offsets[row] - offsets[0] : 0, out_dets, out_vals};

  // ---- diagonal (local_conn.cpp:60-67)
  {
    const double d = diag_value(V, xa, xb);
    const bool keep = (lane == 0) && fabs(d) > kMatThr;
    E.template emit<FILL>(keep, xa, xb, d);
  }

  // ---- singles (local_conn.cpp:69-76; det_ops.hpp:123-159)
  for (int spin = 0; spin < 2; ++spin) {
    const int no = spin ? nb : na, nv = spin ? nvb : nva;
    const int* O = spin ? L.occB : L.occA;
    const int* Vl = spin ? L.virB : L.virA;
    const u64 xs = spin ? xb : xa;
    const int S = no * nv;
    for (int base = 0; base < S; base += kWave) {
      const int s = base + lane;
      bool keep = false;
      u64 ka = xa, kb = xb;
      double v = 0.0;
      if (s < S) {
        const int r = O[s / nv], a = Vl[s % nv];
        const double e = single_value(V, xa, xb, spin, r, a);
        const double ph = (between(xs, a, r) & 1) ? -1.0 : 1.0;
        v = ph * e;
        keep = fabs(v) > sthr;
        const u64 flip = (1ULL << r) | (1ULL << a);
        if (spin == 0)
          ka ^= flip;
        else
          kb ^= flip;
      }
      E.template emit<FILL>(keep, ka, kb, v);
    }
  }

  // ---- doubles
  if (use_hb) {
    // ham_utils.hpp:280-313: pairs (p > q) of the occupied SO list.
    const int ne = na + nb;
    for (int p = 1; p < ne; ++p) {
      const int i = L.occSO[p];
      for (int q = 0; q < p; ++q) {
        const int j = L.occSO[q];
        const int hi = i > j ? i : j, lo = i > j ? j : i;
        const size_t rid = static_cast<size_t>(hi) * (hi - 1) / 2 + lo;
        const u32 len = hb_len[rid];
        const u32 off = V.hb_off[rid];
        const bool same = ((i ^ j) & 1) == 0;
        for (u32 base = 0; base < len; base += kWave) {
          const u32 k = base + lane;
          bool keep = false;
          u64 ka = xa, kb = xb;
          double v = 0.0;
          if (k < len) {
            const u32 ab = V.hb_ab[off + k];
            const int a = static_cast<int>(ab & 0xFFFFu), b = static_cast<int>(ab >> 16);
            const double val = V.hb_v[off + k];
            keep = !so_occupied(xa, xb, a, n_orb) && !so_occupied(xa, xb, b, n_orb) && fabs(val) > kMatThr;
            if (FILL && keep) {
              int perm;
              const int mi = i >> 1, mj = j >> 1, ma = a >> 1, mb = b >> 1;
              if (same) {
                const u64 xs = (i & 1) ? xb : xa;
                perm = parity2(xs, mi < mj ? mi : mj, mi < mj ? mj : mi, ma < mb ? ma : mb, ma < mb ? mb : ma);
                const u64 flip = (1ULL << mi) | (1ULL << mj) | (1ULL << ma) | (1ULL << mb);
                if (i & 1)
                  kb ^= flip;
                else
                  ka ^= flip;
              } else {
                const int ra = (i & 1) ? mj : mi, rb = (i & 1) ? mi : mj;
                const int aa = (a & 1) ? mb : ma, ab_ = (a & 1) ? ma : mb;
                perm = between(xa, aa, ra) + between(xb, ab_, rb);
                ka ^= (1ULL << ra) | (1ULL << aa);
                kb ^= (1ULL << rb) | (1ULL << ab_);
              }
              v = (perm & 1) ? -val : val;
            }
          }
          E.template emit<FILL>(keep, ka, kb, v);
        }
      }
    }
  } else {
    // det_ops.hpp:161-215: same-spin doubles, alpha then beta.
    for (int spin = 0; spin < 2; ++spin) {
      const int no = spin ? nb : na, nv = spin ? nvb : nva;
      if (no < 2 || nv < 2) continue;
      const int* O = spin ? L.occB : L.occA;
      const int* Vl = spin ? L.virB : L.virA;
      const u64 xs = spin ? xb : xa;
      const int NP = nv * (nv - 1) / 2;
      for (int h1i = 0; h1i < no; ++h1i)
        for (int h2i = h1i + 1; h2i < no; ++h2i) {
          const int h0 = O[h1i], h1 = O[h2i];
          for (int base = 0; base < NP; base += kWave) {
            const int t = base + lane;
            bool keep = false;
            u64 ka = xa, kb = xb;
            double v = 0.0;
            if (t < NP) {
              int u, w;
              pair_from_index(t, nv, u, w);
              const int p0 = Vl[u], p1 = Vl[w];
              const double val = g2(V, h0, p0, h1, p1) - g2(V, h0, p1, h1, p0);
              keep = fabs(val) > kMatThr;
              if (FILL) {
                const int perm = parity2(xs, h0, h1, p0, p1);
                v = (perm & 1) ? -val : val;
                const u64 flip = (1ULL << h0) | (1ULL << h1) | (1ULL << p0) | (1ULL << p1);
                if (spin)
                  kb ^= flip;
                else
                  ka ^= flip;
              }
            }
            E.template emit<FILL>(keep, ka, kb, v);
          }
        }
    }
    // det_ops.hpp:217-246: opposite-spin doubles, nested (ha, pa, hb, pb).
    if (na && nva && nb && nvb) {
      const int NB = nb * nvb;
      for (int hai = 0; hai < na; ++hai)
        for (int pai = 0; pai < nva; ++pai) {
          const int ha = L.occA[hai], pa = L.virA[pai];
          const int pa_perm = between(xa, pa, ha);
          for (int base = 0; base < NB; base += kWave) {
            const int t = base + lane;
            bool keep = false;
            u64 ka = xa, kb = xb;
            double v = 0.0;
            if (t < NB) {
              const int hb = L.occB[t / nvb], pb = L.virB[t % nvb];
              const double val = g2(V, ha, pa, hb, pb);
              keep = fabs(val) > kMatThr;
              if (FILL) {
                const int perm = pa_perm + between(xb, pb, hb);
                v = (perm & 1) ? -val : val;
                ka ^= (1ULL << ha) | (1ULL << pa);
                kb ^= (1ULL << hb) | (1ULL << pb);
              }
            }
            E.template emit<FILL>(keep, ka, kb, v);
          }
        }
    }
  }

  if (!FILL && lane == 0) counts[row] = E.pos;
}

__global__ void diag_kernel(DevView V, const uint64_t* __restrict__ dets, int64_t n, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 xa = dets[2 * i], xb = dets[2 * i + 1];
  out[i] = ((xa | xb) & ~low_mask(V.n)) ? __longlong_as_double(0x7FF8000000000000LL) : diag_value(V, xa, xb);
}

// ------------------------------------------------------------------ int64 exclusive scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ int64_t block_exclusive_scan(int64_t x, int64_t* total) {
  __shared__ int64_t wsum[kScanBlock / kWave];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  int64_t incl = x;
  for (int d = 1; d < kWave; d <<= 1) {
    const int64_t y = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += y;
  }
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  int64_t before = 0, all = 0;
  for (int w = 0; w < kScanBlock / kWave; ++w) {
    if (w < wid) before += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + incl - x;
}

__global__ void scan_tiles(const int64_t* __restrict__ in, int64_t n, int64_t* __restrict__ tile_sums) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int64_t s = 0;
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  int64_t tot;
  (void)block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of the tile sums in place.
__global__ void scan_sums(int64_t* __restrict__ sums, int64_t m) {
  int64_t carry = 0;
  for (int64_t base = 0; base < m; base += kScanBlock) {
    const int64_t i = base + threadIdx.x;
    const int64_t x = i < m ? sums[i] : 0;
    int64_t tot;
    const int64_t ex = block_exclusive_scan(x, &tot);
    if (i < m) sums[i] = carry + ex;
    carry += tot;
    __syncthreads();
  }
}

__global__ void scan_apply(const int64_t* __restrict__ in, int64_t n, const int64_t* __restrict__ tile_off,
                           int64_t* __restrict__ out) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t ex = block_exclusive_scan(s, &tot) + tile_off[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanBlock - 1) out[n] = ex;
}

}  // namespace

// out[0..n] = exclusive scan of in[0..n), out[n] = total.  `in` may alias `out`
// only if distinct from it; callers pass separate buffers.
void exclusive_scan_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s) {
  if (n == 0) {
    DQ_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), s));
    return;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  DevBuf<int64_t> sums(static_cast<size_t>(tiles));
  scan_tiles<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, n, sums.p);
  scan_sums<<<1, kScanBlock, 0, s>>>(sums.p, tiles);
  scan_apply<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, n, sums.p, out);
  DQ_HIP(hipGetLastError());
  DQ_HIP(hipStreamSynchronize(s));  // `sums` is freed on return
}

void launch_local_conn(const dq_ctx& c, const DevView& V, const uint64_t* dets, int64_t n, int n_orb, bool use_hb,
                       double eps1, const u32* hb_len, int64_t* counts, const int64_t* offsets,
                       uint64_t* out_dets, double* out_vals, bool fill, hipStream_t s) {
  if (n == 0) return;
  const double sthr = use_hb ? (eps1 > kMatThr ? eps1 : kMatThr) : kMatThr;
  const int64_t blocks = (n + kWavesPerBlock - 1) / kWavesPerBlock;
  DQ_REQUIRE(blocks < (1LL << 31), DQ_EINVAL, "too many rows for one launch");
  if (fill)
    local_conn_kernel<true><<<static_cast<unsigned>(blocks), kWave * kWavesPerBlock, 0, s>>>(
        V, dets, n, n_orb, use_hb ? 1 : 0, sthr, hb_len, nullptr, offsets, out_dets, out_vals, c.d_err.p);
  else
    local_conn_kernel<false><<<static_cast<unsigned>(blocks), kWave * kWavesPerBlock, 0, s>>>(
        V, dets, n, n_orb, use_hb ? 1 : 0, sthr, hb_len, counts, nullptr, nullptr, nullptr, c.d_err.p);
  DQ_HIP(hipGetLastError());
}

void launch_diag(const DevView& V, const uint64_t* dets, int64_t n, double* out, hipStream_t s) {
  if (n == 0) return;
  const int64_t blocks = (n + 255) / 256;
  diag_kernel<<<static_cast<unsigned>(blocks), 256, 0, s>>>(V, dets, n, out);
  DQ_HIP(hipGetLastError());
}

}  // namespace dq

// ----------------------------------------------------------------------------- C-ABI
namespace {
void check_dets_args(const dq_ctx* ctx, const uint64_t* dets, int64_t n, int n_orb) {
  DQ_REQUIRE(ctx, DQ_EINVAL, "null context");
  DQ_REQUIRE(n >= 0, DQ_EINVAL, "negative row count");
  DQ_REQUIRE(n == 0 || dets, DQ_EINVAL, "null dets");
  DQ_REQUIRE(n_orb >= 1 && n_orb <= ctx->n_orb, DQ_EINVAL,
             "n_orb must be in [1, n_orb of the integral context]");
}
}  // namespace

extern "C" {

int dq_ham_diag(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, double* d_out, void* stream) {
  return dq::guarded([&] {
    check_dets_args(ctx, d_dets, n, ctx ? ctx->n_orb : 1);
    DQ_REQUIRE(n == 0 || d_out, DQ_EINVAL, "null output");
    dq::launch_diag(ctx->view(), d_dets, n, d_out, dq::as_stream(stream));
  });
}

int dq_local_conn_count(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, int n_orb, int use_hb, double eps1,
                        int64_t* d_offsets, int64_t* h_total, void* stream) {
  return dq::guarded([&] {
    check_dets_args(ctx, d_dets, n, n_orb);
    DQ_REQUIRE(d_offsets, DQ_EINVAL, "null offsets");
    auto* c = const_cast<dq_ctx*>(ctx);
    const dq::u32* len = nullptr;
    if (use_hb) {
      DQ_REQUIRE(c->hb, DQ_EINVAL, "Heat-Bath table not initialized");
      len = c->hb_lengths(eps1);
    }
    hipStream_t s = dq::as_stream(stream);
    dq::DevBuf<int64_t> counts(static_cast<size_t>(n > 0 ? n : 1));
    dq::launch_local_conn(*c, c->view(), d_dets, n, n_orb, use_hb != 0, eps1, len, counts.p, nullptr, nullptr,
                          nullptr, false, s);
    dq::exclusive_scan_i64(counts.p, n, d_offsets, s);
    c->check_rows(s);
    if (h_total) DQ_HIP(hipMemcpy(h_total, d_offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost));
  });
}

int dq_local_conn_fill(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, int n_orb, int use_hb, double eps1,
                       const int64_t* d_offsets, uint64_t* d_out_dets, double* d_out_vals, void* stream) {
  return dq::guarded([&] {
    check_dets_args(ctx, d_dets, n, n_orb);
    DQ_REQUIRE(d_offsets && d_out_dets && d_out_vals, DQ_EINVAL, "null output");
    auto* c = const_cast<dq_ctx*>(ctx);
    const dq::u32* len = nullptr;
    if (use_hb) {
      DQ_REQUIRE(c->hb, DQ_EINVAL, "Heat-Bath table not initialized");
      len = c->hb_lengths(eps1);
    }
    dq::launch_local_conn(*c, c->view(), d_dets, n, n_orb, use_hb != 0, eps1, len, nullptr, d_offsets, d_out_dets,
                          d_out_vals, true, dq::as_stream(stream));
  });
}

}  // extern "C"
