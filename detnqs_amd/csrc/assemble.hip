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
#include "walk.hpp"

namespace dq {

namespace {

constexpr int kWavesPerBlock = 4;

// get_local_connections row (local_conn.cpp:60-91): diagonal and doubles kept
// if |v| > 1e-15, singles if |v| > sthr (max(eps1, 1e-15) with heat-bath).
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

#ifndef DQ_NT_STORES
#define DQ_NT_STORES 0
#endif

template <bool FILL>
struct ConnSink {
  static constexpr bool kNeedKet = FILL;
  static constexpr bool kCountOnly = !FILL;
  int64_t pos;
  uint64_t* out_dets;
  double* out_vals;
  double sthr;
  u64 xa, xb;
  // Output is written once and not re-read here: non-temporal stores keep the
  // heat-bath rows and integral tables resident in L2.
  __device__ __forceinline__ void emit(bool keep, u64 ka, u64 kb, double v) {
    const u64 mask = __ballot(keep);
    if (FILL && keep) {
      const int64_t at = pos + __popcll(mask & lanemask_lt());
      u64x2 d;
      d.x = ka;
      d.y = kb;
#if DQ_NT_STORES
      __builtin_nontemporal_store(d, reinterpret_cast<u64x2*>(out_dets) + at);
      __builtin_nontemporal_store(v, out_vals + at);
#else
      reinterpret_cast<u64x2*>(out_dets)[at] = d;
      out_vals[at] = v;
#endif
    }
    pos += __popcll(mask);
  }
  __device__ __forceinline__ void add_doubles(int c) { pos += c; }
  __device__ __forceinline__ void diag(double d) { emit(__lane_id() == 0 && fabs(d) > kMatThr, xa, xb, d); }
  __device__ __forceinline__ void single(bool ok, u64 ka, u64 kb, double v) { emit(ok && fabs(v) > sthr, ka, kb, v); }
  __device__ __forceinline__ void dbl(bool ok, u64 ka, u64 kb, double v) { emit(ok && fabs(v) > kMatThr, ka, kb, v); }
};

// ------------------------------------------------------------------ pipelined heat-bath fill
// The fill's heat-bath doubles, software-pipelined: the entry loads of the
// next (pair, chunk) item are issued before the stores of the current one.
// gfx950 counts vector loads and stores on one in-order counter, so a wait for
// loads issued after earlier stores also waits for those stores; here each
// wait covers only loads issued before the previous item's stores.  The waits
// are exact only if every item issues the same number of loads and stores, so
// both go through raw buffer resources: lanes past the row's kept entries (and
// past the pair row) use an offset past the resource's size, which the hardware
// drops (stores) or reads as 0 (loads).  Entries, order, values and output
// positions are those of walk_row's heat-bath loop.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// 64-entry chunks per item.  Six (384 entries) holds a whole same-spin pair
// list at 28 orbitals (C(28,2) = 378) and an opposite-spin one (784) in three
// items: 159.7 ms for the 1M-row fill against 168.5 with four (two, three,
// five, seven and eight chunks: 170.5, 169.1, 185.3, 169.7, 179.5 ms).
#ifndef DQ_FILL_CHUNKS
#define DQ_FILL_CHUNKS 6
#endif
constexpr int kFillChunks = DQ_FILL_CHUNKS;
constexpr int kOOB = static_cast<int>(0x80000000u);

// Raw buffer resource over [base, base + bytes) (gfx9 descriptor word 3).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, u32 bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}

struct HbItem {  // wave-uniform: pair (p > q) of the occupied SO list, chunk base
  int p, q, i, j;
  u32 off, len, base;
};

// Next item in walk order (pairs p = 1.., q < p; chunks of each pair's kept
// prefix); false past the last pair.
__device__ __forceinline__ bool hb_next(HbItem& it, const RowLists& L, int ne, const u32* __restrict__ hb_off,
                                        const u32* __restrict__ hb_len) {
  it.base += kFillChunks * kWave;
  while (it.base >= it.len) {
    if (++it.q >= it.p) {
      if (++it.p >= ne) return false;
      it.q = 0;
    }
    const int i = __builtin_amdgcn_readfirstlane(L.occSO[it.p]);
    const int j = __builtin_amdgcn_readfirstlane(L.occSO[it.q]);
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const size_t rid = static_cast<size_t>(hi) * (hi - 1) / 2 + lo;
    it.i = i;
    it.j = j;
    it.off = hb_off[rid];
    it.len = hb_len[rid] & ~kFullRow;
    it.base = 0;
  }
  return true;
}

struct HbRegs {
  u32 ab[kFillChunks];
  double v[kFillChunks];
};

// Loads of one item; with valid == false every lane reads out of range (no
// memory traffic), so each iteration issues the same loads.
__device__ __forceinline__ void hb_load(const HbItem& it, bool valid, __amdgpu_buffer_rsrc_t rab,
                                        __amdgpu_buffer_rsrc_t rv, HbRegs& r) {
  const int lane = __lane_id();
#pragma unroll
  for (int c = 0; c < kFillChunks; ++c) {
    const u32 k = it.base + c * kWave + lane;
    const bool in = valid && k < it.len;
    r.ab[c] = __builtin_amdgcn_raw_buffer_load_b32(rab, in ? static_cast<int>((it.off + k) * 4u) : kOOB, 0, 0);
    const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(rv, in ? static_cast<int>((it.off + k) * 8u) : kOOB, 0, 0);
    r.v[c] = __builtin_bit_cast(double, w);
  }
}

// Kept doubles of one item, stored at row-local positions pos.. (rdet / rval
// cover the row's output span).
__device__ __forceinline__ void hb_emit(const HbItem& it, const HbRegs& r, u64 xa, u64 xb, u64 ba, u64 bb,
                                        __amdgpu_buffer_rsrc_t rdet, __amdgpu_buffer_rsrc_t rval, u32& pos) {
  const int lane = __lane_id();
  const int i = it.i, j = it.j;
  const bool same = ((i ^ j) & 1) == 0;
  const int mi = i >> 1, mj = j >> 1;
#pragma unroll
  for (int c = 0; c < kFillChunks; ++c) {
    const u32 k = it.base + c * kWave + lane;
    const int a = static_cast<int>(r.ab[c] & 0xFFFFu), b = static_cast<int>(r.ab[c] >> 16);
    const bool ok = k < it.len && !(so_blocked(ba, bb, a) | so_blocked(ba, bb, b));
    u64 ka = xa, kb = xb;
    int perm;
    const int ma = a >> 1, mb = b >> 1;
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
    const double v = (perm & 1) ? -r.v[c] : r.v[c];
    const bool keep = ok && fabs(v) > kMatThr;
    const u64 m = __ballot(keep);
    const u32 at = pos + static_cast<u32>(__popcll(m & lanemask_lt()));
    const u32x4 d = {static_cast<u32>(ka), static_cast<u32>(ka >> 32), static_cast<u32>(kb), static_cast<u32>(kb >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(d, rdet, keep ? static_cast<int>(at * 16u) : kOOB, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rval, keep ? static_cast<int>(at * 8u) : kOOB,
                                          0, 0);
    pos += static_cast<u32>(__popcll(m));
  }
}

__device__ __forceinline__ void fill_hb_doubles(const DevView& V, const RowLists& L, u64 xa, u64 xb, int n_orb,
                                                const u32* __restrict__ hb_off, const u32* __restrict__ hb_len,
                                                uint64_t* row_dets, double* row_vals, u32 row_n, u32 pos) {
  const int ne = __popcll(xa) + __popcll(xb);
  if (ne < 2) return;
  const __amdgpu_buffer_rsrc_t rab = buf_rsrc(V.hb_ab, V.hb_n * 4u);
  const __amdgpu_buffer_rsrc_t rv = buf_rsrc(V.hb_v, V.hb_n * 8u);
  const __amdgpu_buffer_rsrc_t rdet = buf_rsrc(row_dets, row_n * 16u);
  const __amdgpu_buffer_rsrc_t rval = buf_rsrc(row_vals, row_n * 8u);
  const u64 ba = xa | ~low_mask(n_orb), bb = xb | ~low_mask(n_orb);
  pos = __builtin_amdgcn_readfirstlane(pos);
  HbItem cur{1, -1, 0, 0, 0u, 0u, 0u};
  if (!hb_next(cur, L, ne, hb_off, hb_len)) return;
  HbRegs A, B;
  hb_load(cur, true, rab, rv, A);
  // As many (dropped) stores as an item issues, so that on entry, as in every
  // later iteration, an item's loads are followed by one item's stores and
  // then by the next item's loads: the waits then never cover stores.
  // (distinct offsets, all past the resources, so none is merged away)
#pragma unroll
  for (int c = 0; c < kFillChunks; ++c) {
    __builtin_amdgcn_raw_buffer_store_b128(u32x4{0u, 0u, 0u, 0u}, rdet, kOOB + 32 * c, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b64(u32x2{0u, 0u}, rval, kOOB + 32 * c + 16, 0, 0);
  }
  while (true) {
    HbItem nxt = cur;
    const bool more = hb_next(nxt, L, ne, hb_off, hb_len);
    hb_load(nxt, more, rab, rv, B);
    hb_emit(cur, A, xa, xb, ba, bb, rdet, rval, pos);
    if (!more) break;
    cur = nxt;
    A = B;
  }
}

template <bool FILL>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void local_conn_kernel(
    DevView V, const uint64_t* __restrict__ dets, int64_t n_rows, int n_orb, int use_hb, double sthr,
    const u32* __restrict__ hb_off, const u32* __restrict__ hb_len, int64_t* __restrict__ counts,
    const int64_t* __restrict__ offsets, uint64_t* __restrict__ out_dets, double* __restrict__ out_vals,
    int* __restrict__ err) {
  __shared__ RowLists lists[kWavesPerBlock];
  const int wid = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wid;
  RowLists& L = lists[wid];
  const bool active = row < n_rows;
  u64 xa = 0, xb = 0;
  if (active) {  // wave-uniform words: the row's masks live in scalar registers
    xa = uniform_u64(dets[2 * row]);
    xb = uniform_u64(dets[2 * row + 1]);
  }
  build_lists(L, xa, xb, n_orb);
  __syncthreads();
  if (!active) return;
  if ((xa | xb) & ~low_mask(n_orb)) {  // orbitals beyond n_orb: no output, flag the call
    if (!FILL && lane == 0) {
      counts[row] = 0;
      atomicOr(err, 1);
    }
    return;
  }
  // FILL writes row r at offsets[r] - offsets[0], so a caller can stream row
  // chunks (offsets + start) through one output buffer.
  ConnSink<FILL> S{FILL ? offsets[row] - offsets[0] : 0, out_dets, out_vals, sthr, xa, xb};
  const RowEnv R = make_env(V, xa, xb, n_orb, L);
  if (FILL && use_hb) {
    // diagonal and singles through the sink, doubles through the pipeline
    walk_row(R, S, 4, hb_len, 0.0);
    // wave-uniform (the row is the wave's), so the output resources sit in scalar registers
    const int64_t start = uniform_i64(offsets[row] - offsets[0]);
    const u32 row_n = __builtin_amdgcn_readfirstlane(static_cast<u32>(offsets[row + 1] - offsets[row]));
    fill_hb_doubles(V, L, xa, xb, n_orb, hb_off, hb_len, out_dets + 2 * start, out_vals + start, row_n,
                    static_cast<u32>(S.pos - start));
    return;
  }
  walk_row(R, S, use_hb ? 1 : 0, hb_len, 0.0);
  if (!FILL && lane == 0) counts[row] = S.pos;
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

// `in` may equal `out`: each thread reads its items before it writes them.
__global__ void scan_apply(const int64_t* in, int64_t n, const int64_t* __restrict__ tile_off, int64_t* out) {
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

// out[0..n] = exclusive scan of in[0..n), out[n] = total.  `in` may be `out`.
void exclusive_scan_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s) {
  if (n == 0) {
    DQ_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), s));
    return;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  DevBuf<int64_t> sums(static_cast<size_t>(tiles), s);
  scan_tiles<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, n, sums.p);
  scan_sums<<<1, kScanBlock, 0, s>>>(sums.p, tiles);
  scan_apply<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, n, sums.p, out);
  DQ_HIP(hipGetLastError());
}

void launch_local_conn(const dq_ctx& c, const DevView& V, const uint64_t* dets, int64_t n, int n_orb, bool use_hb,
                       double eps1, const u32* hb_len, int64_t* counts, const int64_t* offsets,
                       uint64_t* out_dets, double* out_vals, bool fill, hipStream_t s, double sthr_set) {
  if (n == 0) return;
  const double sthr = sthr_set >= 0.0 ? sthr_set : (use_hb ? (eps1 > kMatThr ? eps1 : kMatThr) : kMatThr);
  const int64_t blocks = (n + kWavesPerBlock - 1) / kWavesPerBlock;
  DQ_REQUIRE(blocks < (1LL << 31), DQ_EINVAL, "too many rows for one launch");
  if (fill)
    local_conn_kernel<true><<<static_cast<unsigned>(blocks), kWave * kWavesPerBlock, 0, s>>>(
        V, dets, n, n_orb, use_hb ? 1 : 0, sthr, V.hb_off, hb_len, nullptr, offsets, out_dets, out_vals, c.d_err.p);
  else
    local_conn_kernel<false><<<static_cast<unsigned>(blocks), kWave * kWavesPerBlock, 0, s>>>(
        V, dets, n, n_orb, use_hb ? 1 : 0, sthr, V.hb_off, hb_len, counts, nullptr, nullptr, nullptr, c.d_err.p);
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
    hipStream_t s = dq::as_stream(stream);
    const dq::u32* len = nullptr;
    if (use_hb) {
      DQ_REQUIRE(c->hb, DQ_EINVAL, "Heat-Bath table not initialized");
      len = c->hb_lengths(eps1, s);
    }
    dq::DevBuf<int64_t> counts(static_cast<size_t>(n > 0 ? n : 1), s);
    dq::launch_local_conn(*c, c->view(), d_dets, n, n_orb, use_hb != 0, eps1, len, counts.p, nullptr, nullptr,
                          nullptr, false, s);
    dq::exclusive_scan_i64(counts.p, n, d_offsets, s);
    c->check_rows(s);
    if (h_total) {
      DQ_HIP(hipMemcpyAsync(h_total, d_offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      DQ_HIP(hipStreamSynchronize(s));
    }
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
      len = c->hb_lengths(eps1, dq::as_stream(stream));
    }
    dq::launch_local_conn(*c, c->view(), d_dets, n, n_orb, use_hb != 0, eps1, len, nullptr, d_offsets, d_out_dets,
                          d_out_vals, true, dq::as_stream(stream));
  });
}

}  // extern "C"
