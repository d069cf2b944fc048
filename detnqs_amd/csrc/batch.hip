// Multi-instance fused local energy: row ranges of many independent instances
// (own integrals, own V, own amplitudes) in one launch.  The serving loop
// (continuous batching: every launch takes the next row ranges of all
// admitted requests) and the batched form (heterogeneous instances in one
// persistent, graph-captured launch) both run on it.
//
// Each work item is one row of one segment; a wavefront computes
// sigma_i = sum_{j in V} H_ij c_j exactly as the single-instance energy kernel
// (local_conn.cpp:148-229: rows with c_i^2 < 1e-12 give 0), so a row's result
// does not depend on which other rows share the launch.  The energy of an
// instance, sum_i c_i sigma_i, is reduced over its whole sigma in the
// single-instance order (256-row blocks by a tree, block sums in order).
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "eloc_strings.hpp"
#include "vvcache.hpp"
#include "walk.hpp"

namespace dq {

namespace {

struct EnergySinkB {
  static constexpr bool kNeedKet = true;
  static constexpr bool kCountOnly = false;
  __device__ __forceinline__ void add_doubles(int) {}
  HashView H;
  const double* c;
  double ci, sthr, acc;
  __device__ __forceinline__ void diag(double d) {
    if (__lane_id() == 0 && fabs(d) > kMatThr) acc += d * ci;
  }
  __device__ __forceinline__ void add(bool keep, u64 ka, u64 kb, double v) {
    if (keep) {
      const long j = hash_find(H, ka, kb);
      if (j >= 0) acc += v * c[j];
    }
  }
  __device__ __forceinline__ void single(bool ok, u64 ka, u64 kb, double v) { add(ok && fabs(v) > sthr, ka, kb, v); }
  __device__ __forceinline__ void dbl(bool ok, u64 ka, u64 kb, double v) { add(ok && fabs(v) > kMatThr, ka, kb, v); }
};

// Largest s with start[s] <= item (start[0] = 0 <= item < start[n]).
__device__ __forceinline__ int find_segment(const int64_t* __restrict__ start, int n, int64_t item) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= item)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Waves of a persistent block take different numbers of rows, so all LDS
// hand-offs use wave-scope barriers only.
template <int METHODS>
__device__ void eloc_row(const SegDev& S, int64_t row, StrWave& W) {
  const int lane = __lane_id();
  const u64 xa = uniform_u64(S.dets[2 * row]), xb = uniform_u64(S.dets[2 * row + 1]);
  RowLists& L = W.L;
  build_lists(L, xa, xb, S.n_orb);
  wave_sync();
  if ((xa | xb) & ~low_mask(S.n_orb)) {
    if (lane == 0) {
      S.sigma[row] = 0.0;
      atomicOr(S.err, 1);
    }
  } else {
    const double ci = S.coeffs[row];
    if (ci * ci < 1e-12) {
      if (lane == 0) S.sigma[row] = 0.0;
    } else {
      double a = 0.0;
      const bool strings = METHODS == 1 || (METHODS == 3 && S.method == DQ_ELOC_STRINGS);
      if (strings) {
        if constexpr ((METHODS & 1) != 0) a = str_energy_row(S.V, S.X, S.C, xa, xb, S.n_orb, ci, S.sthr, S.dmin, W);
      } else {
        if constexpr ((METHODS & 2) != 0) {
          EnergySinkB E{S.H, S.coeffs, ci, S.sthr, 0.0};
          const RowEnv R = make_env(S.V, xa, xb, S.n_orb, L);
          walk_row(R, E, S.dbl_mode, S.hb_len, 0.0);
          a = E.acc;
        }
      }
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, kWave);
      if (lane == 0) S.sigma[row] = a;
    }
  }
  wave_sync();  // the lists are rewritten by the wave's next item
}

template <int METHODS>
__device__ __forceinline__ void eloc_item(const SegDev* __restrict__ segs, int nseg, const int64_t* __restrict__ start,
                                          int64_t item, StrWave& L) {
  // wave-uniform: the segment's fields are read with scalar loads
  const int s = __builtin_amdgcn_readfirstlane(find_segment(start, nseg, item));
  const SegDev S = load_uniform(segs + s);
  const int64_t row = uniform_i64(S.row_begin + (item - start[s]));
  eloc_row<METHODS>(S, row, L);
}

// METHODS: bit 0 = some segment uses the string method, bit 1 = the walk; a
// launch is compiled for the methods it contains.
template <int WPB, int METHODS>
__global__ __launch_bounds__(kWave* WPB) void eloc_kernel(const SegDev* __restrict__ segs, int nseg,
                                                          const int64_t* __restrict__ start, int64_t total,
                                                          int persistent, unsigned long long* __restrict__ counter) {
  __shared__ StrWave lds[WPB];
  StrWave& L = lds[threadIdx.x / kWave];
  if (!persistent) {
    const int64_t item = static_cast<int64_t>(blockIdx.x) * WPB + threadIdx.x / kWave;
    if (item >= total) return;
    eloc_item<METHODS>(segs, nseg, start, item, L);
    return;
  }
  // Work queue: each wave takes the next row until the queue is drained; every
  // wave leaves through the same bound check.
  while (true) {
    unsigned long long it = 0;
    if (__lane_id() == 0) it = atomicAdd(counter, 1ULL);
    const int64_t item = uniform_i64(static_cast<int64_t>(__shfl(it, 0, kWave)));
    if (item >= total) return;
    eloc_item<METHODS>(segs, nseg, start, item, L);
  }
}

// Rows of layout segments: sigma_i from the lane-ordered H_VV, one wavefront
// per row (the arithmetic of vv_sigma in vvcache.hip).  A separate kernel, so
// the persistent walk/string kernel keeps its own register budget.
constexpr int kVvWpb = 4;

__global__ __launch_bounds__(kWave* kVvWpb) void vv_seg_kernel(const SegDev* __restrict__ segs, int nseg,
                                                               const int64_t* __restrict__ start, int64_t total) {
  const int64_t item = static_cast<int64_t>(blockIdx.x) * kVvWpb + threadIdx.x / kWave;
  if (item >= total) return;  // whole waves leave
  const int s = __builtin_amdgcn_readfirstlane(find_segment(start, nseg, item));
  const SegDev S = load_uniform(segs + s);
  const int64_t row = uniform_i64(S.row_begin + (item - start[s]));
  const int lane = __lane_id();
  const double ci = S.coeffs[row];
  if (ci * ci < 1e-12) {  // local_conn.cpp:190
    if (lane == 0) S.sigma[row] = 0.0;
    return;
  }
  double a = vv_row_sum(S.vv_col, S.vv_val, S.vv_off[row], S.vv_off[row + 1], S.coeffs);
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, kWave);
  if (lane == 0) S.sigma[row] = a;
}

__global__ void reset_counter(unsigned long long* c) { *c = 0ULL; }

__global__ void gather_many(const GatherSeg* __restrict__ gs, int ng, const int64_t* __restrict__ gstart,
                            int64_t total) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int e = find_segment(gstart, ng, i);
  const GatherSeg G = gs[e];
  const int64_t k = i - gstart[e];
  G.ca[k] = G.c[G.a_row[k]];
  G.cb[k] = G.c[G.b_row[k]];
}

// Per 256-row chunk of every energy instance: tree sum of c_i * sigma_i.
__global__ __launch_bounds__(256) void energy_chunks(const EnergySeg* __restrict__ es, int ne,
                                                     const int64_t* __restrict__ estart, double* __restrict__ part) {
  __shared__ double buf[256];
  const int64_t chunk = blockIdx.x;
  const int e = find_segment(estart, ne, chunk);
  const EnergySeg E = es[e];
  const int64_t i = (chunk - estart[e]) * 256 + threadIdx.x;
  buf[threadIdx.x] = i < E.n_rows ? E.coeffs[i] * E.sigma[i] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) buf[threadIdx.x] += buf[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[chunk] = buf[0];
}

// One thread per energy instance: chunk sums added in order.
__global__ void energy_sums(const EnergySeg* __restrict__ es, int ne, const int64_t* __restrict__ estart,
                            const double* __restrict__ part, double* __restrict__ energy) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double t = 0.0;
  for (int64_t k = estart[e]; k < estart[e + 1]; ++k) t += part[k];
  energy[es[e].slot] = t;
}

}  // namespace

void validate_tuning(const dq_eloc_tuning& t) {
  DQ_REQUIRE(t.waves_per_block == 1 || t.waves_per_block == 2 || t.waves_per_block == 4 || t.waves_per_block == 8,
             DQ_EINVAL, "waves_per_block must be 1, 2, 4 or 8");
  DQ_REQUIRE(t.grid_blocks >= 0, DQ_EINVAL, "grid_blocks must be >= 0");
}

int default_grid_blocks(int waves_per_block) {
  int dev = 0, cus = 0;
  DQ_HIP(hipGetDevice(&dev));
  DQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(1, cus * 32 / waves_per_block);  // 32 waves per CU
}

void eloc_thresholds(const dq_ctx& c, int use_hb, double eps1, double* sthr, double* dmin) {
  *sthr = use_hb ? std::max(eps1, kMatThr) : kMatThr;
  *dmin = use_hb ? std::max(c.hb_thr, eps1 > 0.0 ? eps1 : 0.0) : 0.0;
}

SegDev make_segment(const dq_eloc_segment& s) {
  DQ_REQUIRE(s.ctx && s.vmap && s.d_dets && s.d_coeffs && s.d_sigma, DQ_EINVAL, "segment has a null pointer");
  DQ_REQUIRE(s.n_orb >= 1 && s.n_orb <= s.ctx->n_orb, DQ_EINVAL, "n_orb must be in [1, n_orb of the integral context]");
  DQ_REQUIRE(s.vmap->n == s.n_rows, DQ_EINVAL, "segment DetMap size does not match n_rows");
  DQ_REQUIRE(0 <= s.row_begin && s.row_begin <= s.row_end && s.row_end <= s.n_rows, DQ_EINVAL,
             "segment row range out of bounds");
  DQ_REQUIRE(s.method == DQ_ELOC_STRINGS || s.method == DQ_ELOC_WALK, DQ_EINVAL, "unknown energy method");
  if (s.vv) {
    DQ_REQUIRE(s.method == DQ_ELOC_STRINGS, DQ_EINVAL, "a cached H_VV serves the string method only");
    DQ_REQUIRE(s.vv->vmap == s.vmap && s.vv->ctx == s.ctx && s.vv->n == s.n_rows && s.vv->n_orb == s.n_orb &&
                   s.vv->use_hb == s.use_hb && s.vv->eps1 == s.eps1,
               DQ_EINVAL, "the cached H_VV was built for another V, context or screening");
  }
  auto* c = const_cast<dq_ctx*>(s.ctx);
  if (s.use_hb) DQ_REQUIRE(c->hb, DQ_EINVAL, "compute_variational_energy: Heat-Bath mode requires non-null hb_table");
  SegDev d{};
  d.X = s.vmap->x.view;
  d.method = s.method;
  eloc_thresholds(*c, s.use_hb, s.eps1, &d.sthr, &d.dmin);
  d.V = c->view();
  d.H = s.vmap->h.view();
  d.dets = s.d_dets;
  d.coeffs = s.d_coeffs;
  d.sigma = s.d_sigma;
  // the plan is launched later, on a stream not known here: wait on the host
  d.hb_len = s.use_hb ? c->hb_lengths(s.eps1, nullptr, true) : nullptr;
  d.row_begin = s.row_begin;
  d.n_orb = s.n_orb;
  d.dbl_mode = s.use_hb ? 1 : 0;
  if (s.vv) {
    d.method = kElocVv;
    d.vv_off = s.vv->off.p;
    d.vv_col = s.vv->col.p;
    d.vv_val = s.vv->val.p;
  }
  return d;
}

void launch_eloc(const PlanDev& P, const dq_eloc_tuning& t, double* d_energy, hipStream_t s) {
  const int wpb = t.waves_per_block;
  const bool pers = t.persistent != 0;
  if (P.gather_total > 0) {
    gather_many<<<static_cast<unsigned>((P.gather_total + 255) / 256), 256, 0, s>>>(P.gsegs, P.n_gather, P.gstart,
                                                                                   P.gather_total);
    DQ_HIP(hipGetLastError());
  }
  if (P.vtotal > 0) {
    vv_seg_kernel<<<static_cast<unsigned>((P.vtotal + kVvWpb - 1) / kVvWpb), kWave * kVvWpb, 0, s>>>(
        P.vsegs, P.n_vseg, P.vstart, P.vtotal);
    DQ_HIP(hipGetLastError());
  }
  if (P.total > 0) {
    unsigned grid;
    if (pers) {
      grid = static_cast<unsigned>(t.grid_blocks > 0 ? t.grid_blocks : default_grid_blocks(wpb));
      reset_counter<<<1, 1, 0, s>>>(P.counter);
    } else {
      grid = static_cast<unsigned>((P.total + wpb - 1) / wpb);
    }
    const int M = P.methods;
    DQ_REQUIRE(M >= 1 && M <= 3, DQ_EINVAL, "launch has no energy method");
#define DQ_EL(W, MM)                                                                                      \
  if (wpb == W && M == MM) {                                                                              \
    eloc_kernel<W, MM><<<grid, kWave * W, 0, s>>>(P.segs, P.nseg, P.start, P.total, pers ? 1 : 0, P.counter); \
    launched = true;                                                                                      \
  }
    bool launched = false;
    DQ_EL(1, 1) DQ_EL(2, 1) DQ_EL(4, 1) DQ_EL(8, 1)
    DQ_EL(1, 2) DQ_EL(2, 2) DQ_EL(4, 2) DQ_EL(8, 2)
    DQ_EL(1, 3) DQ_EL(2, 3) DQ_EL(4, 3) DQ_EL(8, 3)
#undef DQ_EL
    DQ_REQUIRE(launched, DQ_EINVAL, "waves_per_block must be 1, 2, 4 or 8");
    DQ_HIP(hipGetLastError());
  }
  if (d_energy && P.n_energy > 0) {
    if (P.n_chunks > 0)
      energy_chunks<<<static_cast<unsigned>(P.n_chunks), 256, 0, s>>>(P.esegs, P.n_energy, P.estart, P.part);
    energy_sums<<<static_cast<unsigned>((P.n_energy + 63) / 64), 64, 0, s>>>(P.esegs, P.n_energy, P.estart, P.part,
                                                                             d_energy);
    DQ_HIP(hipGetLastError());
  }
}

}  // namespace dq

dq_eloc_plan::~dq_eloc_plan() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
}

// ----------------------------------------------------------------------------- C-ABI
extern "C" {

int dq_detmap_build(const uint64_t* d_dets, int64_t n, int strict, void* stream, dq_detmap** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(out && n >= 0 && (n == 0 || d_dets), DQ_EINVAL, "bad arguments");
    auto m = std::make_unique<dq_detmap>();
    m->n = n;
    hipStream_t s = dq::as_stream(stream);
    const bool dup = m->h.build(d_dets, n, s, false);
    if (dup && strict) throw dq::Error(DQ_EINVAL, "DetMap::from_ordered: duplicate determinant found");
    m->x.build(d_dets, n, m->h.view(), s, false, m->h.vals.p);  // the DetMap then gives coefficient indices
    DQ_HIP(hipStreamSynchronize(s));
    *out = m.release();
  });
}

int dq_detmap_free(dq_detmap* m) {
  return dq::guarded([&] { delete m; });
}

int dq_eloc_plan_create(const dq_eloc_segment* segs, int n_segs, const dq_eloc_tuning* tune, dq_eloc_plan** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(out && n_segs >= 0 && (n_segs == 0 || segs), DQ_EINVAL, "bad arguments");
    auto P = std::make_unique<dq_eloc_plan>();
    if (tune) P->tune = *tune;
    dq::validate_tuning(P->tune);
    std::vector<dq::SegDev> h, hv;
    std::vector<int64_t> start{0}, vstart{0};
    std::vector<dq::EnergySeg> es;
    std::vector<int64_t> estart{0};
    std::vector<dq::GatherSeg> gs;
    std::vector<int64_t> gstart{0};
    std::vector<std::pair<const void*, const double*>> gkey;  // (DetMap, psi) gathered once per launch
    for (int i = 0; i < n_segs; ++i) {
      const auto& s = segs[i];
      dq::SegDev d = dq::make_segment(s);
      if (d.method == DQ_ELOC_STRINGS && s.row_end > s.row_begin) {
        const std::pair<const void*, const double*> key{s.vmap, s.d_coeffs};
        size_t gi = 0;
        while (gi < gkey.size() && gkey[gi] != key) ++gi;
        if (gi == gkey.size()) {
          gkey.push_back(key);
          const int64_t nm = d.X.n;
          auto ca = std::make_unique<dq::DevBuf<double>>(std::max<int64_t>(nm, 1));
          auto cb = std::make_unique<dq::DevBuf<double>>(std::max<int64_t>(nm, 1));
          gs.push_back(dq::GatherSeg{d.X.a_row, d.X.b_row, s.d_coeffs, ca->p, cb->p});
          gstart.push_back(gstart.back() + nm);
          P->gbufs.push_back(std::move(ca));
          P->gbufs.push_back(std::move(cb));
        }
        d.C = dq::StrCoef{gs[gi].ca, gs[gi].cb};
      }
      if (s.row_end > s.row_begin) {
        auto& hh = d.method == dq::kElocVv ? hv : h;
        auto& ss = d.method == dq::kElocVv ? vstart : start;
        hh.push_back(d);
        ss.push_back(ss.back() + (s.row_end - s.row_begin));
      }
      if (s.energy_slot >= 0) {
        es.push_back(dq::EnergySeg{s.d_coeffs, s.d_sigma, s.n_rows, s.energy_slot});
        estart.push_back(estart.back() + (s.n_rows + 255) / 256);
      }
    }
    auto& D = P->dev;
    D.methods = 0;
    for (const auto& d : h) D.methods |= d.method == DQ_ELOC_STRINGS ? 1 : 2;
    if (D.methods == 0) D.methods = 1;
    D.n_vseg = static_cast<int>(hv.size());
    D.vtotal = vstart.back();
    P->vsegs.alloc(std::max<size_t>(hv.size(), 1));
    P->vstart.alloc(vstart.size());
    if (!hv.empty()) DQ_HIP(hipMemcpy(P->vsegs.p, hv.data(), hv.size() * sizeof(dq::SegDev), hipMemcpyHostToDevice));
    DQ_HIP(hipMemcpy(P->vstart.p, vstart.data(), vstart.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    D.vsegs = P->vsegs.p;
    D.vstart = P->vstart.p;
    D.nseg = static_cast<int>(h.size());
    D.total = start.back();
    D.n_energy = static_cast<int>(es.size());
    D.n_chunks = estart.back();
    D.n_gather = static_cast<int>(gs.size());
    D.gather_total = gstart.back();
    P->gsegs.alloc(std::max<size_t>(gs.size(), 1));
    P->gstart.alloc(gstart.size());
    if (!gs.empty()) DQ_HIP(hipMemcpy(P->gsegs.p, gs.data(), gs.size() * sizeof(dq::GatherSeg), hipMemcpyHostToDevice));
    DQ_HIP(hipMemcpy(P->gstart.p, gstart.data(), gstart.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    D.gsegs = P->gsegs.p;
    D.gstart = P->gstart.p;
    P->segs.alloc(std::max<size_t>(h.size(), 1));
    P->start.alloc(start.size());
    P->esegs.alloc(std::max<size_t>(es.size(), 1));
    P->estart.alloc(estart.size());
    P->part.alloc(std::max<int64_t>(D.n_chunks, 1));
    P->counter.alloc(1);
    P->err.alloc(1);
    for (auto& d : h) d.err = P->err.p;
    if (!h.empty()) DQ_HIP(hipMemcpy(P->segs.p, h.data(), h.size() * sizeof(dq::SegDev), hipMemcpyHostToDevice));
    DQ_HIP(hipMemcpy(P->start.p, start.data(), start.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if (!es.empty())
      DQ_HIP(hipMemcpy(P->esegs.p, es.data(), es.size() * sizeof(dq::EnergySeg), hipMemcpyHostToDevice));
    DQ_HIP(hipMemcpy(P->estart.p, estart.data(), estart.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    DQ_HIP(hipMemset(P->counter.p, 0, sizeof(unsigned long long)));
    DQ_HIP(hipMemset(P->err.p, 0, sizeof(int)));
    if (P->tune.persistent && P->tune.grid_blocks == 0)
      P->tune.grid_blocks = dq::default_grid_blocks(P->tune.waves_per_block);
    D.segs = P->segs.p;
    D.start = P->start.p;
    D.esegs = P->esegs.p;
    D.estart = P->estart.p;
    D.part = P->part.p;
    D.counter = P->counter.p;
    D.err = P->err.p;
    *out = P.release();
  });
}

int dq_eloc_plan_info(const dq_eloc_plan* plan, int64_t* total_rows, int* n_segments, int* n_energy,
                      dq_eloc_tuning* tune) {
  return dq::guarded([&] {
    DQ_REQUIRE(plan, DQ_EINVAL, "null plan");
    if (total_rows) *total_rows = plan->dev.total;
    if (n_segments) *n_segments = plan->dev.nseg;
    if (n_energy) *n_energy = plan->dev.n_energy;
    if (tune) *tune = plan->tune;
  });
}

int dq_eloc_plan_launch(dq_eloc_plan* plan, double* d_energy, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(plan, DQ_EINVAL, "null plan");
    DQ_REQUIRE(d_energy || plan->dev.n_energy == 0, DQ_EINVAL, "plan has energy slots but d_energy is NULL");
    dq::launch_eloc(plan->dev, plan->tune, d_energy, dq::as_stream(stream));
  });
}

int dq_eloc_plan_capture(dq_eloc_plan* plan, double* d_energy) {
  return dq::guarded([&] {
    DQ_REQUIRE(plan, DQ_EINVAL, "null plan");
    DQ_REQUIRE(d_energy || plan->dev.n_energy == 0, DQ_EINVAL, "plan has energy slots but d_energy is NULL");
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    plan->exec = nullptr;
    plan->graph = nullptr;
    hipStream_t cs = nullptr;  // the null stream cannot be captured
    DQ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      try {
        dq::launch_eloc(plan->dev, plan->tune, d_energy, cs);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        throw;
      }
      e = hipStreamEndCapture(cs, &plan->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    DQ_HIP(e);
  });
}

int dq_eloc_plan_graph_launch(dq_eloc_plan* plan, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(plan && plan->exec, DQ_EINVAL, "plan has no captured graph");
    DQ_HIP(hipGraphLaunch(plan->exec, dq::as_stream(stream)));
  });
}

int dq_eloc_plan_check(dq_eloc_plan* plan, void* stream) {
  return dq::guarded([&] {
    DQ_REQUIRE(plan, DQ_EINVAL, "null plan");
    hipStream_t s = dq::as_stream(stream);
    int flag = 0;
    DQ_HIP(hipMemcpyAsync(&flag, plan->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    if (flag) {
      DQ_HIP(hipMemsetAsync(plan->err.p, 0, sizeof(int), s));
      DQ_HIP(hipStreamSynchronize(s));
      throw dq::Error(DQ_EINVAL, "determinant occupies orbitals at or above n_orb");
    }
  });
}

int dq_eloc_plan_free(dq_eloc_plan* plan) {
  return dq::guarded([&] { delete plan; });
}

}  // extern "C"
