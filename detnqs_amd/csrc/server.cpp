// Serving loop: continuous batching of local-energy requests on one GPU.
//
// Concurrent VMC runs (each with its own integral context) submit requests
// (V determinants, amplitudes psi_V).  A loop thread owned by the server
//   * admits queued requests in arrival order while their device memory fits
//     in a fixed HBM arena (the budget): V words, DetMap, psi and sigma;
//   * stages the admitted request's data on a copy stream (H2D, DetMap build),
//     overlapping the kernels of requests already running;
//   * launches the multi-instance energy kernel over row ranges of every
//     admitted request (up to a quarter of each launch finishes small
//     requests that arrived behind large ones; the rest goes to the oldest
//     requests first);
//   * forms a request's energy in the launch that computes its last rows and
//     returns it when that launch's event completes.
// Integral contexts (integrals + heat-bath tables) stay resident for the life
// of the caller's dq_ctx; V sets submitted with a nonzero v_key stay resident
// in the arena (determinants + DetMap) and are reused by later requests with
// the same (context, v_key) until the arena needs the space.
//
// A V set kept under a v_key also gets its lane-ordered H_VV in the arena
// (vvcache.hpp: the elements the string kernel generates, in the order it adds
// them): a count pass on the copy stream after the string index, then, once
// the loop has read the size, a fill pass.  The V set's requests wait for it,
// and every request on that V with the same screening is then a product over
// the stored elements, which VMC inner steps (same V, new amplitudes) repeat
// many times.
//
// Each row is computed by the same code as dq_variational_energy (or over a
// layout that reproduces it bit for bit), so a request's sigma and energy are
// bit-identical whether it runs alone or in a mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "assemble.hpp"
#include "batch.hpp"
#include "vvcache.hpp"

namespace dq {
namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation (offsets, 256-B granules).
class Arena {
 public:
  void init(size_t bytes) {
    cap_ = align_up(bytes);
    base_.alloc(cap_);
    free_.clear();
    free_[0] = cap_;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char* ptr(size_t off) const { return base_.p + off; }
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  DevBuf<char> base_;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

// Pinned host buffer (asynchronous copies to and from it do not block the host).
template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void reserve(size_t count) {
    if (count <= n) return;
    release();
    DQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

// Lane-ordered H_VV of a V set, in the arena.
enum class Lay { kNone, kCounting, kFilling, kReady, kFailed };
struct VvLayout {
  Lay st = Lay::kNone;
  int n_orb = 0, use_hb = 0;
  double eps1 = 0.0;
  double sthr = 0.0, dmin = 0.0;
  size_t off_at = 0, off_bytes = 0;  // rounds prefix (n + 1 int64) + error word
  size_t el_at = 0, el_bytes = 0;    // col (u32) and val (f64) slots
  int64_t* off = nullptr;
  int* err = nullptr;
  u32* col = nullptr;
  double* val = nullptr;
  int64_t rounds = 0;
  PinnedBuf<int64_t> h_total;        // rounds total, then the error word
  hipEvent_t counted = nullptr, filled = nullptr;
};

// Device-resident V set: determinants + DetMap.
struct VSet {
  const dq_ctx* ctx = nullptr;
  uint64_t key = 0;  // 0: private to one request
  int64_t n = 0;
  size_t off = 0, bytes = 0;
  u64* dets = nullptr;
  u64* keys = nullptr;
  u32* vals = nullptr;
  u64 mask = 0;
  StrIndex X;  // string index (V-driven method)
  int refs = 0;
  uint64_t last_use = 0;
  hipEvent_t ready = nullptr;  // V staged and DetMap built (copy stream)
  VvLayout lay;
  // a request with this screening would use the layout (built or being built)
  bool layout_for(const dq_server_request& in) const {
    return lay.st != Lay::kNone && lay.st != Lay::kFailed && in.method == DQ_ELOC_STRINGS &&
           in.n_orb == lay.n_orb && in.use_hb == lay.use_hb && in.eps1 == lay.eps1;
  }
  bool layout_busy() const { return lay.st == Lay::kCounting || lay.st == Lay::kFilling; }
};

size_t vset_bytes(int64_t n);

// Arena bytes of one request: its V set (unless cached), psi (host input),
// sigma, psi in member order (two copies), the error word.
size_t request_bytes(int64_t n_rows, int coeffs_on_device) {
  const size_t n = static_cast<size_t>(n_rows);
  return vset_bytes(n_rows) + (coeffs_on_device ? 0 : align_up(8 * n)) + 3 * align_up(8 * n) +
         align_up(sizeof(int));
}

size_t vset_bytes(int64_t n) {
  const u64 cap = hash_capacity(n);
  return align_up(16 * static_cast<size_t>(n)) + align_up(4 * cap) + align_up(16 * cap) +
         align_up(str_index_bytes(n)) + align_up(str_index_scratch_bytes(n));
}

enum class St { kQueued, kAdmitted, kDone };

struct Req {
  int64_t ticket = 0;
  dq_server_request in{};
  std::vector<u64> h_dets;      // host copies of host inputs
  std::vector<double> h_coeffs;
  St st = St::kQueued;
  VSet* v = nullptr;
  size_t off = 0, bytes = 0;  // psi (host input) + sigma + err word
  const double* coeffs = nullptr;
  double* sigma = nullptr;
  double* ca = nullptr;  // psi in string-index member order
  double* cb = nullptr;
  int* err = nullptr;
  hipEvent_t ready = nullptr;  // request data staged
  hipEvent_t input = nullptr;  // device inputs produced (caller's stream)
  int64_t next_row = 0;
  double energy = 0.0;
  int status = DQ_OK;
  Clock::time_point t_submit, t_admit, t_first, t_done;
  bool launched = false;
};

struct Slot {  // one launch in flight
  PinnedBuf<unsigned char> host;  // staged launch description
  DevBuf<unsigned char> dev;
  DevBuf<double> part, energy;    // part: chunk sums + one word for the work counter
  PinnedBuf<double> h_energy;
  PinnedBuf<int> h_err;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Req*> finishing;  // requests whose energy this launch forms
};

}  // namespace
}  // namespace dq

struct dq_server {
  dq_server_config cfg{};
  int device = 0;
  hipStream_t compute = nullptr, copy = nullptr;
  dq::Arena arena;
  std::mutex mu;
  std::condition_variable cv_loop, cv_done;
  std::thread loop;
  bool stop = false;
  int64_t next_ticket = 1;
  std::deque<dq::Req*> queued;
  std::vector<dq::Req*> admitted;  // admission order
  std::unordered_map<int64_t, dq::Req*> by_ticket;
  std::list<dq::VSet*> vsets;      // cached (key != 0) and private sets
  std::vector<std::unique_ptr<dq::Slot>> slots;
  std::deque<int> inflight;        // slot indices in launch order
  uint64_t use_clock = 0;
  dq::Clock::time_point t0;
  dq_server_stats stats{};
  std::string fatal;

  void run();
  bool admit_one(dq::Req* r);
  dq::VSet* find_vset(const dq_ctx* ctx, uint64_t key, int64_t n);
  bool evict_one();
  void drop_vset(dq::VSet* v);
  void start_layout(dq::VSet* v, const dq_server_request& in);
  bool advance_layouts();
  void release_request(dq::Req* r);
  bool launch();
  void complete(dq::Slot& s);
  double secs(dq::Clock::time_point t) const { return std::chrono::duration<double>(t - t0).count(); }
};

namespace dq {
namespace {

template <class T>
T* carve(unsigned char*& p, size_t count) {
  T* r = reinterpret_cast<T*>(p);
  p += align_up(count * sizeof(T));
  return r;
}

}  // namespace
}  // namespace dq

dq::VSet* dq_server::find_vset(const dq_ctx* ctx, uint64_t key, int64_t n) {
  if (key == 0) return nullptr;
  for (auto* v : vsets)
    if (v->key == key && v->ctx == ctx && v->n == n) return v;
  return nullptr;
}

void dq_server::drop_vset(dq::VSet* v) {
  arena.release(v->off, v->bytes);
  if (v->lay.off_bytes) arena.release(v->lay.off_at, v->lay.off_bytes);
  if (v->lay.el_bytes) arena.release(v->lay.el_at, v->lay.el_bytes);
  if (v->lay.counted) (void)hipEventDestroy(v->lay.counted);
  if (v->lay.filled) (void)hipEventDestroy(v->lay.filled);
  if (v->ready) (void)hipEventDestroy(v->ready);
  vsets.remove(v);
  delete v;
}

bool dq_server::evict_one() {
  dq::VSet* best = nullptr;
  for (auto* v : vsets)  // a layout still being built is written by the copy stream
    if (v->refs == 0 && !v->layout_busy() && (!best || v->last_use < best->last_use)) best = v;
  if (!best) return false;
  drop_vset(best);
  ++stats.v_evictions;
  return true;
}

// Count pass of a keyed V set's layout (copy stream, after its string index);
// the loop reads the size when `counted` completes (advance_layouts).  No
// layout if its prefix does not fit the arena now.
void dq_server::start_layout(dq::VSet* v, const dq_server_request& in) {
  using namespace dq;
  VvLayout& L = v->lay;
  L.n_orb = in.n_orb;
  L.use_hb = in.use_hb;
  L.eps1 = in.eps1;
  eloc_thresholds(*in.ctx, in.use_hb, in.eps1, &L.sthr, &L.dmin);
  L.off_bytes = align_up(8 * (static_cast<size_t>(v->n) + 1)) + align_up(sizeof(int));
  if (!arena.alloc(L.off_bytes, &L.off_at)) {
    L.off_bytes = 0;
    L.st = Lay::kFailed;
    return;
  }
  unsigned char* p = reinterpret_cast<unsigned char*>(arena.ptr(L.off_at));
  L.off = carve<int64_t>(p, static_cast<size_t>(v->n) + 1);
  L.err = carve<int>(p, 1);
  L.h_total.reserve(2);
  fill_u32(L.err, 1, 0u, copy);
  vv_count(*in.ctx, v->dets, v->n, in.n_orb, L.sthr, L.dmin, v->X, L.off, nullptr, L.err, copy);
  exclusive_scan_i64(L.off, v->n, L.off, copy);
  DQ_HIP(hipMemcpyAsync(L.h_total.p, L.off + v->n, sizeof(int64_t), hipMemcpyDeviceToHost, copy));
  DQ_HIP(hipMemcpyAsync(L.h_total.p + 1, L.err, sizeof(int), hipMemcpyDeviceToHost, copy));
  DQ_HIP(hipEventCreateWithFlags(&L.counted, hipEventDisableTiming));
  DQ_HIP(hipEventRecord(L.counted, copy));
  L.st = Lay::kCounting;
}

// Moves layouts along: counted -> fill pass queued; filled -> ready.
bool dq_server::advance_layouts() {
  using namespace dq;
  bool moved = false;
  for (auto* v : vsets) {
    VvLayout& L = v->lay;
    if (L.st == Lay::kCounting) {
      const hipError_t q = hipEventQuery(L.counted);
      if (q == hipErrorNotReady) continue;
      DQ_HIP(q);
      moved = true;
      const int bad = static_cast<int>(L.h_total.p[1] & 0xFFFFFFFF);
      L.rounds = L.h_total.p[0];
      const size_t slots = static_cast<size_t>(L.rounds) * kWave;
      L.el_bytes = align_up(4 * std::max<size_t>(slots, 1)) + align_up(8 * std::max<size_t>(slots, 1));
      if (bad || !arena.alloc(L.el_bytes, &L.el_at)) {  // bad rows: the string path reports them
        L.el_bytes = 0;
        L.st = Lay::kFailed;
        continue;
      }
      unsigned char* p = reinterpret_cast<unsigned char*>(arena.ptr(L.el_at));
      L.col = carve<u32>(p, std::max<size_t>(slots, 1));
      L.val = carve<double>(p, std::max<size_t>(slots, 1));
      vv_fill(*v->ctx, v->dets, v->n, L.n_orb, L.sthr, L.dmin, v->X, L.off, L.col, L.val, copy);
      DQ_HIP(hipEventCreateWithFlags(&L.filled, hipEventDisableTiming));
      DQ_HIP(hipEventRecord(L.filled, copy));
      L.st = Lay::kFilling;
    } else if (L.st == Lay::kFilling) {
      const hipError_t q = hipEventQuery(L.filled);
      if (q == hipErrorNotReady) continue;
      DQ_HIP(q);
      L.st = Lay::kReady;
      ++stats.vv_builds;
      moved = true;
    }
  }
  return moved;
}

// Reserves device memory for r and stages its data; false if it does not fit now.
bool dq_server::admit_one(dq::Req* r) {
  using namespace dq;
  const auto& in = r->in;
  if (r->input) DQ_HIP(hipStreamWaitEvent(copy, r->input, 0));
  VSet* v = find_vset(in.ctx, in.v_key, in.n_rows);
  const bool hit = v != nullptr;
  const size_t rbytes = request_bytes(in.n_rows, in.coeffs_on_device) - vset_bytes(in.n_rows);
  const size_t vbytes = hit ? 0 : vset_bytes(in.n_rows);
  size_t voff = 0, roff = 0;
  while (true) {
    bool ok_v = hit || arena.alloc(vbytes, &voff);
    bool ok_r = ok_v && arena.alloc(rbytes, &roff);
    if (ok_r) break;
    if (ok_v && !hit) arena.release(voff, vbytes);
    if (!evict_one()) return false;
  }
  if (!hit) {
    v = new VSet();
    v->ctx = in.ctx;
    v->key = in.v_key;
    v->n = in.n_rows;
    v->off = voff;
    v->bytes = vbytes;
    unsigned char* p = reinterpret_cast<unsigned char*>(arena.ptr(voff));
    const u64 cap = hash_capacity(in.n_rows);
    v->dets = carve<u64>(p, 2 * static_cast<size_t>(in.n_rows));
    v->vals = carve<u32>(p, cap);
    v->keys = carve<u64>(p, 2 * cap);
    v->mask = cap - 1;
    unsigned char* xmem = carve<unsigned char>(p, str_index_bytes(in.n_rows));
    unsigned char* xscr = carve<unsigned char>(p, str_index_scratch_bytes(in.n_rows));
    DQ_HIP(hipEventCreateWithFlags(&v->ready, hipEventDisableTiming));
    const size_t db = 16 * static_cast<size_t>(in.n_rows);
    if (db) {
      if (in.dets_on_device)
        DQ_HIP(hipMemcpyAsync(v->dets, in.dets, db, hipMemcpyDeviceToDevice, copy));
      else
        DQ_HIP(hipMemcpyAsync(v->dets, r->h_dets.data(), db, hipMemcpyHostToDevice, copy));
    }
    hash_build_async(v->dets, in.n_rows, v->keys, v->vals, v->mask, copy);
    v->X = str_index_build(v->dets, in.n_rows, HashView{v->keys, v->vals, v->mask}, xmem, xscr, copy, v->vals);
    DQ_HIP(hipEventRecord(v->ready, copy));
    vsets.push_back(v);
    if (in.v_key) ++stats.v_misses;
    if (cfg.vv_layout && in.v_key && in.method == DQ_ELOC_STRINGS && in.n_rows > 0) start_layout(v, in);
  } else {
    ++stats.v_hits;
  }
  v->refs++;
  v->last_use = ++use_clock;
  r->v = v;
  r->off = roff;
  r->bytes = rbytes;
  unsigned char* p = reinterpret_cast<unsigned char*>(arena.ptr(roff));
  if (in.coeffs_on_device) {
    r->coeffs = in.coeffs;
  } else {
    double* c = carve<double>(p, in.n_rows);
    if (in.n_rows) DQ_HIP(hipMemcpyAsync(c, r->h_coeffs.data(), 8 * in.n_rows, hipMemcpyHostToDevice, copy));
    r->coeffs = c;
  }
  r->sigma = carve<double>(p, in.n_rows);
  r->ca = carve<double>(p, in.n_rows);
  r->cb = carve<double>(p, in.n_rows);
  r->err = carve<int>(p, 1);
  DQ_HIP(hipMemsetAsync(r->err, 0, sizeof(int), copy));
  DQ_HIP(hipEventCreateWithFlags(&r->ready, hipEventDisableTiming));
  DQ_HIP(hipEventRecord(r->ready, copy));
  r->st = St::kAdmitted;
  r->t_admit = Clock::now();
  return true;
}

void dq_server::release_request(dq::Req* r) {
  arena.release(r->off, r->bytes);
  if (r->ready) (void)hipEventDestroy(r->ready);
  if (r->input) (void)hipEventDestroy(r->input);
  r->ready = nullptr;
  r->input = nullptr;
  dq::VSet* v = r->v;
  r->v = nullptr;
  if (v && --v->refs == 0 && v->key == 0) drop_vset(v);
}

// Builds and enqueues one launch over the admitted requests; false if idle.
bool dq_server::launch() {
  using namespace dq;
  std::vector<Req*> run;
  for (auto* r : admitted)  // requests whose layout is being built wait for it
    if ((r->next_row < r->in.n_rows || !r->launched) && !(r->v->layout_busy() && r->v->layout_for(r->in)))
      run.push_back(r);
  if (run.empty()) return false;
  int si = -1;
  for (int k = 0; k < static_cast<int>(slots.size()); ++k)
    if (!slots[k]->busy) {
      si = k;
      break;
    }
  if (si < 0) return false;
  Slot& S = *slots[si];

  // Shares of rows_per_launch: requests behind the oldest whose remaining rows
  // fit in what is left of a quarter of the launch finish in it (a small
  // request that arrives behind large ones is done within a launch or two);
  // the rest go to the requests in arrival order (the large ones keep close to
  // first come, first served).  Equal shares of everything stretched the large
  // requests by the number running.
  const int64_t R = std::max<int64_t>(cfg.rows_per_launch, 1);
  std::vector<int64_t> take(run.size(), 0);
  int64_t left = R;
  int64_t small = R / 4;
  for (size_t k = 1; k < run.size() && small > 0; ++k) {
    const int64_t rem = run[k]->in.n_rows - run[k]->next_row;
    if (rem <= small) {
      take[k] = rem;
      small -= rem;
      left -= rem;
    }
  }
  for (size_t k = 0; k < run.size() && left > 0; ++k) {
    const int64_t t = std::min(run[k]->in.n_rows - run[k]->next_row - take[k], left);
    take[k] += t;
    left -= t;
  }

  std::vector<SegDev> segs, vsegs;
  std::vector<int64_t> start{0}, vstart{0};
  std::vector<EnergySeg> es;
  std::vector<int64_t> estart{0};
  S.finishing.clear();
  for (size_t k = 0; k < run.size(); ++k) {
    Req* r = run[k];
    const bool first = !r->launched;
    const bool last = r->next_row + take[k] == r->in.n_rows;
    if (take[k] == 0 && !(first && last)) continue;  // no share this time (n_rows > 0)
    const bool lay = r->v->lay.st == Lay::kReady && r->v->layout_for(r->in);
    if (first) {
      DQ_HIP(hipStreamWaitEvent(compute, r->v->ready, 0));
      DQ_HIP(hipStreamWaitEvent(compute, r->ready, 0));
      if (r->in.method == DQ_ELOC_STRINGS && !lay) str_gather_coef(r->v->X, r->coeffs, r->ca, r->cb, compute);
      r->launched = true;
      r->t_first = Clock::now();
    }
    if (take[k] > 0 && lay) {  // a product over the stored elements (the layout is complete: polled)
      SegDev d{};
      d.coeffs = r->coeffs;
      d.sigma = r->sigma;
      d.row_begin = r->next_row;
      d.method = kElocVv;
      d.vv_off = r->v->lay.off;
      d.vv_col = r->v->lay.col;
      d.vv_val = r->v->lay.val;
      vsegs.push_back(d);
      vstart.push_back(vstart.back() + take[k]);
      r->next_row += take[k];
      stats.vv_rows += take[k];
    } else if (take[k] > 0) {
      SegDev d{};
      auto* c = const_cast<dq_ctx*>(r->in.ctx);
      d.V = c->view();
      d.H = HashView{r->v->keys, r->v->vals, r->v->mask};
      if (r->in.n_rows == 0) d.H = HashView{};
      d.dets = r->v->dets;
      d.coeffs = r->coeffs;
      d.sigma = r->sigma;
      d.hb_len = r->in.use_hb ? c->hb_lengths(r->in.eps1, compute) : nullptr;
      d.err = r->err;
      d.X = r->v->X;
      d.C = StrCoef{r->ca, r->cb};
      d.row_begin = r->next_row;
      eloc_thresholds(*c, r->in.use_hb, r->in.eps1, &d.sthr, &d.dmin);
      d.n_orb = r->in.n_orb;
      d.dbl_mode = r->in.use_hb ? 1 : 0;
      d.method = r->in.method;
      segs.push_back(d);
      start.push_back(start.back() + take[k]);
      r->next_row += take[k];
    }
    if (last) {
      es.push_back(EnergySeg{r->coeffs, r->sigma, r->in.n_rows, static_cast<int>(S.finishing.size())});
      estart.push_back(estart.back() + (r->in.n_rows + 255) / 256);
      S.finishing.push_back(r);
    }
  }
  if (segs.empty() && vsegs.empty() && es.empty()) return false;

  // Stage the launch description: segs | start | vsegs | vstart | esegs | estart.
  const size_t nseg = segs.size(), nv = vsegs.size(), ne = es.size();
  const size_t bytes = align_up(nseg * sizeof(SegDev)) + align_up((nseg + 1) * 8) + align_up(nv * sizeof(SegDev)) +
                       align_up((nv + 1) * 8) + align_up(ne * sizeof(EnergySeg)) + align_up((ne + 1) * 8);
  // The slot is idle (its last launch completed), so its buffers may be reused.
  S.host.reserve(bytes);
  if (S.dev.n < bytes) S.dev.alloc(bytes);
  const int64_t n_chunks = estart.back();
  if (S.part.n < static_cast<size_t>(n_chunks) + 1) S.part.alloc(static_cast<size_t>(n_chunks) + 1);
  if (S.energy.n < std::max<size_t>(ne, 1)) S.energy.alloc(std::max<size_t>(ne, 1));
  S.h_energy.reserve(std::max<size_t>(ne, 1));
  S.h_err.reserve(std::max<size_t>(ne, 1));
  unsigned char* hp = S.host.p;
  unsigned char* dp = S.dev.p;
  auto put = [&](const void* src, size_t b) {
    if (b) std::memcpy(hp, src, b);
    unsigned char* d = dp;
    hp += align_up(b);
    dp += align_up(b);
    return d;
  };
  PlanDev P;
  P.segs = reinterpret_cast<const SegDev*>(put(segs.data(), nseg * sizeof(SegDev)));
  P.start = reinterpret_cast<const int64_t*>(put(start.data(), (nseg + 1) * 8));
  P.vsegs = reinterpret_cast<const SegDev*>(put(vsegs.data(), nv * sizeof(SegDev)));
  P.vstart = reinterpret_cast<const int64_t*>(put(vstart.data(), (nv + 1) * 8));
  P.n_vseg = static_cast<int>(nv);
  P.vtotal = vstart.back();
  P.esegs = reinterpret_cast<const EnergySeg*>(put(es.data(), ne * sizeof(EnergySeg)));
  P.estart = reinterpret_cast<const int64_t*>(put(estart.data(), (ne + 1) * 8));
  P.part = S.part.p;
  P.counter = reinterpret_cast<unsigned long long*>(S.part.p + n_chunks);
  P.methods = 0;
  for (const auto& d : segs) P.methods |= d.method == DQ_ELOC_STRINGS ? 1 : 2;
  if (P.methods == 0) P.methods = 1;
  P.nseg = static_cast<int>(nseg);
  P.total = start.back();
  P.n_energy = static_cast<int>(ne);
  P.n_chunks = n_chunks;
  DQ_HIP(hipMemcpyAsync(S.dev.p, S.host.p, bytes, hipMemcpyHostToDevice, compute));
  launch_eloc(P, cfg.tune, S.energy.p, compute);
  if (ne) DQ_HIP(hipMemcpyAsync(S.h_energy.p, S.energy.p, ne * sizeof(double), hipMemcpyDeviceToHost, compute));
  for (size_t k = 0; k < ne; ++k)
    DQ_HIP(hipMemcpyAsync(S.h_err.p + k, S.finishing[k]->err, sizeof(int), hipMemcpyDeviceToHost, compute));
  if (!S.done) DQ_HIP(hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
  DQ_HIP(hipEventRecord(S.done, compute));
  S.busy = true;
  inflight.push_back(si);
  stats.launches++;
  stats.rows += P.total + P.vtotal;
  return true;
}

void dq_server::complete(dq::Slot& S) {
  const auto now = dq::Clock::now();
  for (size_t k = 0; k < S.finishing.size(); ++k) {
    dq::Req* r = S.finishing[k];
    r->energy = S.h_energy.p[k];
    r->status = S.h_err.p[k] ? DQ_EINVAL : DQ_OK;
    r->t_done = now;
    r->st = dq::St::kDone;
    admitted.erase(std::find(admitted.begin(), admitted.end(), r));
    release_request(r);
    stats.completed++;
  }
  S.finishing.clear();
  S.busy = false;
}

void dq_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    bool progressed = false;
    try {
      // completions, in launch order
      while (!inflight.empty()) {
        dq::Slot& S = *slots[inflight.front()];
        const hipError_t q = hipEventQuery(S.done);
        if (q == hipErrorNotReady) break;
        DQ_HIP(q);
        complete(S);
        inflight.pop_front();
        progressed = true;
        cv_done.notify_all();
      }
      if (advance_layouts()) progressed = true;
      // admission in arrival order while memory allows
      while (!queued.empty() && admit_one(queued.front())) {
        admitted.push_back(queued.front());
        queued.pop_front();
        progressed = true;
      }
      // keep up to max_inflight launches queued on the device
      while (static_cast<int>(inflight.size()) < cfg.max_inflight && launch()) progressed = true;
    } catch (const std::exception& e) {
      fatal = e.what();
      for (auto& kv : by_ticket) {
        if (kv.second->st != dq::St::kDone) {
          kv.second->status = DQ_ERUNTIME;
          kv.second->st = dq::St::kDone;
        }
      }
      cv_done.notify_all();
      return;
    }
    const bool idle = inflight.empty() && queued.empty() && admitted.empty();
    if (stop && idle) break;
    if (progressed) continue;
    if (idle) {
      cv_loop.wait(lk, [&] { return stop || !queued.empty(); });
    } else {
      // GPU busy: poll completions without holding the lock
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      lk.lock();
    }
  }
}

// ----------------------------------------------------------------------------- C-ABI
extern "C" {

int dq_server_create(const dq_server_config* cfg, dq_server** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(cfg && out, DQ_EINVAL, "null argument");
    DQ_REQUIRE(cfg->hbm_budget_bytes > 0, DQ_EINVAL, "hbm_budget_bytes must be positive");
    DQ_REQUIRE(cfg->rows_per_launch > 0, DQ_EINVAL, "rows_per_launch must be positive");
    DQ_REQUIRE(cfg->max_inflight >= 1 && cfg->max_inflight <= 8, DQ_EINVAL, "max_inflight must be in [1, 8]");
    dq::validate_tuning(cfg->tune);
    auto s = std::make_unique<dq_server>();
    s->cfg = *cfg;
    DQ_HIP(hipGetDevice(&s->device));
    if (s->cfg.tune.persistent && s->cfg.tune.grid_blocks == 0)
      s->cfg.tune.grid_blocks = dq::default_grid_blocks(s->cfg.tune.waves_per_block);
    s->arena.init(static_cast<size_t>(cfg->hbm_budget_bytes));
    DQ_HIP(hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking));
    DQ_HIP(hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking));
    for (int k = 0; k < cfg->max_inflight; ++k) s->slots.push_back(std::make_unique<dq::Slot>());
    s->t0 = dq::Clock::now();
    dq_server* raw = s.get();
    s->loop = std::thread([raw] { raw->run(); });
    *out = s.release();
  });
}

int dq_server_request_bytes(int64_t n_rows, int coeffs_on_device, int64_t* bytes) {
  return dq::guarded([&] {
    DQ_REQUIRE(bytes && n_rows >= 0, DQ_EINVAL, "bad arguments");
    const size_t n = static_cast<size_t>(n_rows);
    (void)n;
    *bytes = static_cast<int64_t>(dq::request_bytes(n_rows, coeffs_on_device));
  });
}

int dq_server_submit(dq_server* s, const dq_server_request* req, int64_t* ticket) {
  return dq::guarded([&] {
    DQ_REQUIRE(s && req && ticket, DQ_EINVAL, "null argument");
    const auto& in = *req;
    DQ_REQUIRE(in.ctx, DQ_EINVAL, "null context");
    DQ_REQUIRE(in.n_rows >= 0 && (in.n_rows == 0 || (in.dets && in.coeffs)), DQ_EINVAL, "bad arguments");
    DQ_REQUIRE(in.n_orb >= 1 && in.n_orb <= in.ctx->n_orb, DQ_EINVAL,
               "n_orb must be in [1, n_orb of the integral context]");
    if (in.use_hb)
      DQ_REQUIRE(in.ctx->hb, DQ_EINVAL, "compute_variational_energy: Heat-Bath mode requires non-null hb_table");
    DQ_REQUIRE(in.method == DQ_ELOC_STRINGS || in.method == DQ_ELOC_WALK, DQ_EINVAL, "unknown energy method");
    const size_t need = dq::request_bytes(in.n_rows, in.coeffs_on_device);
    DQ_REQUIRE(need <= s->arena.capacity(), DQ_ERANGE, "request does not fit in the server's HBM budget");
    auto r = std::make_unique<dq::Req>();
    r->in = in;
    if (!in.dets_on_device && in.n_rows) r->h_dets.assign(in.dets, in.dets + 2 * in.n_rows);
    if (!in.coeffs_on_device && in.n_rows) r->h_coeffs.assign(in.coeffs, in.coeffs + in.n_rows);
    if (in.dets_on_device || in.coeffs_on_device) {
      DQ_HIP(hipEventCreateWithFlags(&r->input, hipEventDisableTiming));
      DQ_HIP(hipEventRecord(r->input, dq::as_stream(in.stream)));
    }
    r->t_submit = dq::Clock::now();
    std::lock_guard<std::mutex> g(s->mu);
    DQ_REQUIRE(s->fatal.empty(), DQ_ERUNTIME, "server stopped: " + s->fatal);
    DQ_REQUIRE(!s->stop, DQ_EINVAL, "server is shutting down");
    // cache before the loop needs it (the loop's stream waits for the upload)
    if (in.use_hb) const_cast<dq_ctx*>(in.ctx)->hb_lengths(in.eps1, s->copy);
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->by_ticket[r->ticket] = r.get();
    s->queued.push_back(r.release());
    s->stats.submitted++;
    s->cv_loop.notify_one();
  });
}

int dq_server_wait(dq_server* s, int64_t ticket, double timeout_s, dq_server_result* res) {
  return dq::guarded([&] {
    DQ_REQUIRE(s && res, DQ_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->by_ticket.find(ticket);
    DQ_REQUIRE(it != s->by_ticket.end(), DQ_EINVAL, "unknown ticket");
    dq::Req* r = it->second;
    auto ready = [&] { return r->st == dq::St::kDone; };
    if (timeout_s < 0) {
      s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
      throw dq::Error(DQ_ETIMEDOUT, "request not finished within the timeout");
    }
    res->energy = r->energy;
    res->status = r->status;
    res->t_submit = s->secs(r->t_submit);
    res->t_admit = s->secs(r->t_admit);
    res->t_first_launch = s->secs(r->t_first);
    res->t_done = s->secs(r->t_done);
    s->by_ticket.erase(it);
    const int st = r->status;
    delete r;
    if (st == DQ_EINVAL) throw dq::Error(DQ_EINVAL, "determinant occupies orbitals at or above n_orb");
    if (st != DQ_OK) throw dq::Error(st, "server stopped: " + s->fatal);
  });
}

int dq_server_stats_get(dq_server* s, dq_server_stats* out) {
  return dq::guarded([&] {
    DQ_REQUIRE(s && out, DQ_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    out->arena_bytes = static_cast<int64_t>(s->arena.capacity());
    out->arena_used = static_cast<int64_t>(s->arena.used());
    out->arena_peak = static_cast<int64_t>(s->arena.peak());
    out->queued = static_cast<int64_t>(s->queued.size());
    out->running = static_cast<int64_t>(s->admitted.size());
    int64_t vc = 0;
    for (auto* v : s->vsets)
      if (v->key) ++vc;
    out->v_cached = vc;
  });
}

int dq_server_destroy(dq_server* s) {
  return dq::guarded([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stop = true;
      s->cv_loop.notify_one();
    }
    if (s->loop.joinable()) s->loop.join();
    (void)hipStreamSynchronize(s->compute);
    (void)hipStreamSynchronize(s->copy);
    for (auto& kv : s->by_ticket) delete kv.second;
    while (!s->vsets.empty()) s->drop_vset(s->vsets.front());
    for (auto& sl : s->slots)
      if (sl->done) (void)hipEventDestroy(sl->done);
    (void)hipStreamDestroy(s->compute);
    (void)hipStreamDestroy(s->copy);
    delete s;
  });
}

}  // extern "C"
