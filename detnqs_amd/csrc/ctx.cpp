// Integral context: FCIDUMP parsing, dense gather tables, heat-bath table.
#include "ctx.hpp"

#include <omp.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

namespace dq {

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

struct CachedBlock {
  void* p;
  size_t bytes;
  int dev;
  hipEvent_t ev;
};

std::mutex g_cache_mu;
std::vector<CachedBlock> g_cache;  // oldest release first

// Bytes the cache may keep per device: a quarter of the device's memory.
size_t cache_cap(int dev) {
  static size_t cap[64] = {};
  if (dev < 0 || dev >= 64) return 0;
  if (!cap[dev]) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      return 0;
    }
    cap[dev] = total_b / 4;
  }
  return cap[dev];
}

constexpr size_t kCacheGranule = size_t(2) << 20;

size_t round_up(size_t b) { return (b + kCacheGranule - 1) / kCacheGranule * kCacheGranule; }

// Frees the cached blocks of device `dev` (all devices when dev < 0); the
// caller holds the lock.
void drop_cached(int dev) {
  std::vector<CachedBlock> keep;
  for (auto& b : g_cache) {
    if (dev >= 0 && b.dev != dev) {
      keep.push_back(b);
      continue;
    }
    if (b.ev) {
      (void)hipEventSynchronize(b.ev);
      (void)hipEventDestroy(b.ev);
    }
    (void)hipFree(b.p);
  }
  g_cache.swap(keep);
}

void keep_block(const CachedBlock& blk);

}  // namespace

void* cache_alloc(size_t bytes, hipStream_t s) {
  int dev = 0;
  DQ_HIP(hipGetDevice(&dev));
  const size_t want = round_up(bytes);
  {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    size_t best = g_cache.size();
    for (size_t i = 0; i < g_cache.size(); ++i) {
      const auto& b = g_cache[i];
      if (b.dev == dev && b.bytes >= want && b.bytes <= 2 * want &&
          (best == g_cache.size() || b.bytes < g_cache[best].bytes))
        best = i;
    }
    if (best < g_cache.size()) {
      CachedBlock b = g_cache[best];
      g_cache.erase(g_cache.begin() + static_cast<std::ptrdiff_t>(best));
      if (b.ev) {
        const hipError_t e = s ? hipStreamWaitEvent(s, b.ev, 0) : hipEventSynchronize(b.ev);
        (void)hipEventDestroy(b.ev);
        if (e != hipSuccess) {
          (void)hipFree(b.p);
          throw Error(DQ_ERUNTIME, std::string("waiting for a cached block: ") + hipGetErrorString(e));
        }
      }
      return b.p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lock(g_cache_mu);
      drop_cached(dev);
    }
    DQ_HIP(hipMalloc(&p, want));
  }
  return p;
}

void cache_release(void* p, size_t bytes, hipStream_t s) {
  if (!p) return;
  int dev = 0;
  hipEvent_t ev = nullptr;
  if (hipGetDevice(&dev) != hipSuccess || hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
      hipEventRecord(ev, s) != hipSuccess) {
    // cannot order a later reuse after this release: wait and free instead
    if (ev) (void)hipEventDestroy(ev);
    (void)hipStreamSynchronize(s);
    (void)hipFree(p);
    return;
  }
  keep_block(CachedBlock{p, round_up(bytes), dev, ev});
}

void cache_release_idle(void* p, size_t bytes) {
  if (!p) return;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipFree(p);
    return;
  }
  keep_block(CachedBlock{p, round_up(bytes), dev, nullptr});
}

namespace {

void keep_block(const CachedBlock& blk) {
  const int dev = blk.dev;
  const size_t cap = cache_cap(dev);
  std::lock_guard<std::mutex> lock(g_cache_mu);
  g_cache.push_back(blk);
  size_t held = 0;
  for (const auto& b : g_cache)
    if (b.dev == dev) held += b.bytes;
  // over the cap: return the oldest blocks of this device (after their work)
  for (size_t i = 0; held > cap && i < g_cache.size();) {
    if (g_cache[i].dev != dev) {
      ++i;
      continue;
    }
    if (g_cache[i].ev) {
      (void)hipEventSynchronize(g_cache[i].ev);
      (void)hipEventDestroy(g_cache[i].ev);
    }
    (void)hipFree(g_cache[i].p);
    held -= g_cache[i].bytes;
    g_cache.erase(g_cache.begin() + static_cast<std::ptrdiff_t>(i));
  }
}

}  // namespace

void cache_trim() {
  std::lock_guard<std::mutex> lock(g_cache_mu);
  drop_cached(-1);
}

namespace {

// FCIDUMP reader with the reference's rules (integral_mo.cpp:38-182):
// the &FCI namelist gives NORB (must match) and NELEC; integral lines are
// "value p q r s" (1-based), "0 0 0 0" is e_nuc, "p q 0 0" is h_pq; lines
// starting with ! * C c # or containing '&' are skipped; Fortran D exponents
// are accepted; later lines overwrite earlier ones.
void parse_fcidump(dq_ctx& c, const std::string& path) {
  std::ifstream f(path);
  DQ_REQUIRE(f.good(), DQ_ERUNTIME, "Cannot open FCIDUMP: " + path);
  std::string line, nl;
  bool in = false;
  while (std::getline(f, line)) {
    if (!in && (line.find("&FCI") != std::string::npos || line.find("&fci") != std::string::npos)) in = true;
    if (in) {
      nl += line + " ";
      if (line.find("&END") != std::string::npos || line.find("&end") != std::string::npos) break;
    }
  }
  DQ_REQUIRE(in, DQ_ERUNTIME, "No &FCI namelist found in FCIDUMP");
  std::string up = nl;
  for (auto& ch : up) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
  auto param = [&](const char* key, int& v) {
    std::string k = std::string(key) + "=";
    size_t pos = up.find(k);
    if (pos == std::string::npos) return false;
    pos += k.size();
    while (pos < nl.size() && std::isspace(static_cast<unsigned char>(nl[pos]))) ++pos;
    size_t e = pos;
    while (e < nl.size() && (std::isdigit(static_cast<unsigned char>(nl[e])) || nl[e] == '-' || nl[e] == '+')) ++e;
    if (e == pos) return false;
    v = std::stoi(nl.substr(pos, e - pos));
    return true;
  };
  int norb = 0;
  if (param("NORB", norb) && norb != c.n_orb)
    throw Error(DQ_ERUNTIME, "NORB mismatch: expected " + std::to_string(c.n_orb) + ", got " +
                                 std::to_string(norb));
  DQ_REQUIRE(param("NELEC", c.n_elec) && c.n_elec >= 0, DQ_ERUNTIME, "Invalid or missing NELEC in namelist");
  f.clear();
  f.seekg(0);
  const int n = c.n_orb;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    const char c0 = line[0];
    if (c0 == '!' || c0 == '*' || c0 == 'C' || c0 == 'c' || c0 == '#') continue;
    if (line.find('&') != std::string::npos) continue;
    for (auto& ch : line)
      if (ch == 'D' || ch == 'd') ch = 'E';
    std::istringstream is(line);
    double v;
    int p, q, r, s;
    if (!(is >> v >> p >> q >> r >> s)) continue;
    if (p == 0 && q == 0 && r == 0 && s == 0) {
      c.e_nuc = v;
    } else if (r == 0 && s == 0) {
      DQ_REQUIRE(p >= 1 && q >= 1 && p <= n && q <= n, DQ_ERANGE, "FCIDUMP index out of range");
      c.h1e[key1(p - 1, q - 1)] = v;
    } else {
      DQ_REQUIRE(p >= 1 && q >= 1 && r >= 1 && s >= 1 && p <= n && q <= n && r <= n && s <= n, DQ_ERANGE,
                 "FCIDUMP index out of range");
      c.h2e[key2(p - 1, q - 1, r - 1, s - 1)] = v;
    }
  }
}

void alloc_packed(dq_ctx& c, int n_orb) {
  DQ_REQUIRE(n_orb >= 1 && n_orb <= 64, DQ_EINVAL, "Number of orbitals must be in [1, 64]");
  c.n_orb = n_orb;
  const size_t n1 = static_cast<size_t>(n_orb) * (n_orb + 1) / 2;
  c.h1e.assign(n1, 0.0);
  c.h2e.assign(n1 * (n1 + 1) / 2, 0.0);
}

template <class T>
void upload(DevBuf<T>& d, const std::vector<T>& h) {
  d.alloc(h.size());
  if (!h.empty()) DQ_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace

void upload_tables(dq_ctx& c) {
  c.d_err.alloc(1);
  DQ_HIP(hipMemset(c.d_err.p, 0, sizeof(int)));
  const int n = c.n_orb;
  const size_t n2 = static_cast<size_t>(n) * n, n3 = n2 * n;
  std::vector<double> h1(n2), Js(n3), Ks(n3), Jd(n2), Kd(n2);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) {
      h1[p * n + q] = c.g1(p, q);
      Jd[p * n + q] = c.g2(p, p, q, q);
      Kd[p * n + q] = c.g2(p, q, q, p);
    }
  for (int r = 0; r < n; ++r)
    for (int k = 0; k < n; ++k)
      for (int a = 0; a < n; ++a) {
        Js[(static_cast<size_t>(r) * n + k) * n + a] = c.g2(r, a, k, k);
        Ks[(static_cast<size_t>(r) * n + k) * n + a] = c.g2(r, k, k, a);
      }
  // pair masks (see DevView::pm)
  const int nso = 2 * n;
  const size_t nrows = static_cast<size_t>(nso) * (nso - 1) / 2;
  std::vector<u64> pm(nrows * n, 0);
  c.pm_cnt.assign(nrows, 0);
#pragma omp parallel for schedule(dynamic)
  for (int i = 1; i < nso; ++i) {
    for (int j = 0; j < i; ++j) {
      const size_t r = static_cast<size_t>(i) * (i - 1) / 2 + j;
      const int mi = i >> 1, mj = j >> 1;
      u32 cnt = 0;
      if ((i & 1) == (j & 1)) {
        const int r0 = mj < mi ? mj : mi, r1 = mj < mi ? mi : mj;
        for (int a = 0; a < n; ++a) {
          if (a == mi || a == mj) continue;
          for (int b = 0; b < a; ++b) {
            if (b == mi || b == mj) continue;
            const double V = c.g2(r0, b, r1, a) - c.g2(r0, a, r1, b);  // added b < a
            if (std::abs(V) > kMatThr) {
              pm[r * n + a] |= 1ULL << b;
              ++cnt;
            }
          }
        }
      } else {
        const int ra = (i & 1) ? mj : mi, rb = (i & 1) ? mi : mj;
        for (int a = 0; a < n; ++a) {
          if (a == ra) continue;
          for (int b = 0; b < n; ++b) {
            if (b == rb) continue;
            if (std::abs(c.g2(ra, a, rb, b)) > kMatThr) {
              pm[r * n + a] |= 1ULL << b;
              ++cnt;
            }
          }
        }
      }
      c.pm_cnt[r] = cnt;
    }
  }
  upload(c.d_pm, pm);
  upload(c.d_pm_cnt, c.pm_cnt);
  upload(c.d_h1, h1);
  upload(c.d_h2, c.h2e);
  upload(c.d_Js, Js);
  upload(c.d_Ks, Ks);
  upload(c.d_Jd, Jd);
  upload(c.d_Kd, Kd);
}

// Heat-bath table (hb_table.cpp:39-168).  Row {i>j} lists the virtual pairs
// (a>b), disjoint from {i,j} and spin-compatible, with w = |<ij||ab>| >= thr,
// sorted by descending w with std::sort; the reference sorts the same input
// sequence with std::ranges::sort (the same libstdc++ introsort), so equal
// weights come out in the reference's order.
void build_hb(dq_ctx& c, double thr) {
  const int nso = 2 * c.n_orb;
  DQ_REQUIRE(nso >= 2, DQ_EINVAL, "HeatBathTable requires n_so >= 2");
  const size_t nrows = static_cast<size_t>(nso) * (nso - 1) / 2;
  struct E {
    int a, b;
    double w;
  };
  auto phys = [&](int mu, int ka, int nu, int la) -> double {  // integral_so.hpp:89-97
    if ((mu & 1) != (nu & 1) || (ka & 1) != (la & 1)) return 0.0;
    return c.g2(mu >> 1, nu >> 1, ka >> 1, la >> 1);
  };
  std::vector<std::vector<E>> rows(nrows);
#pragma omp parallel for schedule(dynamic)
  for (int i = 1; i < nso; ++i) {
    for (int j = 0; j < i; ++j) {
      std::vector<E> buf;
      const int si = i & 1, sj = j & 1;
      for (int a = 1; a < nso; ++a)
        for (int b = 0; b < a; ++b) {
          if (a == i || a == j || b == i || b == j) continue;
          const int sa = a & 1, sb = b & 1;
          if (!((sa == si && sb == sj) || (sa == sj && sb == si))) continue;
          const double w = std::abs(phys(i, j, a, b) - phys(i, j, b, a));
          if (w >= thr) buf.push_back({a, b, w});
        }
      std::sort(buf.begin(), buf.end(), [](const E& x, const E& y) { return x.w > y.w; });
      rows[static_cast<size_t>(i) * (i - 1) / 2 + j] = std::move(buf);
    }
  }
  auto hb = std::make_unique<HostHB>();
  hb->nso = nso;
  hb->off.assign(nrows + 1, 0);
  size_t tot = 0;
  for (size_t r = 0; r < nrows; ++r) {
    hb->off[r] = static_cast<u32>(tot);
    tot += rows[r].size();
    DQ_REQUIRE(tot < 0xFFFFFFFFull, DQ_ERUNTIME, "heat-bath table exceeds 2^32 entries");
  }
  hb->off[nrows] = static_cast<u32>(tot);
  hb->a.resize(tot);
  hb->b.resize(tot);
  hb->w.resize(tot);
  hb->v.resize(tot);
  std::vector<u32> ab(tot);
#pragma omp parallel for schedule(dynamic)
  for (int i = 1; i < nso; ++i) {
    for (int j = 0; j < i; ++j) {
      const size_t r = static_cast<size_t>(i) * (i - 1) / 2 + j;
      const size_t o = hb->off[r];
      for (size_t k = 0; k < rows[r].size(); ++k) {
        const int a = rows[r][k].a, b = rows[r][k].b;
        hb->a[o + k] = a;
        hb->b[o + k] = b;
        hb->w[o + k] = rows[r][k].w;
        ab[o + k] = static_cast<u32>(a) | (static_cast<u32>(b) << 16);
        // Value HamEval::compute_double returns for removing {i,j} and adding
        // {a,b} (ham_eval.cpp:117-147): with R = removed, A = added (MO,
        // ascending within a spin), same spin: (R0 A0|R1 A1) - (R0 A1|R1 A0);
        // mixed: (Ra Aa|Rb Ab).
        double V;
        if ((i & 1) == (j & 1)) {
          const int r0 = j >> 1, r1 = i >> 1, a0 = b >> 1, a1 = a >> 1;  // i>j, a>b, same spin
          V = c.g2(r0, a0, r1, a1) - c.g2(r0, a1, r1, a0);
        } else {
          const int ra = ((i & 1) == 0 ? i : j) >> 1, rb = ((i & 1) == 1 ? i : j) >> 1;
          const int aa = ((a & 1) == 0 ? a : b) >> 1, ab_ = ((a & 1) == 1 ? a : b) >> 1;
          V = c.g2(ra, aa, rb, ab_);
        }
        hb->v[o + k] = V;
      }
    }
  }
  upload(c.d_hb_off, hb->off);
  upload(c.d_hb_ab, ab);
  upload(c.d_hb_v, hb->v);
  upload(c.d_hb_w, hb->w);
  c.hb = std::move(hb);
  c.hb_thr = thr;
  {
    std::lock_guard<std::mutex> lock(c.hb_mu);
    c.hb_len_cache.clear();
  }
}

}  // namespace dq

dq::DevView dq_ctx::view() const {
  dq::DevView v;
  v.n = n_orb;
  v.h1 = d_h1.p;
  v.h2 = d_h2.p;
  v.Js = d_Js.p;
  v.Ks = d_Ks.p;
  v.Jd = d_Jd.p;
  v.Kd = d_Kd.p;
  v.pm = d_pm.p;
  v.pm_cnt = d_pm_cnt.p;
  if (hb) {
    v.nso = hb->nso;
    v.hb_n = hb->off.empty() ? 0u : hb->off.back();
    v.hb_off = d_hb_off.p;
    v.hb_ab = d_hb_ab.p;
    v.hb_v = d_hb_v.p;
    v.hb_w = d_hb_w.p;
  }
  return v;
}

// Per-row number of entries kept by the cutoff (RowView::with_cutoff,
// hb_table.hpp:131-146: prefix with w >= eps; eps <= 0 keeps the whole row),
// further limited to w > 1e-15 (the element threshold; w equals |element|).
const dq::u32* dq_ctx::hb_lengths(double eps, hipStream_t s, bool host_wait) {
  DQ_REQUIRE(hb != nullptr, DQ_EINVAL, "Heat-Bath table not initialized");
  std::lock_guard<std::mutex> lock(hb_mu);
  for (auto& e : hb_len_cache)
    if (e->eps == eps) {
      DQ_HIP(host_wait ? hipEventSynchronize(e->ready) : hipStreamWaitEvent(s, e->ready, 0));
      return e->dev.p;
    }
  const size_t nrows = hb->off.size() - 1;
  auto e = std::make_unique<dq::HbLengths>();
  e->eps = eps;
  e->host.resize(nrows);
  for (size_t r = 0; r < nrows; ++r) {
    const size_t o = hb->off[r], L = hb->off[r + 1] - o;
    const double* w = hb->w.data() + o;
    auto prefix = [&](double t, bool strict) {
      size_t lo = 0, hi = L;
      while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (strict ? (w[mid] > t) : (w[mid] >= t))
          lo = mid + 1;
        else
          hi = mid;
      }
      return lo;
    };
    size_t keep = eps > 0.0 ? prefix(eps, false) : L;
    keep = std::min(keep, prefix(dq::kMatThr, true));
    e->host[r] = static_cast<dq::u32>(keep);
    // kept entries are a subset of the pair-mask set; equal sizes mean equal sets
    if (keep == pm_cnt[r]) e->host[r] |= dq::kFullRow;
  }
  e->dev.alloc(nrows);
  DQ_HIP(hipMemcpyAsync(e->dev.p, e->host.data(), nrows * sizeof(dq::u32), hipMemcpyHostToDevice, s));
  DQ_HIP(hipEventCreateWithFlags(&e->ready, hipEventDisableTiming));
  DQ_HIP(hipEventRecord(e->ready, s));
  if (host_wait) DQ_HIP(hipEventSynchronize(e->ready));
  hb_len_cache.push_back(std::move(e));
  return hb_len_cache.back()->dev.p;
}

void dq_ctx::check_rows(hipStream_t s) {
  int flag = 0;
  DQ_HIP(hipMemcpyAsync(&flag, d_err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  DQ_HIP(hipStreamSynchronize(s));
  if (flag) {
    DQ_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), s));
    DQ_HIP(hipStreamSynchronize(s));
    if (flag & 1) throw dq::Error(DQ_EINVAL, "determinant occupies orbitals at or above n_orb");
    if (flag & 4) throw dq::Error(DQ_ERUNTIME, "internal: a row's connection count disagrees with its walk");
    throw dq::Error(DQ_ERUNTIME, "internal: two kept kets of one row share an order position");
  }
}

// ----------------------------------------------------------------------------- C-ABI
extern "C" {

const char* dq_last_error(void) { return dq::g_last_error.c_str(); }

int dq_trim_memory(void) {
  return dq::guarded([&] { dq::cache_trim(); });
}
int dq_version(void) { return 1; }

int dq_ctx_create_fcidump(const char* path, int n_orb, dq_ctx** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(path && out, DQ_EINVAL, "null argument");
    auto c = std::make_unique<dq_ctx>();
    dq::alloc_packed(*c, n_orb);
    dq::parse_fcidump(*c, path);
    DQ_HIP(hipGetDevice(&c->device));
    dq::upload_tables(*c);
    *out = c.release();
  });
}

int dq_ctx_create(int n_orb, int n_elec, const double* h_h1e, const double* h_h2e, double e_nuc,
                  dq_ctx** out) {
  return dq::guarded([&] {
    DQ_REQUIRE(h_h1e && h_h2e && out, DQ_EINVAL, "null argument");
    auto c = std::make_unique<dq_ctx>();
    dq::alloc_packed(*c, n_orb);
    std::memcpy(c->h1e.data(), h_h1e, c->h1e.size() * sizeof(double));
    std::memcpy(c->h2e.data(), h_h2e, c->h2e.size() * sizeof(double));
    c->n_elec = n_elec;
    c->e_nuc = e_nuc;
    DQ_HIP(hipGetDevice(&c->device));
    dq::upload_tables(*c);
    *out = c.release();
  });
}

int dq_ctx_destroy(dq_ctx* ctx) {
  return dq::guarded([&] { delete ctx; });
}

int dq_ctx_info(const dq_ctx* ctx, int* n_orb, int* n_elec, double* e_nuc) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx, DQ_EINVAL, "null context");
    if (n_orb) *n_orb = ctx->n_orb;
    if (n_elec) *n_elec = ctx->n_elec;
    if (e_nuc) *e_nuc = ctx->e_nuc;
  });
}

int dq_hb_prepare(dq_ctx* ctx, double threshold) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx, DQ_EINVAL, "null context");
    dq::build_hb(*ctx, threshold);
  });
}

int dq_hb_clear(dq_ctx* ctx) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx, DQ_EINVAL, "null context");
    ctx->hb.reset();
    ctx->d_hb_off.release();
    ctx->d_hb_ab.release();
    ctx->d_hb_v.release();
    ctx->d_hb_w.release();
    std::lock_guard<std::mutex> lock(ctx->hb_mu);
    ctx->hb_len_cache.clear();
  });
}

int dq_hb_export(const dq_ctx* ctx, int64_t* n_entries, uint64_t* h_row_offsets, int32_t* h_a, int32_t* h_b,
                 double* h_w) {
  return dq::guarded([&] {
    DQ_REQUIRE(ctx && n_entries, DQ_EINVAL, "null argument");
    if (!ctx->hb) {
      *n_entries = 0;
      return;
    }
    const auto& hb = *ctx->hb;
    *n_entries = static_cast<int64_t>(hb.a.size());
    if (h_row_offsets)
      for (size_t r = 0; r < hb.off.size(); ++r) h_row_offsets[r] = hb.off[r];
    if (h_a) std::memcpy(h_a, hb.a.data(), hb.a.size() * sizeof(int32_t));
    if (h_b) std::memcpy(h_b, hb.b.data(), hb.b.size() * sizeof(int32_t));
    if (h_w) std::memcpy(h_w, hb.w.data(), hb.w.size() * sizeof(double));
  });
}

}  // extern "C"
