// ORACLE — test infrastructure only.
//
// A CPU restatement of the DetNQS local-energy path, used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
// The product (detnqs_amd/, libdetnqs_hip.so) never links or calls it.
//
// Every function cites the reference file:line whose semantics it restates.
// Where the reference leaves something implementation-defined the choice made
// here is stated.  Parity status of each function (what pins it) is listed in
// the header comment of oracle/oracle.py.
//
// Build: oracle/Makefile -> oracle/liboracle.so (g++ -O2 -fopenmp, no FMA).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

using u64 = std::uint64_t;
using u32 = std::uint32_t;

thread_local std::string g_err;

// ---------------------------------------------------------------- bits
// bit_utils.hpp:31-190
inline int popc(u64 x) { return __builtin_popcountll(x); }
inline int ctz64(u64 x) { return __builtin_ctzll(x); }
inline u64 low_mask(int n) {  // make_mask, bit_utils.hpp:115-119 (n>=64 -> all ones)
  if (n <= 0) return 0;
  if (n >= 64) return ~0ULL;
  return (1ULL << n) - 1;
}

struct Det {
  u64 a, b;
  bool operator==(const Det& o) const { return a == o.a && b == o.b; }
  bool operator<(const Det& o) const { return a != o.a ? a < o.a : b < o.b; }  // det.hpp:30-36
};
struct DetHash {
  size_t operator()(const Det& d) const {
    u64 x = d.a * 0x9E3779B97F4A7C15ULL ^ (d.b + 0x632BE59BD9B4E019ULL + (d.a << 6) + (d.a >> 2));
    x ^= x >> 31;
    return static_cast<size_t>(x * 0xBF58476D1CE4E5B9ULL);
  }
};

// ---------------------------------------------------------------- integrals
// integral_mo.hpp:30-49 packed keys; integral_mo.cpp:25-182 storage and FCIDUMP.
inline size_t k1(int p, int q) {
  if (p < q) std::swap(p, q);
  return static_cast<size_t>(p) * (p + 1) / 2 + q;
}
inline size_t k2(int p, int q, int r, int s) {
  size_t x = k1(p, q), y = k1(r, s);
  if (x < y) std::swap(x, y);
  return x * (x + 1) / 2 + y;
}

struct Ints {
  int n = 0;
  int nelec = 0;
  double enuc = 0.0;
  std::vector<double> h1, h2;
  explicit Ints(int n_orb) : n(n_orb) {
    if (n_orb <= 0 || n_orb > 64) throw std::invalid_argument("Number of orbitals must be in [1, 64]");
    size_t n1 = static_cast<size_t>(n) * (n + 1) / 2;
    h1.assign(n1, 0.0);
    h2.assign(n1 * (n1 + 1) / 2, 0.0);
  }
  double g1(int p, int q) const { return h1[k1(p, q)]; }
  double g2(int p, int q, int r, int s) const { return h2[k2(p, q, r, s)]; }

  // Spin-orbital view, integral_so.hpp:77-118: so = 2*mo + spin.
  double h1so(int mu, int nu) const {
    if ((mu & 1) != (nu & 1)) return 0.0;
    return g1(mu >> 1, nu >> 1);
  }
  // <mu kappa | nu lambda> = (mu nu | kappa lambda), zero on spin mismatch.
  double phys(int mu, int ka, int nu, int la) const {
    if ((mu & 1) != (nu & 1) || (ka & 1) != (la & 1)) return 0.0;
    return g2(mu >> 1, nu >> 1, ka >> 1, la >> 1);
  }
  double anti(int mu, int ka, int nu, int la) const {
    return phys(mu, ka, nu, la) - phys(mu, ka, la, nu);
  }
};

// FCIDUMP parser with the reference rules (integral_mo.cpp:38-182).
void load_fcidump(Ints& I, const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("Cannot open FCIDUMP: " + path);
  std::string line, nl;
  bool in = false;
  while (std::getline(f, line)) {
    if (!in && (line.find("&FCI") != std::string::npos || line.find("&fci") != std::string::npos)) in = true;
    if (in) {
      nl += line + " ";
      if (line.find("&END") != std::string::npos || line.find("&end") != std::string::npos) break;
    }
  }
  if (!in) throw std::runtime_error("No &FCI namelist found in FCIDUMP");
  std::string up = nl;
  for (auto& c : up) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  auto param = [&](const char* key, int& v) {
    std::string k = std::string(key) + "=";
    auto pos = up.find(k);
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
  if (param("NORB", norb) && norb != I.n) throw std::runtime_error("NORB mismatch");
  if (!param("NELEC", I.nelec) || I.nelec < 0) throw std::runtime_error("Invalid or missing NELEC in namelist");
  f.clear();
  f.seekg(0);
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    char c0 = line[0];
    if (c0 == '!' || c0 == '*' || c0 == 'C' || c0 == 'c' || c0 == '#') continue;
    if (line.find('&') != std::string::npos) continue;
    for (auto& c : line)
      if (c == 'D' || c == 'd') c = 'E';
    std::istringstream is(line);
    double v;
    int p, q, r, s;
    if (!(is >> v >> p >> q >> r >> s)) continue;
    if (p == 0 && q == 0 && r == 0 && s == 0)
      I.enuc = v;
    else if (r == 0 && s == 0)
      I.h1[k1(p - 1, q - 1)] = v;
    else
      I.h2[k2(p - 1, q - 1, r - 1, s - 1)] = v;
  }
}

// ---------------------------------------------------------------- excitations
// det_ops.cpp:45-196 (Scemama–Giner phase).
struct Exc {
  int degree = 0, na = 0, nb = 0;
  int ha[2]{}, pa[2]{}, hb[2]{}, pb[2]{};  // holes = ket \ bra, particles = bra \ ket
  double phase = 1.0;
};

inline int between(u64 occ, int x, int y) {  // det_ops.cpp:85-93
  if (x == y) return 0;
  int lo = std::min(x, y), hi = std::max(x, y);
  return popc(occ & low_mask(hi) & ~low_mask(lo + 1));
}

inline int spin_diff(u64 bra, u64 ket, int* h, int* p) {  // det_ops.cpp:45-77
  u64 hm = ket & ~bra, pm = bra & ~ket;
  int nh = popc(hm), np = popc(pm);
  if (nh != np) return 3;
  if (nh > 2) return nh;
  int i = 0;
  for (u64 m = hm; m; m &= m - 1) h[i++] = ctz64(m);
  i = 0;
  for (u64 m = pm; m; m &= m - 1) p[i++] = ctz64(m);
  return nh;
}

inline int perm_one_spin(u64 bra, int n, const int* h, const int* p) {  // det_ops.cpp:101-127
  int c = 0;
  for (int i = 0; i < n; ++i) c += between(bra, h[i], p[i]);
  if (n == 2) {
    int a = std::min(h[0], p[0]), b = std::max(h[0], p[0]);
    int cc = std::min(h[1], p[1]), d = std::max(h[1], p[1]);
    if ((cc > a && cc < b && d > b) || (a > cc && a < d && b > d)) ++c;
  }
  return c;
}

Exc analyze(const Det& bra, const Det& ket) {  // det_ops.cpp:159-196
  Exc e;
  e.na = spin_diff(bra.a, ket.a, e.ha, e.pa);
  e.nb = spin_diff(bra.b, ket.b, e.hb, e.pb);
  e.degree = e.na + e.nb;
  if (e.degree > 2) {
    e.phase = 0.0;
    return e;
  }
  int n = perm_one_spin(bra.a, e.na, e.ha, e.pa) + perm_one_spin(bra.b, e.nb, e.hb, e.pb);
  e.phase = (n & 1) ? -1.0 : 1.0;
  return e;
}

// ---------------------------------------------------------------- Slater–Condon
// ham_eval.cpp:33-147.
struct Ham {
  const Ints& I;
  explicit Ham(const Ints& i) : I(i) {}

  double diag(const Det& d) const {  // ham_eval.cpp:53-86 (occ list alpha then beta)
    int occ[128];
    int n = 0;
    for (u64 m = d.a; m; m &= m - 1) occ[n++] = 2 * ctz64(m);
    for (u64 m = d.b; m; m &= m - 1) occ[n++] = 2 * ctz64(m) + 1;
    double e1 = 0.0;
    for (int i = 0; i < n; ++i) e1 += I.h1so(occ[i], occ[i]);
    double e2 = 0.0;
    for (int i = 1; i < n; ++i)
      for (int j = 0; j < i; ++j) e2 += I.anti(occ[i], occ[j], occ[i], occ[j]);
    return e1 + e2;
  }

  double single(const Exc& x, const Det& ket) const {  // ham_eval.cpp:88-115
    bool alpha = x.na == 1;
    int i = alpha ? 2 * x.ha[0] : 2 * x.hb[0] + 1;
    int a = alpha ? 2 * x.pa[0] : 2 * x.pb[0] + 1;
    double v = I.h1so(a, i);
    for (u64 m = ket.a; m; m &= m - 1) {
      int k = 2 * ctz64(m);
      if (k != i) v += I.anti(a, k, i, k);
    }
    for (u64 m = ket.b; m; m &= m - 1) {
      int k = 2 * ctz64(m) + 1;
      if (k != i) v += I.anti(a, k, i, k);
    }
    return v;
  }

  double dbl(const Exc& x) const {  // ham_eval.cpp:117-147
    if (x.na == 2) return I.anti(2 * x.pa[0], 2 * x.pa[1], 2 * x.ha[0], 2 * x.ha[1]);
    if (x.nb == 2) return I.anti(2 * x.pb[0] + 1, 2 * x.pb[1] + 1, 2 * x.hb[0] + 1, 2 * x.hb[1] + 1);
    return I.phys(2 * x.pa[0], 2 * x.pb[0] + 1, 2 * x.ha[0], 2 * x.hb[0] + 1);
  }

  double elem(const Det& bra, const Det& ket) const {  // ham_eval.cpp:33-51
    Exc x = analyze(bra, ket);
    if (x.degree > 2) return 0.0;
    if (x.degree == 0) return diag(ket);
    if (x.degree == 1) return x.phase * single(x, ket);
    return x.phase * dbl(x);
  }
};

// ---------------------------------------------------------------- generators
// det_ops.hpp:123-268 (singles a, singles b; doubles aa, bb, ab).
template <class F>
void singles(const Det& k, int n, F&& f) {
  const u64 full = low_mask(n);
  for (int s = 0; s < 2; ++s) {
    u64 occ = s ? k.b : k.a, vir = ~occ & full;
    if (!occ || !vir) continue;
    for (u64 h = occ; h; h &= h - 1) {
      u64 hb = h & (~h + 1);
      for (u64 p = vir; p; p &= p - 1) {
        u64 pb = p & (~p + 1);
        if (s == 0)
          f(Det{k.a ^ hb ^ pb, k.b});
        else
          f(Det{k.a, k.b ^ hb ^ pb});
      }
    }
  }
}

template <class F>
void doubles(const Det& k, int n, F&& f) {
  const u64 full = low_mask(n);
  for (int s = 0; s < 2; ++s) {
    u64 occ = s ? k.b : k.a, vir = ~occ & full;
    if (popc(occ) < 2 || popc(vir) < 2) continue;
    for (u64 h1 = occ; h1; h1 &= h1 - 1) {
      u64 b1 = h1 & (~h1 + 1);
      for (u64 h2 = h1 & (h1 - 1); h2; h2 &= h2 - 1) {
        u64 holes = b1 | (h2 & (~h2 + 1));
        for (u64 p1 = vir; p1; p1 &= p1 - 1) {
          u64 c1 = p1 & (~p1 + 1);
          for (u64 p2 = p1 & (p1 - 1); p2; p2 &= p2 - 1) {
            u64 parts = c1 | (p2 & (~p2 + 1));
            if (s == 0)
              f(Det{k.a ^ holes ^ parts, k.b});
            else
              f(Det{k.a, k.b ^ holes ^ parts});
          }
        }
      }
    }
  }
  u64 va = ~k.a & full, vb = ~k.b & full;
  if (!k.a || !va || !k.b || !vb) return;
  for (u64 ha = k.a; ha; ha &= ha - 1)
    for (u64 pa = va; pa; pa &= pa - 1) {
      u64 na = k.a ^ (ha & (~ha + 1)) ^ (pa & (~pa + 1));
      for (u64 hb = k.b; hb; hb &= hb - 1)
        for (u64 pb = vb; pb; pb &= pb - 1) f(Det{na, k.b ^ (hb & (~hb + 1)) ^ (pb & (~pb + 1))});
    }
}

// ---------------------------------------------------------------- heat-bath table
// hb_table.cpp:39-168, hb_table.hpp:131-175.  Row {i>j} holds (a>b) disjoint
// from {i,j}, spin-compatible, w = |<ij||ab>| >= thr, sorted descending by w
// with std::sort (the reference's std::ranges::sort is the same libstdc++
// introsort, so equal weights come out in the same order).
struct HB {
  int nso = 0;
  std::vector<size_t> off;
  std::vector<int> a, b;
  std::vector<double> w;
  static size_t rid(int i, int j) {
    if (i < j) std::swap(i, j);
    return static_cast<size_t>(i) * (i - 1) / 2 + j;
  }
  void build(const Ints& I, double thr) {
    nso = 2 * I.n;
    if (nso < 2) throw std::invalid_argument("HeatBathTable requires n_so >= 2");
    size_t nrows = static_cast<size_t>(nso) * (nso - 1) / 2;
    struct E {
      int a, b;
      double w;
    };
    std::vector<std::vector<E>> rows(nrows);
#pragma omp parallel for schedule(dynamic)
    for (int i = 1; i < nso; ++i) {
      for (int j = 0; j < i; ++j) {
        std::vector<E> buf;
        int si = i & 1, sj = j & 1;
        for (int x = 1; x < nso; ++x)
          for (int y = 0; y < x; ++y) {
            if (x == i || x == j || y == i || y == j) continue;
            int sx = x & 1, sy = y & 1;
            if (!((sx == si && sy == sj) || (sx == sj && sy == si))) continue;
            double wt = std::abs(I.anti(i, j, x, y));
            if (wt >= thr) buf.push_back({x, y, wt});
          }
        std::sort(buf.begin(), buf.end(), [](const E& p, const E& q) { return p.w > q.w; });
        rows[rid(i, j)] = std::move(buf);
      }
    }
    off.assign(nrows + 1, 0);
    for (size_t r = 0; r < nrows; ++r) off[r + 1] = off[r] + rows[r].size();
    a.resize(off[nrows]);
    b.resize(off[nrows]);
    w.resize(off[nrows]);
    for (size_t r = 0; r < nrows; ++r)
      for (size_t k = 0; k < rows[r].size(); ++k) {
        a[off[r] + k] = rows[r][k].a;
        b[off[r] + k] = rows[r][k].b;
        w[off[r] + k] = rows[r][k].w;
      }
  }
  // Prefix length with w >= eps (hb_table.hpp:131-146; eps <= 0 keeps all).
  size_t cut(size_t r, double eps) const {
    size_t len = off[r + 1] - off[r];
    if (eps <= 0.0 || len == 0) return len;
    size_t lo = 0, hi = len;
    const double* ww = w.data() + off[r];
    while (lo < hi) {
      size_t mid = lo + (hi - lo) / 2;
      if (ww[mid] >= eps)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  }
};

// ham_utils.hpp:280-313 with ham_utils.cpp:453-506: occupied SOs alpha then beta,
// out-of-range orbitals count as occupied.
template <class F>
void doubles_hb(const Det& bra, int n_orb, const HB& hb, double cutoff, F&& f) {
  int occ[128];
  int n = 0;
  for (u64 m = bra.a; m; m &= m - 1) occ[n++] = 2 * ctz64(m);
  for (u64 m = bra.b; m; m &= m - 1) occ[n++] = 2 * ctz64(m) + 1;
  auto is_occ = [&](int so) {
    int mo = so >> 1;
    if (mo < 0 || mo >= n_orb) return true;
    u64 bit = 1ULL << mo;
    return ((so & 1) ? (bra.b & bit) : (bra.a & bit)) != 0;
  };
  for (int p = 1; p < n; ++p)
    for (int q = 0; q < p; ++q) {
      int i = occ[p], j = occ[q];
      size_t r = HB::rid(i, j);
      size_t len = hb.cut(r, cutoff);
      for (size_t k = 0; k < len; ++k) {
        int x = hb.a[hb.off[r] + k], y = hb.b[hb.off[r] + k];
        if (is_occ(x) || is_occ(y)) continue;
        Det out = bra;
        auto flip = [&](int so, bool set) {
          u64 bit = 1ULL << (so >> 1);
          u64& w = (so & 1) ? out.b : out.a;
          w = set ? (w | bit) : (w & ~bit);
        };
        flip(i, false);
        flip(j, false);
        flip(x, true);
        flip(y, true);
        f(out);
      }
    }
}

constexpr double kMatThr = 1e-15;  // ham_utils.hpp:39

// ---------------------------------------------------------------- context
struct Ctx {
  Ints I;
  Ham H;
  std::unique_ptr<HB> hb;
  explicit Ctx(int n) : I(n), H(I) {}
};

// local_conn.cpp:36-94.
void local_row(const Ctx& c, const Det& bra, int n_orb, bool use_hb, double eps,
               std::vector<Det>& dets, std::vector<double>& vals) {
  const double sthr = use_hb ? std::max(eps, kMatThr) : kMatThr;
  double hd = c.H.diag(bra);
  if (std::abs(hd) > kMatThr) {
    dets.push_back(bra);
    vals.push_back(hd);
  }
  singles(bra, n_orb, [&](const Det& k) {
    double v = c.H.elem(bra, k);
    if (std::abs(v) > sthr) {
      dets.push_back(k);
      vals.push_back(v);
    }
  });
  auto vd = [&](const Det& k) {
    double v = c.H.elem(bra, k);
    if (std::abs(v) > kMatThr) {
      dets.push_back(k);
      vals.push_back(v);
    }
  };
  if (use_hb)
    doubles_hb(bra, n_orb, *c.hb, eps, vd);
  else
    doubles(bra, n_orb, vd);
}

// DetMap (det_space.cpp:29-65): from_ordered keeps input order.  In the
// tolerant mode later duplicates are dropped and the rest numbered
// consecutively (det_space.cpp:40-54), so a determinant's index is its rank
// among first occurrences, not its row; the energy and PT2 read coeffs[index]
// (local_conn.cpp:192-194, 322-333), i.e. another row's coefficient once a
// duplicate precedes it.
struct DetMap {
  std::vector<Det> dets;
  std::unordered_map<Det, u32, DetHash> idx;
  static DetMap ordered(const std::vector<Det>& v, bool strict) {
    DetMap m;
    m.idx.reserve(v.size() * 2);
    for (const auto& d : v) {
      auto it = m.idx.find(d);
      if (it != m.idx.end()) {
        if (strict) throw std::invalid_argument("DetMap::from_ordered: duplicate determinant found");
        continue;
      }
      m.idx.emplace(d, static_cast<u32>(m.dets.size()));
      m.dets.push_back(d);
    }
    return m;
  }
  static DetMap sorted_unique(std::vector<Det> v) {  // det_space.cpp:23-27
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    DetMap m;
    m.dets = std::move(v);
    for (size_t i = 0; i < m.dets.size(); ++i) m.idx.emplace(m.dets[i], static_cast<u32>(i));
    return m;
  }
  long find(const Det& d) const {
    auto it = idx.find(d);
    return it == idx.end() ? -1 : static_cast<long>(it->second);
  }
};

struct COO {
  std::vector<u32> r, c;
  std::vector<double> v;
  u32 nr = 0, nc = 0;
};

// sort_and_merge_coo, ham_utils.cpp:41-83 (sort by (row,col), sum dups, drop 0).
void sort_merge(COO& m) {
  size_t n = m.v.size();
  std::vector<size_t> id(n);
  for (size_t i = 0; i < n; ++i) id[i] = i;
  std::sort(id.begin(), id.end(), [&](size_t x, size_t y) {
    return m.r[x] != m.r[y] ? m.r[x] < m.r[y] : m.c[x] < m.c[y];
  });
  COO o;
  o.nr = m.nr;
  o.nc = m.nc;
  for (size_t i = 0; i < n;) {
    size_t k = id[i];
    u32 r = m.r[k], c = m.c[k];
    double v = m.v[k];
    ++i;
    while (i < n && m.r[id[i]] == r && m.c[id[i]] == c) v += m.v[id[i++]];
    if (std::abs(v) > 0.0) {
      o.r.push_back(r);
      o.c.push_back(c);
      o.v.push_back(v);
    }
  }
  m = std::move(o);
}

// stream_build (build_ham.cpp:213-314) for the three policies (build_ham.cpp:78-196).
enum Policy { KNOWN, STATIC_HB, DYN_AMP };
struct Blocks {
  COO vv, vp;
  std::vector<Det> P;
};

Blocks assemble(const Ctx& c, const std::vector<Det>& V, const std::vector<Det>* ext, int n_orb,
                Policy pol, bool use_hb, double eps, const double* psi) {
  Blocks out;
  DetMap mv = DetMap::ordered(V, true);
  DetMap me;
  bool has_ext = ext && !ext->empty();
  if (has_ext) me = DetMap::ordered(*ext, true);
  const size_t n = V.size();
  std::vector<COO> tvv(n), tvi(n);
  struct Keyed {
    u32 row;
    Det k;
    double v;
  };
  std::vector<std::vector<Keyed>> tvk(n);
#pragma omp parallel for schedule(dynamic, 16)
  for (long ii = 0; ii < static_cast<long>(n); ++ii) {
    const u32 row = static_cast<u32>(ii);
    const Det& bra = V[ii];
    COO& vv = tvv[ii];
    COO& vi = tvi[ii];
    auto& vk = tvk[ii];
    double h = c.H.diag(bra);
    if (std::abs(h) > kMatThr) {
      vv.r.push_back(row);
      vv.c.push_back(row);
      vv.v.push_back(h);
    }
    auto route = [&](const Det& k, double val) {
      long j = mv.find(k);
      if (j >= 0) {
        vv.r.push_back(row);
        vv.c.push_back(static_cast<u32>(j));
        vv.v.push_back(val);
        return;
      }
      if (pol == KNOWN) {
        long e = has_ext ? me.find(k) : -1;
        if (e >= 0) {
          vi.r.push_back(row);
          vi.c.push_back(static_cast<u32>(e));
          vi.v.push_back(val);
        }
        return;
      }
      vk.push_back({row, k, val});
    };
    auto classify_drop = [&](const Det& k) {
      if (pol != KNOWN) return false;
      return mv.find(k) < 0 && !(has_ext && me.find(k) >= 0);
    };
    singles(bra, n_orb, [&](const Det& k) {
      if (classify_drop(k)) return;
      double v = c.H.elem(bra, k);
      if (std::abs(v) <= kMatThr) return;
      if (pol == DYN_AMP && !(std::abs(v * psi[ii]) >= eps)) return;  // build_ham.cpp:186-188
      route(k, v);
    });
    auto vd = [&](const Det& k) {
      if (classify_drop(k)) return;
      double v = c.H.elem(bra, k);
      if (std::abs(v) <= kMatThr) return;
      route(k, v);
    };
    if (pol == KNOWN) {
      doubles(bra, n_orb, vd);
    } else if (pol == STATIC_HB) {
      if (use_hb && c.hb)
        doubles_hb(bra, n_orb, *c.hb, eps, vd);
      else
        doubles(bra, n_orb, vd);
    } else {
      if (!c.hb) {
        doubles(bra, n_orb, vd);
      } else {
        double amp = std::max(std::abs(psi[ii]), 1e-12);  // build_ham.cpp:181-183
        doubles_hb(bra, n_orb, *c.hb, eps / amp, vd);
      }
    }
  }
  for (size_t i = 0; i < n; ++i) {
    out.vv.r.insert(out.vv.r.end(), tvv[i].r.begin(), tvv[i].r.end());
    out.vv.c.insert(out.vv.c.end(), tvv[i].c.begin(), tvv[i].c.end());
    out.vv.v.insert(out.vv.v.end(), tvv[i].v.begin(), tvv[i].v.end());
  }
  out.vv.nr = out.vv.nc = static_cast<u32>(n);
  sort_merge(out.vv);
  if (pol == KNOWN) {
    if (has_ext) {
      for (size_t i = 0; i < n; ++i) {
        out.vp.r.insert(out.vp.r.end(), tvi[i].r.begin(), tvi[i].r.end());
        out.vp.c.insert(out.vp.c.end(), tvi[i].c.begin(), tvi[i].c.end());
        out.vp.v.insert(out.vp.v.end(), tvi[i].v.begin(), tvi[i].v.end());
      }
      out.vp.nr = static_cast<u32>(n);
      out.vp.nc = static_cast<u32>(ext->size());
      sort_merge(out.vp);
      out.P = me.dets;
    }
    return out;
  }
  // finalize_vp_deferred, build_ham.cpp:346-380: P = sorted unique keys.
  std::vector<Det> keys;
  for (auto& v : tvk)
    for (auto& x : v) keys.push_back(x.k);
  DetMap mp = DetMap::sorted_unique(std::move(keys));
  for (auto& v : tvk)
    for (auto& x : v) {
      out.vp.r.push_back(x.row);
      out.vp.c.push_back(static_cast<u32>(mp.find(x.k)));
      out.vp.v.push_back(x.v);
    }
  out.vp.nr = static_cast<u32>(n);
  out.vp.nc = static_cast<u32>(mp.dets.size());
  sort_merge(out.vp);
  out.P = std::move(mp.dets);
  return out;
}

// ---------------------------------------------------------------- SpGEMM
// ham_eff.cpp:35-233.
struct CSR {
  std::vector<size_t> ptr;
  std::vector<u32> col;
  std::vector<double> val;
};

// coo_to_csr / coo_to_csc by buckets (ham_utils.cpp:183-302): per bucket a
// stable order by index, duplicates summed in bucket order, exact zeros dropped.
CSR buckets(const COO& m, u32 nmaj, bool by_row) {
  std::vector<std::vector<std::pair<u32, double>>> bk(nmaj);
  for (size_t k = 0; k < m.v.size(); ++k) {
    u32 maj = by_row ? m.r[k] : m.c[k], mi = by_row ? m.c[k] : m.r[k];
    if (maj >= nmaj) throw std::out_of_range("coo_to_csr: index out of bounds");
    bk[maj].emplace_back(mi, m.v[k]);
  }
  CSR o;
  o.ptr.assign(nmaj + 1, 0);
  for (u32 i = 0; i < nmaj; ++i) {
    auto& b = bk[i];
    std::stable_sort(b.begin(), b.end(), [](auto& x, auto& y) { return x.first < y.first; });
    for (size_t t = 0; t < b.size();) {
      u32 id = b[t].first;
      double acc = b[t].second;
      ++t;
      while (t < b.size() && b[t].first == id) acc += b[t++].second;
      if (acc != 0.0) {
        o.col.push_back(id);
        o.val.push_back(acc);
      }
    }
    o.ptr[i + 1] = o.col.size();
  }
  return o;
}

COO ham_eff(const COO& VV, const COO& VP, const std::vector<double>& hpp, double eref, bool sigma,
            double eps) {
  const size_t np = hpp.size();
  if (np == 0) {
    COO h = VV;
    sort_merge(h);
    return h;
  }
  if (VP.nc != static_cast<u32>(np)) throw std::invalid_argument("H_VP columns != h_pp_diag size");
  const u32 nv = VV.nr ? VV.nr : VP.nr;
  if (VP.nr != nv) throw std::invalid_argument("H_VP rows != H_VV rows");
  CSR R = buckets(VP, nv, true);
  CSR C = buckets(VP, static_cast<u32>(np), false);
  std::vector<double> dinv(np);
  for (size_t j = 0; j < np; ++j) {
    double d = eref - hpp[j];
    dinv[j] = sigma ? d / (d * d + eps * eps) : 1.0 / (d + std::copysign(eps, d));
  }
  std::vector<char> skip(np, 0);
  for (size_t j = 0; j < np; ++j) {
    double s = 0.0;
    for (size_t q = C.ptr[j]; q < C.ptr[j + 1]; ++q) s += C.val[q] * C.val[q];
    if (std::abs(dinv[j]) * s < 1e-12) skip[j] = 1;  // MICRO_CONTRIB_THRESH
  }
  // ADD[i,j] = sum_k (H_VP[i,k]*dinv[k]) * H_VP[j,k], accumulated in ascending k.
  std::vector<std::map<u32, double>> add(nv);
#pragma omp parallel for schedule(dynamic, 16)
  for (long ii = 0; ii < static_cast<long>(nv); ++ii) {
    auto& row = add[ii];
    for (size_t p = R.ptr[ii]; p < R.ptr[ii + 1]; ++p) {
      u32 k = R.col[p];
      if (skip[k]) continue;
      for (size_t q = C.ptr[k]; q < C.ptr[k + 1]; ++q) row.emplace(C.col[q], 0.0);
    }
    for (size_t p = R.ptr[ii]; p < R.ptr[ii + 1]; ++p) {
      u32 k = R.col[p];
      if (skip[k]) continue;
      double a = R.val[p];
      if (a == 0.0) continue;
      double s = dinv[k];
      if (s == 0.0) continue;
      double as = a * s;
      for (size_t q = C.ptr[k]; q < C.ptr[k + 1]; ++q) row[C.col[q]] += as * C.val[q];
    }
  }
  // csr_add(VV, ADD, 1e-15) then csr_to_coo (ham_utils.cpp:330-442).
  CSR V = buckets(VV, nv, true);
  COO o;
  o.nr = o.nc = nv;
  for (u32 i = 0; i < nv; ++i) {
    size_t pa = V.ptr[i], qa = V.ptr[i + 1];
    auto it = add[i].begin(), ie = add[i].end();
    auto emit = [&](u32 c, double v) {
      if (std::abs(v) > kMatThr) {
        o.r.push_back(i);
        o.c.push_back(c);
        o.v.push_back(v);
      }
    };
    while (pa < qa && it != ie) {
      if (V.col[pa] < it->first) {
        emit(V.col[pa], V.val[pa]);
        ++pa;
      } else if (it->first < V.col[pa]) {
        emit(it->first, it->second);
        ++it;
      } else {
        emit(V.col[pa], V.val[pa] + it->second);
        ++pa;
        ++it;
      }
    }
    for (; pa < qa; ++pa) emit(V.col[pa], V.val[pa]);
    for (; it != ie; ++it) emit(it->first, it->second);
  }
  return o;
}

// ---------------------------------------------------------------- slogdet
// Semantics of the reference FFI (logdet_cuda.cu:187-291, 293-386): sign and
// log|det| from an LU with partial pivoting; a zero pivot or a non-finite
// log|u_kk| gives (0, -inf); backward grad = cot_logabs * A^{-T}, zero when
// singular.  The factorisation restates LAPACK dgetf2 (Golub & Van Loan
// Alg. 3.4.1): pivot = first row of maximal |a| in the current order,
// multipliers by the reciprocal of the pivot when |pivot| >= DBL_MIN.
struct LU {
  int n;
  std::vector<double> a;  // row-major, in place
  std::vector<int> perm;  // perm[k] = original row now at position k
  bool singular = false;
  int swaps = 0;
};

LU lu_factor(const double* A, int n) {
  LU f;
  f.n = n;
  f.a.assign(A, A + static_cast<size_t>(n) * n);
  f.perm.resize(n);
  for (int i = 0; i < n; ++i) f.perm[i] = i;
  double* a = f.a.data();
  for (int k = 0; k < n; ++k) {
    int p = k;
    double m = std::abs(a[k * n + k]);
    for (int r = k + 1; r < n; ++r)
      if (std::abs(a[r * n + k]) > m) {
        m = std::abs(a[r * n + k]);
        p = r;
      }
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      std::swap(f.perm[k], f.perm[p]);
      ++f.swaps;
    }
    double piv = a[k * n + k];
    if (piv == 0.0) {
      f.singular = true;
      continue;
    }
    const bool recip = std::abs(piv) >= std::numeric_limits<double>::min();
    const double rp = 1.0 / piv;
    for (int r = k + 1; r < n; ++r) {
      double l = recip ? a[r * n + k] * rp : a[r * n + k] / piv;
      a[r * n + k] = l;
      for (int j = k + 1; j < n; ++j) a[r * n + j] = a[r * n + j] - l * a[k * n + j];
    }
  }
  return f;
}

void slogdet_one(const double* A, int n, double* sign, double* logabs) {
  LU f = lu_factor(A, n);
  if (f.singular) {
    *sign = 0.0;
    *logabs = -INFINITY;
    return;
  }
  double s = (f.swaps & 1) ? -1.0 : 1.0, la = 0.0;
  for (int k = 0; k < n; ++k) {
    double d = f.a[k * n + k], ad = std::abs(d);
    if (!(ad > 0.0 && std::isfinite(ad))) {
      *sign = 0.0;
      *logabs = -INFINITY;
      return;
    }
    la += std::log(ad);
    if (d < 0.0) s = -s;
  }
  *sign = s;
  *logabs = la;
}

// X = A^{-T} from the LU (dgetrs 'T'): A^T X = I with P A = L U.
void inv_t_one(const double* A, int n, double cot, double* G) {
  LU f = lu_factor(A, n);
  if (f.singular) {
    std::fill(G, G + static_cast<size_t>(n) * n, 0.0);
    return;
  }
  const double* a = f.a.data();
  std::vector<double> z(n);
  for (int col = 0; col < n; ++col) {
    // solve U^T y = e_col
    for (int i = 0; i < n; ++i) {
      double s = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= a[k * n + i] * z[k];
      z[i] = s / a[i * n + i];
    }
    // solve L^T w = y (unit diagonal)
    for (int i = n - 1; i >= 0; --i) {
      double s = z[i];
      for (int k = i + 1; k < n; ++k) s -= a[k * n + i] * z[k];
      z[i] = s;
    }
    // x = P^T w : x[perm[k]] = w[k]; column col of X
    for (int k = 0; k < n; ++k) G[static_cast<size_t>(f.perm[k]) * n + col] = cot * z[k];
  }
}

// ---------------------------------------------------------------- bundle
struct Bundle {
  std::map<std::string, std::vector<unsigned char>> arr;
  template <class T>
  void put(const std::string& k, const std::vector<T>& v) {
    auto& d = arr[k];
    d.resize(v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(d.data(), v.data(), d.size());
  }
  void dets(const std::string& k, const std::vector<Det>& v) {
    std::vector<u64> f(2 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      f[2 * i] = v[i].a;
      f[2 * i + 1] = v[i].b;
    }
    put(k, f);
  }
  void coo(const std::string& p, const COO& c) {
    put(p + ".row", c.r);
    put(p + ".col", c.c);
    put(p + ".val", c.v);
    put(p + ".shape", std::vector<u32>{c.nr, c.nc});
  }
};

std::vector<Det> to_dets(const u64* d, size_t n) {
  std::vector<Det> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = Det{d[2 * i], d[2 * i + 1]};
  return v;
}

template <class F>
void* guard(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_ctx_from_fcidump(const char* path, int n_orb) {
  return guard([&]() -> void* {
    auto* c = new Ctx(n_orb);
    try {
      load_fcidump(c->I, path);
    } catch (...) {
      delete c;
      throw;
    }
    return c;
  });
}

void* orc_ctx_from_arrays(int n_orb, const double* h1, const double* h2, double enuc, int nelec) {
  return guard([&]() -> void* {
    auto* c = new Ctx(n_orb);
    std::memcpy(c->I.h1.data(), h1, c->I.h1.size() * sizeof(double));
    std::memcpy(c->I.h2.data(), h2, c->I.h2.size() * sizeof(double));
    c->I.enuc = enuc;
    c->I.nelec = nelec;
    return c;
  });
}

void orc_ctx_destroy(void* c) { delete static_cast<Ctx*>(c); }
double orc_ctx_e_nuc(void* c) { return static_cast<Ctx*>(c)->I.enuc; }
int orc_ctx_nelec(void* c) { return static_cast<Ctx*>(c)->I.nelec; }
double orc_get_h1e(void* c, int p, int q) { return static_cast<Ctx*>(c)->I.g1(p, q); }
double orc_get_h2e(void* c, int p, int q, int r, int s) { return static_cast<Ctx*>(c)->I.g2(p, q, r, s); }

int orc_hb_prepare(void* c, double thr) {
  return guard([&]() -> void* {
           auto* x = static_cast<Ctx*>(c);
           x->hb = std::make_unique<HB>();
           x->hb->build(x->I, thr);
           return x;
         })
             ? 0
             : -1;
}

void* orc_hb_export(void* c) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (!x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto* B = new Bundle;
    std::vector<u64> off(x->hb->off.begin(), x->hb->off.end());
    B->put("row_offsets", off);
    std::vector<int32_t> a(x->hb->a.begin(), x->hb->a.end()), b(x->hb->b.begin(), x->hb->b.end());
    B->put("a", a);
    B->put("b", b);
    B->put("w", x->hb->w);
    return B;
  });
}

int orc_analyze(const u64* bra, const u64* ket, int* out3, double* phase) {
  Exc e = analyze(Det{bra[0], bra[1]}, Det{ket[0], ket[1]});
  out3[0] = e.degree;
  out3[1] = e.na;
  out3[2] = e.nb;
  *phase = e.phase;
  return 0;
}

int orc_compute_elem(void* c, const u64* bra, const u64* ket, double* out) {
  *out = static_cast<Ctx*>(c)->H.elem(Det{bra[0], bra[1]}, Det{ket[0], ket[1]});
  return 0;
}

void* orc_ham_diag(void* c, const u64* dets, size_t n) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    std::vector<double> d(n);
#pragma omp parallel for if (n > 256)
    for (long i = 0; i < static_cast<long>(n); ++i) d[i] = x->H.diag(Det{dets[2 * i], dets[2 * i + 1]});
    auto* B = new Bundle;
    B->put("diag", d);
    return B;
  });
}

// local_conn.cpp:96-142 (64-bit offsets here; the reference uses int32).
void* orc_local_connections(void* c, const u64* dets, size_t n, int n_orb, int use_hb, double eps) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    std::vector<std::vector<Det>> rd(n);
    std::vector<std::vector<double>> rv(n);
#pragma omp parallel for schedule(dynamic, 32)
    for (long i = 0; i < static_cast<long>(n); ++i)
      local_row(*x, Det{dets[2 * i], dets[2 * i + 1]}, n_orb, use_hb != 0, eps, rd[i], rv[i]);
    std::vector<int64_t> off(n + 1, 0);
    for (size_t i = 0; i < n; ++i) off[i + 1] = off[i] + static_cast<int64_t>(rd[i].size());
    std::vector<Det> all;
    std::vector<double> vals;
    all.reserve(off[n]);
    vals.reserve(off[n]);
    for (size_t i = 0; i < n; ++i) {
      all.insert(all.end(), rd[i].begin(), rd[i].end());
      vals.insert(vals.end(), rv[i].begin(), rv[i].end());
    }
    auto* B = new Bundle;
    B->put("offsets", off);
    B->dets("dets", all);
    B->put("values", vals);
    return B;
  });
}

// Row counts and per-row value sums (fixed order), for size-independent checks.
int orc_local_conn_stats(void* c, const u64* dets, size_t n, int n_orb, int use_hb, double eps,
                         int64_t* counts, double* vsum) {
  auto* x = static_cast<Ctx*>(c);
  if (use_hb && !x->hb) {
    g_err = "Heat-Bath table not initialized";
    return -1;
  }
#pragma omp parallel for schedule(dynamic, 8)
  for (long i = 0; i < static_cast<long>(n); ++i) {
    std::vector<Det> d;
    std::vector<double> v;
    local_row(*x, Det{dets[2 * i], dets[2 * i + 1]}, n_orb, use_hb != 0, eps, d, v);
    counts[i] = static_cast<int64_t>(v.size());
    double s = 0.0;
    for (double t : v) s += t;
    vsum[i] = s;
  }
  return 0;
}

void* orc_ham_block(void* c, const u64* dv, size_t nv, const u64* dp, size_t np, int n_orb) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    auto V = to_dets(dv, nv);
    auto P = to_dets(dp, np);
    Blocks b = assemble(*x, V, dp ? &P : nullptr, n_orb, KNOWN, false, 0.0, nullptr);
    auto* B = new Bundle;
    B->coo("H_VV", b.vv);
    B->coo("H_VP", b.vp);
    B->dets("det_P", b.P);
    return B;
  });
}

void* orc_ham_conn(void* c, const u64* dets, size_t n, int n_orb, int use_hb, double eps) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto V = to_dets(dets, n);
    Blocks b = assemble(*x, V, nullptr, n_orb, STATIC_HB, use_hb != 0, eps, nullptr);
    auto* B = new Bundle;
    B->coo("H_VV", b.vv);
    B->coo("H_VP", b.vp);
    B->dets("det_P", b.P);
    return B;
  });
}

void* orc_ham_conn_amp(void* c, const u64* dets, const double* psi, size_t n, int n_orb, double eps) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (!x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto V = to_dets(dets, n);
    Blocks b = assemble(*x, V, nullptr, n_orb, DYN_AMP, true, eps, psi);
    auto* B = new Bundle;
    B->coo("H_VV", b.vv);
    B->coo("H_VP", b.vp);
    B->dets("det_P", b.P);
    return B;
  });
}

void* orc_ham_eff(const u32* vr, const u32* vc, const double* vv, size_t vnnz, u32 vnr, u32 vnc,
                  const u32* pr, const u32* pc, const double* pv, size_t pnnz, u32 pnr, u32 pnc,
                  const double* hpp, size_t np, double eref, int sigma, double eps) {
  return guard([&]() -> void* {
    COO VV, VP;
    VV.r.assign(vr, vr + vnnz);
    VV.c.assign(vc, vc + vnnz);
    VV.v.assign(vv, vv + vnnz);
    VV.nr = vnr;
    VV.nc = vnc;
    VP.r.assign(pr, pr + pnnz);
    VP.c.assign(pc, pc + pnnz);
    VP.v.assign(pv, pv + pnnz);
    VP.nr = pnr;
    VP.nc = pnc;
    std::vector<double> d(hpp, hpp + np);
    COO H = ham_eff(VV, VP, d, eref, sigma != 0, eps);
    auto* B = new Bundle;
    B->coo("H", H);
    return B;
  });
}

// local_conn.cpp:148-229, with rows reduced in index order (the reference's
// OpenMP reduction order is unspecified).  sigma_out (optional) gets the per-row
// sums sigma_i = sum_{j in V} H_ij c_j (0 for skipped rows).
// sigma_i of compute_variational_energy (local_conn.cpp:176-223) for row i of V.
static double sigma_row(const Ctx& x, const std::vector<Det>& V, const DetMap& m, const double* coeffs, long i,
                        int n_orb, bool use_hb, double eps, double sthr) {
  const Det& bra = V[i];
  const double ci = coeffs[i];
  if (ci * ci < 1e-12) return 0.0;
  double s = 0.0;
  auto acc = [&](const Det& k, double h) {
    long j = m.find(k);
    if (j >= 0) s += h * coeffs[j];
  };
  double hd = x.H.diag(bra);
  if (std::abs(hd) > kMatThr) s += hd * ci;
  singles(bra, n_orb, [&](const Det& k) {
    double v = x.H.elem(bra, k);
    if (std::abs(v) > sthr) acc(k, v);
  });
  auto vd = [&](const Det& k) {
    double v = x.H.elem(bra, k);
    if (std::abs(v) > kMatThr) acc(k, v);
  };
  if (use_hb)
    doubles_hb(bra, n_orb, *x.hb, eps, vd);
  else
    doubles(bra, n_orb, vd);
  return s;
}

int orc_variational_energy(void* c, const u64* dets, const double* coeffs, size_t n, int n_orb,
                           int use_hb, double eps, double* out, double* sigma_out) {
  auto* x = static_cast<Ctx*>(c);
  void* ok = guard([&]() -> void* {
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath mode requires non-null hb_table");
    auto V = to_dets(dets, n);
    DetMap m = DetMap::ordered(V, false);
    std::vector<double> sig(n, 0.0);
    const double sthr = use_hb ? std::max(eps, kMatThr) : kMatThr;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < static_cast<long>(n); ++i)
      sig[i] = sigma_row(*x, V, m, coeffs, i, n_orb, use_hb != 0, eps, sthr);
    double e = 0.0;
    for (size_t i = 0; i < n; ++i) e += coeffs[i] * sig[i];
    *out = e;
    if (sigma_out) std::memcpy(sigma_out, sig.data(), n * sizeof(double));
    return x;
  });
  return ok ? 0 : -1;
}

// sigma of the listed rows only (same arithmetic as orc_variational_energy), for
// checks on spaces whose full energy is too slow on the host.
int orc_sigma_rows(void* c, const u64* dets, const double* coeffs, size_t n, int n_orb, int use_hb, double eps,
                   const int64_t* rows, size_t m_rows, double* sigma_out) {
  auto* x = static_cast<Ctx*>(c);
  void* ok = guard([&]() -> void* {
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath mode requires non-null hb_table");
    auto V = to_dets(dets, n);
    DetMap m = DetMap::ordered(V, false);
    const double sthr = use_hb ? std::max(eps, kMatThr) : kMatThr;
    for (size_t k = 0; k < m_rows; ++k)
      if (rows[k] < 0 || static_cast<size_t>(rows[k]) >= n) throw std::out_of_range("row index out of range");
#pragma omp parallel for schedule(dynamic, 1)
    for (long k = 0; k < static_cast<long>(m_rows); ++k)
      sigma_out[k] = sigma_row(*x, V, m, coeffs, static_cast<long>(rows[k]), n_orb, use_hb != 0, eps, sthr);
    return x;
  });
  return ok ? 0 : -1;
}

// ---------------------------------------------------------------- det batch features
// prepare_det_batch, det_batch.cpp:189-319 (checks in the same order, same
// messages).  Spin-orbital order is the block one: beta orbital mo is n_orb + mo.
int orc_prepare_det_batch(const u64* dets, size_t B, u64 ref_a, u64 ref_b, int n_orb, int n_alpha, int n_beta,
                          int kmax, int flags, int32_t* occ, int8_t* k_out, int8_t* ph_out, int32_t* holes,
                          int32_t* parts, uint8_t* mask, int32_t* hpos, int32_t* ppos) {
  void* ok = guard([&]() -> void* {
    const bool need_k = (flags & 1) || (flags & 4), need_ph = flags & 2, need_hp = flags & 4, need_pos = flags & 8;
    if (n_orb < 0 || n_orb > 64) throw std::invalid_argument("n_orb out of range [0,64]");
    if (need_pos && !need_hp) throw std::invalid_argument("need_hp_pos requires need_hp");
    if (need_hp && kmax < 0) throw std::invalid_argument("kmax must be non-negative");
    const u64 m = low_mask(n_orb);
    const u64 ra = ref_a & m, rb = ref_b & m;
    // reference occupied / virtual position maps (det_batch.cpp:105-143)
    std::vector<int32_t> opos(128, -1), vpos(128, -1);
    if (need_pos) {
      if (popc(ra) != n_alpha || popc(rb) != n_beta) throw std::invalid_argument("reference electron count mismatch");
      int p = 0;
      for (int o = 0; o < n_orb; ++o)
        if (ra >> o & 1) opos[o] = p++;
      for (int o = 0; o < n_orb; ++o)
        if (rb >> o & 1) opos[n_orb + o] = p++;
      int v = 0;
      for (int so = 0; so < 2 * n_orb; ++so)
        if (opos[so] < 0) vpos[so] = v++;
    }
    auto spin_sign = [&](u64 r, u64 d) {  // det_batch.cpp:61-97
      u64 h = r & ~d, q = d & ~r;
      if (popc(h) != popc(q)) throw std::invalid_argument("holes != parts in spin channel");
      u64 o = r;
      int n = 0;
      for (int x = 63; x >= 0; --x)
        if (h >> x & 1) {
          n += popc(o & low_mask(x));
          o ^= 1ULL << x;
        }
      for (int x = 0; x < 64; ++x)
        if (q >> x & 1) {
          n += popc(o & low_mask(x));
          o ^= 1ULL << x;
        }
      return (n & 1) ? -1 : 1;
    };
    const int ne = n_alpha + n_beta;
    for (size_t i = 0; i < B; ++i) {
      const u64 a = dets[2 * i], b = dets[2 * i + 1];
      if ((a & ~m) || (b & ~m)) throw std::invalid_argument("det bits outside n_orb range");
      if (popc(a) != n_alpha || popc(b) != n_beta) throw std::invalid_argument("electron count mismatch");
      int j = 0;
      for (int o = 0; o < n_orb; ++o)
        if (a >> o & 1) occ[i * ne + j++] = o;
      for (int o = 0; o < n_orb; ++o)
        if (b >> o & 1) occ[i * ne + j++] = n_orb + o;
      const u64 ha = ra & ~a, hb = rb & ~b;
      const int kk = popc(ha) + popc(hb);
      if (need_k) k_out[i] = static_cast<int8_t>(std::min(kk, 127));
      if (need_ph) ph_out[i] = static_cast<int8_t>(spin_sign(ra, a) * spin_sign(rb, b));
      if (need_hp) {
        std::vector<int> hl, pl;
        for (int o = n_orb - 1; o >= 0; --o)
          if (ha >> o & 1) hl.push_back(o);
        for (int o = n_orb - 1; o >= 0; --o)
          if (hb >> o & 1) hl.push_back(n_orb + o);
        const u64 pa = a & ~ra, pb = b & ~rb;
        for (int o = 0; o < n_orb; ++o)
          if (pa >> o & 1) pl.push_back(o);
        for (int o = 0; o < n_orb; ++o)
          if (pb >> o & 1) pl.push_back(n_orb + o);
        for (int x = 0; x < kmax; ++x) {
          const size_t at = i * static_cast<size_t>(kmax) + x;
          holes[at] = x < static_cast<int>(hl.size()) ? hl[x] : -1;
          parts[at] = x < static_cast<int>(pl.size()) ? pl[x] : -1;
          mask[at] = x < kk ? 1 : 0;
          if (need_pos) {
            hpos[at] = holes[at] < 0 ? -1 : opos[holes[at]];
            ppos[at] = parts[at] < 0 ? -1 : vpos[parts[at]];
          }
        }
      }
    }
    static int done;
    return &done;
  });
  return ok ? 0 : -1;
}

// FCISpace, det_enum.cpp:44-107: words with k of n bits set in ascending order
// (Gosper's successor), beta words outer.
void* orc_gen_fci(int n_orb, int na, int nb) {
  return guard([&]() -> void* {
    if (n_orb <= 0 || n_orb > 64) throw std::invalid_argument("FCISpace: n_orb out of range");
    if (na < 0 || na > n_orb || nb < 0 || nb > n_orb) throw std::invalid_argument("FCISpace: invalid electrons");
    auto words = [&](int k) {
      std::vector<u64> w;
      if (k == 0) return std::vector<u64>{0};
      u64 x = low_mask(k);
      const u64 top = low_mask(n_orb);
      while (true) {
        w.push_back(x);
        const u64 c = x & (~x + 1), r = x + c;
        if (r == 0 || (r & ~top)) break;
        x = (((x ^ r) >> 2) / c) | r;
      }
      return w;
    };
    const auto wa = words(na), wb = words(nb);
    std::vector<Det> out;
    out.reserve(wa.size() * wb.size());
    for (u64 b : wb)
      for (u64 a : wa) out.push_back(Det{a, b});
    auto* Bd = new Bundle;
    Bd->dets("dets", out);
    return Bd;
  });
}

// ---------------------------------------------------------------- connected / perturbative spaces
// gen_connected_dets (det_space.cpp:219-227 + canonicalize) and
// gen_perturbative_dets (bridge.cpp:278-338; det_space.cpp:233-255 for mode 0,
// ham_utils.cpp:546-701 for modes 1 static and 2 dynamic).
std::vector<Det> sorted_unique(std::vector<Det> v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

void* orc_gen_connected(const u64* dets, size_t n, int n_orb) {
  return guard([&]() -> void* {
    std::vector<Det> out;
    auto V = to_dets(dets, n);
    for (const auto& k : V) {
      singles(k, n_orb, [&](const Det& d) { out.push_back(d); });
      doubles(k, n_orb, [&](const Det& d) { out.push_back(d); });
    }
    auto* B = new Bundle;
    B->dets("dets", sorted_unique(std::move(out)));
    return B;
  });
}

void* orc_gen_perturbative(void* c, const u64* dets, size_t n, int n_orb, const double* psi, int mode, double eps) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    auto V = to_dets(dets, n);
    DetMap mv = DetMap::ordered(V, true);
    if (mode < 0 || mode > 2) throw std::invalid_argument("invalid screen mode");
    std::vector<Det> cand;
    if (mode == 0) {
      for (const auto& k : V) {
        singles(k, n_orb, [&](const Det& d) { cand.push_back(d); });
        doubles(k, n_orb, [&](const Det& d) { cand.push_back(d); });
      }
    } else {
      if (!x || !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
      for (size_t i = 0; i < n; ++i) {
        const Det& bra = V[i];
        singles(bra, n_orb, [&](const Det& k) {
          const double h = x->H.elem(bra, k);
          if (std::abs(h) <= kMatThr) return;
          if (mode == 2 && std::abs(h * psi[i]) < eps) return;
          cand.push_back(k);
        });
        const double cut = mode == 2 ? eps / std::max(std::abs(psi[i]), 1e-12) : eps;
        doubles_hb(bra, n_orb, *x->hb, cut, [&](const Det& k) { cand.push_back(k); });
      }
    }
    std::vector<Det> out;
    for (const auto& d : sorted_unique(std::move(cand)))
      if (mv.find(d) < 0) out.push_back(d);
    auto* B = new Bundle;
    B->dets("dets", out);
    return B;
  });
}

// ---------------------------------------------------------------- EN-PT2
// compute_pt2, local_conn.cpp:247-418, with the semantics of one thread: rows
// in index order, sigma_i accumulated diagonal-first then singles then doubles
// in generation order, v_a accumulated in row order (then generation order).
// The external term is summed over a in ascending (alpha, beta) order (the
// reference iterates an unordered_map).  Returns "e_int", "e_ext", "n_ext",
// "ext_dets" (sorted) and "ext_vals" (v_a).
void* orc_compute_pt2(void* c, const u64* dets, const double* coeffs, size_t n, int n_orb, double e_ref,
                      int use_hb, double eps) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto V = to_dets(dets, n);
    DetMap m = DetMap::ordered(V, false);
    auto clamp = [](double d) {
      if (std::abs(d) < 1e-15) d = d >= 0.0 ? 1e-15 : -1e-15;
      return d;
    };
    std::map<Det, double> ext;
    double e_int = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const Det& bra = V[i];
      const double ci = coeffs[i];
      if (ci * ci < 1e-15) continue;
      const double hii = x->H.diag(bra);
      double sig = hii * ci;
      auto inner = [&](const Det& k) {
        const long j = m.find(k);
        if (j < 0) return;
        const double h = x->H.elem(bra, k);
        if (std::abs(h) > kMatThr) sig += h * coeffs[j];
      };
      singles(bra, n_orb, inner);
      doubles(bra, n_orb, inner);
      const double r = sig - e_ref * ci;
      e_int += (r * r) / clamp(e_ref - hii);
      auto outer = [&](const Det& k) {
        if (m.find(k) >= 0) return;
        const double h = x->H.elem(bra, k);
        if (std::abs(h) <= kMatThr) return;
        if (std::abs(h * ci) < eps) return;
        ext[k] += h * ci;
      };
      singles(bra, n_orb, outer);
      if (use_hb)
        doubles_hb(bra, n_orb, *x->hb, eps / std::max(std::abs(ci), 1e-12), outer);
      else
        doubles(bra, n_orb, outer);
    }
    double e_ext = 0.0;
    std::vector<Det> ed;
    std::vector<double> ev;
    for (const auto& [a, v] : ext) {
      e_ext += (v * v) / clamp(e_ref - x->H.diag(a));
      ed.push_back(a);
      ev.push_back(v);
    }
    auto* B = new Bundle;
    B->put("e_int", std::vector<double>{e_int});
    B->put("e_ext", std::vector<double>{e_ext});
    B->put("n_ext", std::vector<int64_t>{static_cast<int64_t>(ed.size())});
    B->dets("ext_dets", ed);
    B->put("ext_vals", ev);
    return B;
  });
}

int orc_slogdet(const double* A, int64_t B, int n, double* sign, double* logabs) {
  if (n < 1) {
    g_err = "N must be >= 1";
    return -1;
  }
#pragma omp parallel for schedule(static) if (B > 1024)
  for (int64_t b = 0; b < B; ++b)
    slogdet_one(A + static_cast<size_t>(b) * n * n, n, sign + b, logabs + b);
  return 0;
}

int orc_slogdet_bwd(const double* A, const double* cot, int64_t B, int n, double* G) {
#pragma omp parallel for schedule(static) if (B > 1024)
  for (int64_t b = 0; b < B; ++b)
    inv_t_one(A + static_cast<size_t>(b) * n * n, n, cot[b], G + static_cast<size_t>(b) * n * n);
  return 0;
}

int64_t orc_bundle_size(void* b, const char* key) {
  auto* B = static_cast<Bundle*>(b);
  auto it = B->arr.find(key);
  return it == B->arr.end() ? -1 : static_cast<int64_t>(it->second.size());
}
int orc_bundle_copy(void* b, const char* key, void* dst) {
  auto* B = static_cast<Bundle*>(b);
  auto it = B->arr.find(key);
  if (it == B->arr.end()) return -1;
  if (!it->second.empty()) std::memcpy(dst, it->second.data(), it->second.size());
  return 0;
}
void orc_bundle_free(void* b) { delete static_cast<Bundle*>(b); }

}  // extern "C"
