// C-ABI driver around the reference DetNQS C++ core, used ONLY as a test checker.
//
// This file is compiled together with the reference's own hot-path sources, in
// place under /root/reference/csrc, by oracle/build_ref.sh into oracle/_ref/.
// Nothing in the product links it.  It exposes the same operations as the
// reference nanobind bridge (csrc/bridge/bridge.cpp:235-772) through plain C
// functions so that Python tests can drive them with ctypes:
//
//   IntCtx / hb_prepare            bridge.cpp:108-128
//   get_local_connections          bridge.cpp:635-673 -> local_conn.cpp:96-142
//   get_ham_diag / vv / block      bridge.cpp:450-507 -> build_ham.cpp:388-460
//   get_ham_conn / conn_amp        bridge.cpp:509-586 -> build_ham.cpp:462-535
//   get_ham_eff                    bridge.cpp:593-629 -> ham_eff.cpp:64-233
//   compute_variational_energy     bridge.cpp:680-714 -> local_conn.cpp:148-229
//   prepare_det_batch              bridge.cpp:347-444 -> det_batch.cpp:189-319
//   gen_connected_dets             bridge.cpp:262-271 -> det_space.cpp:219-227
//   gen_perturbative_dets          bridge.cpp:278-338 -> det_space.cpp:233-255,
//                                                        ham_utils.cpp:546-701
//   compute_pt2                    bridge.cpp:721-771 -> local_conn.cpp:247-418
//
// det_enum.cpp (gen_fci_dets) is not built: it needs <format>, which this
// image's g++ 11.4 lacks.
//
// Results are returned through an opaque "bundle" of named byte arrays.

#include <detnqs/determinant/det.hpp>
#include <detnqs/determinant/det_batch.hpp>
#include <detnqs/determinant/det_ops.hpp>
#include <detnqs/determinant/det_space.hpp>
#include <detnqs/hamiltonian/build_ham.hpp>
#include <detnqs/hamiltonian/ham_eff.hpp>
#include <detnqs/hamiltonian/ham_eval.hpp>
#include <detnqs/hamiltonian/ham_utils.hpp>
#include <detnqs/hamiltonian/local_conn.hpp>
#include <detnqs/integral/hb_table.hpp>
#include <detnqs/integral/integral_mo.hpp>
#include <detnqs/integral/integral_so.hpp>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include <omp.h>

using namespace detnqs;

namespace {

thread_local std::string g_err;

struct Ctx {
  IntegralMO mo;
  IntegralSO so;
  HamEval ham;
  std::unique_ptr<HeatBathTable> hb;
  Ctx(const std::string& path, int n) : mo(n), so(mo), ham(so) { mo.load_from_fcidump(path); }
};

struct Bundle {
  std::map<std::string, std::vector<unsigned char>> arr;
  template <class T>
  void put(const std::string& k, const std::vector<T>& v) {
    auto& dst = arr[k];
    dst.resize(v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(dst.data(), v.data(), dst.size());
  }
  void put_dets(const std::string& k, const std::vector<Det>& v) {
    std::vector<uint64_t> flat(2 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      flat[2 * i] = v[i].alpha;
      flat[2 * i + 1] = v[i].beta;
    }
    put(k, flat);
  }
  void put_coo(const std::string& p, const COOMatrix& c) {
    put(p + ".row", c.rows);
    put(p + ".col", c.cols);
    put(p + ".val", c.vals);
    put(p + ".shape", std::vector<uint32_t>{c.n_rows, c.n_cols});
  }
};

std::vector<Det> to_dets(const uint64_t* d, size_t n) {
  std::vector<Det> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = Det{d[2 * i], d[2 * i + 1]};
  return v;
}

COOMatrix make_coo(const uint32_t* r, const uint32_t* c, const double* v, size_t nnz,
                   uint32_t nr, uint32_t nc) {
  COOMatrix m;
  m.rows.assign(r, r + nnz);
  m.cols.assign(c, c + nnz);
  m.vals.assign(v, v + nnz);
  m.n_rows = nr;
  m.n_cols = nc;
  return m;
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

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_ctx_create(const char* path, int n_orb) {
  return guard([&]() -> void* { return new Ctx(path, n_orb); });
}
void ref_ctx_destroy(void* c) { delete static_cast<Ctx*>(c); }
double ref_ctx_e_nuc(void* c) { return static_cast<Ctx*>(c)->so.get_e_nuc(); }
int ref_ctx_nelec(void* c) { return static_cast<Ctx*>(c)->mo.n_elecs; }

double ref_get_h1e(void* c, int p, int q) { return static_cast<Ctx*>(c)->mo.get_h1e(p, q); }
double ref_get_h2e(void* c, int p, int q, int r, int s) {
  return static_cast<Ctx*>(c)->mo.get_h2e(p, q, r, s);
}

int ref_hb_prepare(void* c, double thr) {
  return guard([&]() -> void* {
           auto* x = static_cast<Ctx*>(c);
           x->hb = std::make_unique<HeatBathTable>(x->so, HBBuildOptions{.threshold = thr});
           x->hb->build();
           return x;
         }) != nullptr
             ? 0
             : -1;
}

// Exports the HB table as CSR: row_offsets[num_rows+1], a, b, w.
void* ref_hb_export(void* c) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (!x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    const int nso = x->hb->n_so();
    std::vector<uint64_t> off{0};
    std::vector<int32_t> a, b;
    std::vector<double> w;
    for (int i = 1; i < nso; ++i) {
      for (int j = 0; j < i; ++j) {
        auto rv = x->hb->row_view(i, j);
        for (size_t k = 0; k < rv.len; ++k) {
          a.push_back(rv.a[k]);
          b.push_back(rv.b[k]);
          w.push_back(rv.w[k]);
        }
        off.push_back(a.size());
      }
    }
    auto* B = new Bundle;
    B->put("row_offsets", off);
    B->put("a", a);
    B->put("b", b);
    B->put("w", w);
    return B;
  });
}

void* ref_local_connections(void* c, const uint64_t* dets, size_t n, int n_orb, int use_hb,
                            double eps1) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto v = to_dets(dets, n);
    auto r = get_local_connections(std::span<const Det>(v), x->ham, n_orb,
                                   use_hb ? x->hb.get() : nullptr, eps1, use_hb != 0);
    auto* B = new Bundle;
    B->put("offsets", r.offsets);
    B->put_dets("dets", r.dets);
    B->put("values", r.values);
    return B;
  });
}

// Streaming count + checksum over rows (used only to time the reference on
// workloads whose full output would not fit in host memory).
int ref_local_conn_stream(void* c, const uint64_t* dets, size_t n, int n_orb, int use_hb,
                          double eps1, uint64_t* nnz_out, double* valsum_out) {
  auto* x = static_cast<Ctx*>(c);
  if (use_hb && !x->hb) {
    g_err = "Heat-Bath table not initialized";
    return -1;
  }
  uint64_t nnz = 0;
  double vs = 0.0;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : nnz, vs)
  for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(n); ++i) {
    Det bra{dets[2 * i], dets[2 * i + 1]};
    auto row = get_local_conn(bra, x->ham, n_orb, use_hb ? x->hb.get() : nullptr, eps1,
                              use_hb != 0);
    nnz += row.values.size();
    for (double v : row.values) vs += v;
  }
  *nnz_out = nnz;
  *valsum_out = vs;
  return 0;
}

void* ref_ham_diag(void* c, const uint64_t* dets, size_t n) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    auto v = to_dets(dets, n);
    auto d = get_ham_diag(v, x->ham);
    auto* B = new Bundle;
    B->put("diag", d);
    return B;
  });
}

int ref_compute_elem(void* c, const uint64_t* bra, const uint64_t* ket, double* out) {
  auto* x = static_cast<Ctx*>(c);
  *out = x->ham.compute_elem(Det{bra[0], bra[1]}, Det{ket[0], ket[1]});
  return 0;
}

// analyze_excitation (det_ops.cpp:159-196): out = {degree, n_alpha_exc, n_beta_exc}, phase.
int ref_analyze(const uint64_t* bra, const uint64_t* ket, int* out3, double* phase) {
  auto info = det_ops::analyze_excitation(Det{bra[0], bra[1]}, Det{ket[0], ket[1]});
  out3[0] = info.degree;
  out3[1] = info.n_alpha_exc;
  out3[2] = info.n_beta_exc;
  *phase = info.phase;
  return 0;
}

void* ref_ham_vv(void* c, const uint64_t* dets, size_t n, int n_orb) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    auto v = to_dets(dets, n);
    auto coo = get_ham_vv(v, x->ham, n_orb);
    auto* B = new Bundle;
    B->put_coo("H", coo);
    return B;
  });
}

void* ref_ham_block(void* c, const uint64_t* dv, size_t nv, const uint64_t* dp, size_t np,
                    int n_orb) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    auto V = to_dets(dv, nv);
    auto P = to_dets(dp, np);
    std::optional<std::span<const Det>> ps;
    if (dp) ps = std::span<const Det>(P);
    auto blk = get_ham_block(V, ps, x->ham, n_orb);
    auto* B = new Bundle;
    B->put_coo("H_VV", blk.H_VV);
    B->put_coo("H_VP", blk.H_VP);
    B->put_dets("det_P", blk.map_P.all_dets());
    return B;
  });
}

void* ref_ham_conn(void* c, const uint64_t* dets, size_t n, int n_orb, int use_hb,
                   double eps1) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto v = to_dets(dets, n);
    auto blk = get_ham_conn(v, x->ham, n_orb, x->hb.get(), eps1, use_hb != 0);
    auto* B = new Bundle;
    B->put_coo("H_VV", blk.H_VV);
    B->put_coo("H_VP", blk.H_VP);
    B->put_dets("det_P", blk.map_P.all_dets());
    return B;
  });
}

void* ref_ham_conn_amp(void* c, const uint64_t* dets, const double* psi, size_t n, int n_orb,
                       double eps1) {
  return guard([&]() -> void* {
    auto* x = static_cast<Ctx*>(c);
    if (!x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto v = to_dets(dets, n);
    std::vector<double> p(psi, psi + n);
    auto blk = get_ham_conn_amp(v, p, x->ham, n_orb, x->hb.get(), eps1);
    auto* B = new Bundle;
    B->put_coo("H_VV", blk.H_VV);
    B->put_coo("H_VP", blk.H_VP);
    B->put_dets("det_P", blk.map_P.all_dets());
    return B;
  });
}

void* ref_ham_eff(const uint32_t* vr, const uint32_t* vc, const double* vv, size_t vnnz,
                  uint32_t vnr, uint32_t vnc, const uint32_t* pr, const uint32_t* pc,
                  const double* pv, size_t pnnz, uint32_t pnr, uint32_t pnc,
                  const double* hpp, size_t np, double e_ref, int reg_sigma, double eps) {
  return guard([&]() -> void* {
    auto VV = make_coo(vr, vc, vv, vnnz, vnr, vnc);
    auto VP = make_coo(pr, pc, pv, pnnz, pnr, pnc);
    std::vector<double> d(hpp, hpp + np);
    HeffConfig cfg{.reg_type = reg_sigma ? Regularization::Sigma : Regularization::LinearShift,
                   .epsilon = eps};
    auto H = get_ham_eff(VV, VP, d, e_ref, cfg);
    auto* B = new Bundle;
    B->put_coo("H", H);
    return B;
  });
}

int ref_variational_energy(void* c, const uint64_t* dets, const double* coeffs, size_t n,
                           int n_orb, int use_hb, double eps1, double* out) {
  auto* x = static_cast<Ctx*>(c);
  void* ok = guard([&]() -> void* {
    auto v = to_dets(dets, n);
    *out = compute_variational_energy(v, std::span<const double>(coeffs, n), x->ham, n_orb,
                                      use_hb ? x->hb.get() : nullptr, eps1, use_hb != 0);
    return x;
  });
  return ok ? 0 : -1;
}

// prepare_det_batch through the bridge's option handling (bridge.cpp:360-411).
int ref_prepare_det_batch(const uint64_t* dets, size_t B, uint64_t ref_a, uint64_t ref_b, int n_orb,
                          int n_alpha, int n_beta, int kmax, int flags, int32_t* occ, int8_t* k,
                          int8_t* ph, int32_t* holes, int32_t* parts, uint8_t* mask, int32_t* hpos,
                          int32_t* ppos) {
  static int done;
  void* ok = guard([&]() -> void* {
    det_batch::PrepareOptions opt;
    opt.kmax = kmax;
    opt.need_hp = (flags & 4) != 0;
    opt.need_k = (flags & 1) || opt.need_hp;
    opt.need_phase = (flags & 2) != 0;
    opt.need_hp_pos = (flags & 8) != 0;
    det_batch::prepare_det_batch(dets, B, Det{ref_a, ref_b}, n_orb, n_alpha, n_beta, opt, occ,
                                 opt.need_k ? k : nullptr, opt.need_phase ? ph : nullptr,
                                 opt.need_hp ? holes : nullptr, opt.need_hp ? parts : nullptr,
                                 opt.need_hp ? reinterpret_cast<bool*>(mask) : nullptr,
                                 opt.need_hp_pos ? hpos : nullptr, opt.need_hp_pos ? ppos : nullptr);
    return &done;
  });
  return ok ? 0 : -1;
}

void* ref_gen_connected(const uint64_t* dets, size_t n, int n_orb) {
  return guard([&]() -> void* {
    auto v = to_dets(dets, n);
    auto out = det_space::canonicalize(det_space::generate_connected(v, n_orb));
    auto* B = new Bundle;
    B->put_dets("dets", out);
    return B;
  });
}

// The body of the bridge's gen_perturbative_dets (bridge.cpp:285-330);
// mode 0 none, 1 static, 2 dynamic.  c may be NULL for mode 0.
void* ref_gen_perturbative(void* c, const uint64_t* dets, size_t n, int n_orb, const double* psi,
                           int mode, double eps1) {
  return guard([&]() -> void* {
    auto V = to_dets(dets, n);
    const DetMap map_V = DetMap::from_ordered({V.begin(), V.end()}, true);
    std::vector<Det> comp;
    if (mode == 0) {
      comp = det_space::generate_complement(std::span<const Det>(V.data(), V.size()), n_orb, map_V, true);
    } else {
      auto* x = static_cast<Ctx*>(c);
      if (!x || !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
      std::span<const double> ps;
      if (mode == 2) ps = std::span<const double>(psi, n);
      comp = generate_complement_screened(std::span<const Det>(V.data(), V.size()), n_orb, map_V, x->ham,
                                          x->hb.get(), mode == 2 ? ScreenMode::Dynamic : ScreenMode::Static,
                                          ps, eps1);
    }
    auto* B = new Bundle;
    B->put_dets("dets", comp);
    return B;
  });
}

// compute_pt2 with `threads` OpenMP threads (1 gives the reference's
// sequential accumulation order).  out3 = {internal, external, n_ext}.
int ref_compute_pt2(void* c, const uint64_t* dets, const double* coeffs, size_t n, int n_orb, double e_ref,
                    int use_hb, double eps1, int threads, double* out3) {
  auto* x = static_cast<Ctx*>(c);
  void* ok = guard([&]() -> void* {
    if (use_hb && !x->hb) throw std::invalid_argument("Heat-Bath table not initialized");
    auto v = to_dets(dets, n);
    const int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    Pt2Result r{};
    try {
      r = compute_pt2(v, std::span<const double>(coeffs, n), x->ham, n_orb, e_ref,
                      use_hb ? x->hb.get() : nullptr, eps1, use_hb != 0);
    } catch (...) {
      omp_set_num_threads(saved);
      throw;
    }
    omp_set_num_threads(saved);
    out3[0] = r.e_pt2_internal;
    out3[1] = r.e_pt2_external;
    out3[2] = static_cast<double>(r.n_ext);
    return x;
  });
  return ok ? 0 : -1;
}

// Bundle accessors.
int64_t ref_bundle_size(void* b, const char* key) {
  auto* B = static_cast<Bundle*>(b);
  auto it = B->arr.find(key);
  return it == B->arr.end() ? -1 : static_cast<int64_t>(it->second.size());
}
int ref_bundle_copy(void* b, const char* key, void* dst) {
  auto* B = static_cast<Bundle*>(b);
  auto it = B->arr.find(key);
  if (it == B->arr.end()) return -1;
  if (!it->second.empty()) std::memcpy(dst, it->second.data(), it->second.size());
  return 0;
}
void ref_bundle_free(void* b) { delete static_cast<Bundle*>(b); }

}  // extern "C"
