// Integral context: host copies plus HBM-resident tables for the kernels.
#pragma once

#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "dq_common.hpp"

namespace dq {

// Read-only view passed by value to the assembly kernels.
//
// All FP64 tables hold exact copies of packed integrals, so every matrix
// element is formed from the same operands, in the same order, as the
// reference's HamEval (ham_eval.cpp:33-147):
//   h1[r*n+a]        = h_ra
//   Js[(r*n+k)*n+a]  = (r a|k k)        single excitation r->a, direct term
//   Ks[(r*n+k)*n+a]  = (r k|k a)        single excitation r->a, exchange term
//   Jd[p*n+q]        = (p p|q q)        diagonal, direct
//   Kd[p*n+q]        = (p q|q p)        diagonal, exchange
//   h2               = packed (pq|rs)   doubles without a heat-bath table
// Heat-bath rows (one per spin-orbital pair i>j, id i(i-1)/2+j) store the
// added pair packed as a | b<<16 (SO indices, a>b) and the signed double
// excitation value V that HamEval::compute_double returns for that
// excitation; the matrix element is phase * V.
struct DevView {
  int n = 0;
  const double* h1 = nullptr;
  const double* h2 = nullptr;
  const double* Js = nullptr;
  const double* Ks = nullptr;
  const double* Jd = nullptr;
  const double* Kd = nullptr;
  // heat-bath
  int nso = 0;
  u32 hb_n = 0;  // stored entries (hb_off[last])
  const u32* hb_off = nullptr;
  const u32* hb_ab = nullptr;
  const double* hb_v = nullptr;
  const double* hb_w = nullptr;
  // Pair masks, one row per SO pair i>j (id i(i-1)/2+j), n words per row:
  // same spin s: pm[row*n + a] has bit b (b < a) set iff removing {i,j} and
  // adding {a,b} (spin s) gives |element| > 1e-15; mixed spin: pm[row*n + a]
  // (a alpha) has bit b (b beta) set under the same rule.  pm_cnt[row] is the
  // number of such pairs when every orbital other than i, j is virtual.
  const u64* pm = nullptr;
  const u32* pm_cnt = nullptr;
};

// hb_len entries carry this flag when the row's kept prefix is the whole
// pair-mask set, so a count can use the masks instead of the entries.
constexpr u32 kFullRow = 0x80000000u;

struct HostHB {
  int nso = 0;
  std::vector<u32> off;  // nrows + 1
  std::vector<int32_t> a, b;
  std::vector<double> w, v;
};

// Kept-prefix lengths of the heat-bath rows for one cutoff: uploaded once on
// the stream of the call that first asked for them; `ready` marks the end of
// that upload, and every later user waits for it on its own stream.
struct HbLengths {
  double eps = 0.0;
  std::vector<u32> host;  // stays alive until the upload has completed
  DevBuf<u32> dev;
  hipEvent_t ready = nullptr;
  HbLengths() = default;
  HbLengths(const HbLengths&) = delete;
  HbLengths& operator=(const HbLengths&) = delete;
  ~HbLengths() {
    if (ready) {
      (void)hipEventSynchronize(ready);
      (void)hipEventDestroy(ready);
    }
  }
};

}  // namespace dq

struct dq_ctx {
  int n_orb = 0;
  int n_elec = 0;
  double e_nuc = 0.0;
  int device = 0;
  std::vector<double> h1e, h2e;  // packed, host
  // device tables
  dq::DevBuf<double> d_h1, d_h2, d_Js, d_Ks, d_Jd, d_Kd;
  dq::DevBuf<dq::u64> d_pm;
  dq::DevBuf<dq::u32> d_pm_cnt;
  std::vector<dq::u32> pm_cnt;
  // heat-bath
  std::unique_ptr<dq::HostHB> hb;
  double hb_thr = 0.0;  // build threshold: the table holds excitations with weight >= hb_thr
  dq::DevBuf<dq::u32> d_hb_off, d_hb_ab;
  dq::DevBuf<double> d_hb_v, d_hb_w;
  // kept-prefix lengths, one device array per eps used; pointers stay valid
  // until the heat-bath table is rebuilt or cleared.  hb_mu guards the list:
  // calls on one context from several threads may ask for new cutoffs at once.
  std::mutex hb_mu;
  std::vector<std::unique_ptr<dq::HbLengths>> hb_len_cache;
  // set by kernels that meet a determinant with orbitals at or above n_orb
  dq::DevBuf<int> d_err;

  dq::DevView view() const;
  double g1(int p, int q) const { return h1e[dq::key1(p, q)]; }
  double g2(int p, int q, int r, int s) const { return h2e[dq::key2(p, q, r, s)]; }
  // Device array of per-row kept-prefix lengths for cutoff eps (cached per eps).
  // Work queued on `s` after this call sees the array; with host_wait the
  // array is also complete when the call returns (for launches on streams
  // not known yet).  Thread-safe; rebuilding or clearing the heat-bath table
  // while other calls use the context is not.
  const dq::u32* hb_lengths(double eps, hipStream_t s, bool host_wait = false);
  // Throws DQ_EINVAL (and clears the flag) if a kernel flagged a bad row.
  void check_rows(hipStream_t s);
};

namespace dq {
void upload_tables(dq_ctx& c);
void build_hb(dq_ctx& c, double thr);
}  // namespace dq
