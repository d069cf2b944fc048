// Multi-instance fused local energy: the device-side description of a launch
// that covers row ranges of many independent instances (each with its own
// integral context, selected space V, DetMap and amplitudes).
#pragma once
#include <memory>
#include <vector>

#include "ctx.hpp"
#include "detnqs_hip.h"
#include "hash.hpp"
#include "strings.hpp"

namespace dq {

// One row range [row_begin, row_begin + rows) of one instance.
struct SegDev {
  DevView V;
  HashView H;
  const u64* dets;      // V of the instance, n_rows x 2 words
  const double* coeffs;  // psi_V
  double* sigma;        // (H psi)_i, written for the range
  const u32* hb_len;    // kept-prefix lengths for eps1 (heat-bath mode)
  int* err;             // set to 1 if a row has orbitals at or above n_orb
  StrIndex X;           // string index of V (V-driven method)
  StrCoef C;            // psi in member order (V-driven method)
  int64_t row_begin;
  double sthr;          // single-excitation threshold
  double dmin;          // heat-bath mode: smallest kept |double|, else 0
  int n_orb;
  int dbl_mode;         // 0 exhaustive doubles, 1 heat-bath
  int method;           // DQ_ELOC_STRINGS, DQ_ELOC_WALK or kElocVv
  // kElocVv: lane-ordered H_VV of V (vvcache.hpp), rows in rounds of 64 slots
  const int64_t* vv_off;
  const u32* vv_col;
  const double* vv_val;
};

// Segment method: a product over the lane-ordered H_VV (bit-identical to the
// string method on the same V).
constexpr int kElocVv = 2;

// An instance whose energy sum_i c_i sigma_i is formed after the launch's rows:
// blocks of 256 rows reduced by a tree, block sums added in order (the same
// arithmetic as the single-instance path, so the result is bit-identical).
struct EnergySeg {
  const double* coeffs;
  const double* sigma;
  int64_t n_rows;
  int slot;  // index into the caller's energy array
};

// psi of one instance gathered into member order at the start of a launch.
struct GatherSeg {
  const u32* a_row;
  const u32* b_row;
  const double* c;
  double* ca;
  double* cb;
};

// Device pointers of one launch (host copy holds the sizes).
struct PlanDev {
  const GatherSeg* gsegs = nullptr;
  const int64_t* gstart = nullptr;  // n_gather + 1 prefix of member counts
  int n_gather = 0;
  int64_t gather_total = 0;
  const SegDev* segs = nullptr;
  const int64_t* start = nullptr;   // nseg + 1 row-item prefix
  // segments whose rows are a product over a lane-ordered H_VV (kElocVv),
  // run by their own kernel before the walk/string kernel
  const SegDev* vsegs = nullptr;
  const int64_t* vstart = nullptr;  // n_vseg + 1 row-item prefix
  int n_vseg = 0;
  int64_t vtotal = 0;
  const EnergySeg* esegs = nullptr;
  const int64_t* estart = nullptr;  // n_energy + 1 prefix of 256-row chunks
  double* part = nullptr;           // n_chunks partial sums
  unsigned long long* counter = nullptr;
  int* err = nullptr;
  int nseg = 0, n_energy = 0;
  int64_t total = 0, n_chunks = 0;
  int methods = 1;  // bit 0: string method present, bit 1: walk present
};

void launch_eloc(const PlanDev& P, const dq_eloc_tuning& t, double* d_energy, hipStream_t s);
void validate_tuning(const dq_eloc_tuning& t);
// Host-side segment for one instance range; throws on invalid input.
SegDev make_segment(const dq_eloc_segment& s);
// Screening thresholds of an instance: single-excitation |v| bound and the
// smallest kept |double| (heat-bath mode).
void eloc_thresholds(const dq_ctx& c, int use_hb, double eps1, double* sthr, double* dmin);
int default_grid_blocks(int waves_per_block);

}  // namespace dq

struct dq_detmap {
  dq::DetHash h;
  dq::StrIndexBuf x;
  int64_t n = 0;
};

struct dq_eloc_plan {
  dq::DevBuf<dq::GatherSeg> gsegs;
  dq::DevBuf<int64_t> gstart;
  std::vector<std::unique_ptr<dq::DevBuf<double>>> gbufs;  // ca, cb per gathered instance
  dq::DevBuf<dq::SegDev> segs, vsegs;
  dq::DevBuf<int64_t> vstart;
  dq::DevBuf<int64_t> start, estart;
  dq::DevBuf<dq::EnergySeg> esegs;
  dq::DevBuf<double> part;
  dq::DevBuf<unsigned long long> counter;
  dq::DevBuf<int> err;
  dq::PlanDev dev;
  dq_eloc_tuning tune{4, 0, 0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double* graph_energy = nullptr;
  ~dq_eloc_plan();
};
