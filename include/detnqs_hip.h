/*
 * detnqs_hip.h — C-ABI of libdetnqs_hip.so, the MI355X (gfx950) local-energy
 * path of DetNQS: Slater–Condon row assembly (heat-bath screened), projection
 * onto the selected space, H_eff down-folding, fused energy, and the batched
 * FP64 slogdet forward/backward.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure;
 *    dq_last_error() returns the message of the calling thread's last failure.
 *    Codes: DQ_EINVAL (bad argument, the reference's ValueError /
 *    invalid_argument / ffi kInvalidArgument), DQ_ERANGE (index out of range),
 *    DQ_ERUNTIME (HIP or allocation failure, ffi kInternal).
 *  - Pointers named d_* are device pointers (HBM), h_* are host pointers.
 *  - `stream` is a hipStream_t passed as void* (0 = the null stream).  Entry
 *    points that produce fixed-size outputs are asynchronous on `stream`;
 *    those that must learn an output size synchronise `stream` (stated below).
 *  - Determinants are pairs of uint64 words (alpha, beta), bit i = spatial
 *    orbital i (csrc/detnqs/determinant/det.hpp:30-36); n_orb <= 64.
 *  - Numeric failure is data, not an error: a singular matrix gives
 *    sign = 0, logabs = -inf and a zero gradient (logdet_cuda.cu:249-289,374-377).
 */
#ifndef DETNQS_HIP_H
#define DETNQS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DQ_OK 0
#define DQ_EINVAL (-1)
#define DQ_ERANGE (-2)
#define DQ_ERUNTIME (-3)
#define DQ_ETIMEDOUT (-4)

typedef struct dq_ctx dq_ctx;       /* integrals + heat-bath table, device resident */
typedef struct dq_blocks dq_blocks; /* device-resident H_VV / H_VP / P-space result */

const char* dq_last_error(void);
int dq_version(void);
/* Large temporaries and dq_detset results come from the library's per-device
 * block cache, which keeps released blocks for reuse (a reuse on another stream
 * waits for the work queued before the release); dq_trim_memory waits for
 * those releases and returns every cached block to the driver.  No
 * counterpart in the reference (host memory only). */
int dq_trim_memory(void);

/* ---- integral context --------------------------------------------------
 * Replaces IntCtx(fcidump_path, num_orb) (csrc/bridge/bridge.cpp:108-128,
 * 242-248; integral_mo.cpp:25-182).  dq_ctx_create takes the packed arrays
 * h1e[n(n+1)/2], h2e[n1e(n1e+1)/2] in the reference's key order
 * (integral_mo.hpp:30-49).  Uses the current HIP device. */
int dq_ctx_create_fcidump(const char* path, int n_orb, dq_ctx** out);
int dq_ctx_create(int n_orb, int n_elec, const double* h_h1e, const double* h_h2e, double e_nuc,
                  dq_ctx** out);
int dq_ctx_destroy(dq_ctx* ctx);
int dq_ctx_info(const dq_ctx* ctx, int* n_orb, int* n_elec, double* e_nuc);

/* IntCtx.hb_prepare(threshold) / hb_clear() (bridge.cpp:115-121;
 * hb_table.cpp:39-168).  Rows are ordered exactly as the reference's. */
int dq_hb_prepare(dq_ctx* ctx, double threshold);
int dq_hb_clear(dq_ctx* ctx);
/* Number of stored entries (0 when no table); optionally exports the table
 * to host arrays of that size (row_offsets has n_so(n_so-1)/2 + 1 entries). */
int dq_hb_export(const dq_ctx* ctx, int64_t* n_entries, uint64_t* h_row_offsets, int32_t* h_a,
                 int32_t* h_b, double* h_w);

/* ---- diagonal ------------------------------------------------------------
 * get_ham_diag (bridge.cpp:450-457; build_ham.cpp:388-400; ham_eval.cpp:53-86). */
int dq_ham_diag(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, double* d_out, void* stream);

/* ---- local connections ----------------------------------------------------
 * get_local_connections (bridge.cpp:635-673; local_conn.cpp:36-142), split into
 * a count pass and a fill pass so that the caller owns every buffer.
 *   count: d_offsets[n+1] receives the exclusive prefix sum of the row sizes
 *          (int64, the reference's offsets are int32: local_conn.hpp:38);
 *          if h_total != NULL the stream is synchronised and *h_total = nnz.
 *   fill : d_out_dets[nnz][2], d_out_vals[nnz] in the reference's row order
 *          (diagonal, singles a then b, doubles in generation / heat-bath order);
 *          row r is written at index d_offsets[r] - d_offsets[0], so passing
 *          (d_dets + 2*start, d_offsets + start) streams a chunk of rows into a
 *          buffer sized for that chunk.
 * Both calls must get the same dets / n_orb / use_hb / eps1. */
int dq_local_conn_count(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, int n_orb, int use_hb,
                        double eps1, int64_t* d_offsets, int64_t* h_total, void* stream);
int dq_local_conn_fill(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, int n_orb, int use_hb,
                       double eps1, const int64_t* d_offsets, uint64_t* d_out_dets, double* d_out_vals,
                       void* stream);

/* ---- projected assembly (H_VV, H_VP, P space) -----------------------------
 * mode 0: KnownSets  -> get_ham_vv / get_ham_block (bridge.cpp:459-507;
 *                       build_ham.cpp:402-460); d_ext may be NULL.
 * mode 1: StaticHB   -> get_ham_conn (bridge.cpp:509-541; build_ham.cpp:462-495).
 * mode 2: DynamicAmp -> get_ham_conn_amp (bridge.cpp:547-586;
 *                       build_ham.cpp:497-535); needs d_psi[n] and a HB table.
 * Outputs are canonical COO sorted by (row, col) (ham_utils.cpp:41-83), P sorted
 * ascending (build_ham.cpp:346-380).  Synchronises `stream`. */
int dq_ham_blocks(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, const uint64_t* d_ext,
                  int64_t n_ext, int n_orb, int mode, int use_hb, double eps1, const double* d_psi,
                  void* stream, dq_blocks** out);
int dq_blocks_sizes(const dq_blocks* b, int64_t* nnz_vv, int64_t* nnz_vp, int64_t* n_p);
/* Copies the result; pointers may be NULL to skip an array.  dst_on_device
 * selects device (1) or host (0) destinations.  Synchronises `stream`. */
int dq_blocks_copy(const dq_blocks* b, uint32_t* vv_row, uint32_t* vv_col, double* vv_val,
                   uint32_t* vp_row, uint32_t* vp_col, double* vp_val, uint64_t* det_p,
                   int dst_on_device, void* stream);
int dq_blocks_free(dq_blocks* b);

/* ---- effective Hamiltonian (SpGEMM down-folding) ---------------------------
 * get_ham_eff (bridge.cpp:593-629; ham_eff.cpp:35-233): H_eff = H_VV + H_VP
 * diag(dinv) H_VP^T, reg_sigma = 1 for 'sigma' (d/(d^2+eps^2)), 0 for
 * 'linear_shift' (1/(d + eps sgn d)).  Inputs are device COO arrays; the result
 * is a dq_blocks whose H_VV part holds H_eff (n_p = 0).  Synchronises. */
int dq_ham_eff(const uint32_t* d_vv_row, const uint32_t* d_vv_col, const double* d_vv_val,
               int64_t nnz_vv, const uint32_t* d_vp_row, const uint32_t* d_vp_col,
               const double* d_vp_val, int64_t nnz_vp, int64_t n_v, const double* d_h_pp,
               int64_t n_p, double e_ref, int reg_sigma, double epsilon, void* stream,
               dq_blocks** out);

/* ---- fused local energy ------------------------------------------------------
 * compute_variational_energy (bridge.cpp:680-714; local_conn.cpp:148-229):
 * sigma_i = sum_{j in V} H_ij c_j (rows with c_i^2 < 1e-12 give 0) into
 * d_sigma[n] (may be NULL); *h_energy = sum_i c_i sigma_i (fixed order).
 * Synchronises. */
int dq_variational_energy(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs,
                          int64_t n, int n_orb, int use_hb, double eps1, double* d_sigma,
                          double* h_energy, void* stream);
/* The same with an explicit method (DQ_ELOC_STRINGS or DQ_ELOC_WALK); no
 * counterpart in the reference, which has the walk only. */
int dq_variational_energy_method(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs,
                                 int64_t n, int n_orb, int use_hb, double eps1, int method,
                                 double* d_sigma, double* h_energy, void* stream);

/* Allocation-free compute_variational_energy (string method).  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Results are those of dq_variational_energy_method with
 * DQ_ELOC_STRINGS: d_sigma[n] (may be NULL) and *d_energy, a DEVICE double,
 * written on `stream`.  Every temporary lives in a library-owned static
 * device array of DQ_NOALLOC_SCRATCH_BYTES per GPU (held until the process
 * ends); at most DQ_NOALLOC_MAX_ROWS rows (DQ_ERANGE beyond).  Calls that
 * overlap on one GPU share that array and race: serialise them (one stream,
 * or the caller's own ordering).  Not capturing: synchronises to report rows
 * with orbitals at or above n_orb; capturing: asynchronous. */
#define DQ_NOALLOC_MAX_ROWS 262144
#define DQ_NOALLOC_SCRATCH_BYTES (64u << 20)
int dq_variational_energy_noalloc(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs,
                                  int64_t n, int n_orb, int use_hb, double eps1, double* d_sigma,
                                  double* d_energy, void* stream);
/* The limits above, for bindings that cannot read the macros. */
int dq_noalloc_limits(int64_t* max_rows, int64_t* scratch_bytes);

/* ---- multi-instance fused local energy (serving and batched forms) ----------
 * No counterpart in the reference, which evaluates one instance per call
 * (local_conn.cpp:148-229).  A plan covers row ranges of many instances, each
 * with its own context, selected space V (device dets + its DetMap) and
 * amplitudes; one launch computes sigma_i = sum_{j in V} H_ij c_j for every
 * row of every range with the arithmetic of dq_variational_energy, so a row's
 * sigma is bit-identical whatever else shares the launch.  A segment with
 * energy_slot >= 0 also gets d_energy[energy_slot] = sum_i c_i sigma_i over all
 * n_rows of its instance, formed after the launch's rows (every row of sigma
 * must be written by then, by this or an earlier launch on the stream); its
 * value is bit-identical to dq_variational_energy's. */
/* Energy methods.  DQ_ELOC_WALK generates every connected determinant of a row
 * and looks it up in V (local_conn.cpp:148-229).  DQ_ELOC_STRINGS (default)
 * visits only the determinants of V that share an alpha or beta word with the
 * row or with one of its alpha single excitations, through the string index of
 * V; it finds the same (y, H_xy) pairs with the same screening, and sums them
 * in another order (results agree to rounding, 1e-12 relative). */
#define DQ_ELOC_STRINGS 0
#define DQ_ELOC_WALK 1

typedef struct dq_detmap dq_detmap;
/* DetMap of d_dets[n][2] (det_space.cpp:29-65) and the string index of the
 * same V; strict = 1 rejects duplicates (from_ordered), 0 keeps the first
 * occurrence.  Synchronises the stream. */
int dq_detmap_build(const uint64_t* d_dets, int64_t n, int strict, void* stream, dq_detmap** out);
int dq_detmap_free(dq_detmap* m);

/* ---- lane-ordered H_VV (repeated energies on one V) --------------------------
 * No counterpart in the reference's fused energy, which regenerates every
 * element per call (local_conn.cpp:148-229); its driver instead builds H_VV once
 * per outer step and multiplies it in every inner step (operator/kernel.py:
 * 48-145).  The cache holds the elements dq_variational_energy_method
 * (DQ_ELOC_STRINGS) would generate for V, in the order the string kernel adds
 * them: row i owns rounds [off[i], off[i+1]) of 64 slots; lane l's j-th kept
 * pair is slot 64 (off[i] + j) + l; unused slots have col 0xFFFFFFFF.  col is a
 * row of V (the first occurrence; the diagonal names row i), val is H_ij.
 * dq_vvcache_energy then gives the same sigma and energy bits as
 * dq_variational_energy_method on any coefficients.  Build synchronises; the
 * cache keeps device pointers to nothing but its own arrays (ctx and vmap are
 * only checked against segments).  Fewer than 2^31 rows. */
typedef struct dq_vvcache dq_vvcache;
int dq_vvcache_build(const dq_ctx* ctx, const dq_detmap* vmap, const uint64_t* d_dets, int64_t n,
                     int n_orb, int use_hb, double eps1, void* stream, dq_vvcache** out);
/* rounds: 64-slot rounds (layout size 64 * rounds); nnz: kept pairs. */
int dq_vvcache_info(const dq_vvcache* vv, int64_t* n_rows, int64_t* rounds, int64_t* nnz);
/* Copies off[n+1] (int64), col[64 rounds], val[64 rounds] (any may be NULL). */
int dq_vvcache_export(const dq_vvcache* vv, int64_t* d_off, uint32_t* d_col, double* d_val, void* stream);
/* sigma (may be NULL) and *h_energy (may be NULL; synchronises if given). */
int dq_vvcache_energy(const dq_vvcache* vv, const double* d_coeffs, double* d_sigma, double* h_energy,
                      void* stream);
int dq_vvcache_free(dq_vvcache* vv);

typedef struct dq_eloc_segment {
  const dq_ctx* ctx;       /* integrals (+ heat-bath table when use_hb) */
  const dq_detmap* vmap;   /* DetMap of the instance's V (n_rows entries) */
  const uint64_t* d_dets;  /* V, n_rows x 2 words */
  const double* d_coeffs;  /* psi_V, n_rows */
  double* d_sigma;         /* out, n_rows (rows [row_begin, row_end) written) */
  int64_t n_rows;
  int64_t row_begin, row_end;
  int n_orb;
  int use_hb;
  double eps1;
  int energy_slot;         /* -1: none */
  int method;              /* DQ_ELOC_STRINGS (0) or DQ_ELOC_WALK (1) */
  const struct dq_vvcache* vv; /* NULL, or the lane-ordered H_VV of this V
                                  (dq_vvcache_build with the same ctx, vmap,
                                  n_orb, use_hb, eps1; string method only):
                                  the rows are then a product over it, with
                                  the same bits */
} dq_eloc_segment;

/* Launch shape: waves_per_block in {1,2,4,8}; persistent = 1 runs grid_blocks
 * blocks (0: 32 waves per CU) whose waves pull rows from a work counter,
 * persistent = 0 runs one wave per row. */
typedef struct dq_eloc_tuning {
  int waves_per_block;
  int persistent;
  int grid_blocks;
} dq_eloc_tuning;

typedef struct dq_eloc_plan dq_eloc_plan;
/* Validates the segments and uploads the launch description (synchronous).
 * The plan keeps device pointers: contexts, DetMaps, H_VV caches and buffers must outlive it,
 * and a context's heat-bath table must not be rebuilt while a plan uses it. */
int dq_eloc_plan_create(const dq_eloc_segment* segs, int n_segs, const dq_eloc_tuning* tune,
                        dq_eloc_plan** out);
int dq_eloc_plan_info(const dq_eloc_plan* plan, int64_t* total_rows, int* n_segments, int* n_energy,
                      dq_eloc_tuning* tune);
/* Asynchronous launch on the stream; allocates nothing. */
int dq_eloc_plan_launch(dq_eloc_plan* plan, double* d_energy, void* stream);
/* Captures the launch (kernels only: counter reset, row kernel, energy
 * reduction) into a HIP graph owned by the plan; dq_eloc_plan_graph_launch
 * replays it.  Launches of one plan on overlapping streams race. */
int dq_eloc_plan_capture(dq_eloc_plan* plan, double* d_energy);
int dq_eloc_plan_graph_launch(dq_eloc_plan* plan, void* stream);
/* Synchronises; DQ_EINVAL if a launch met a determinant with orbitals at or
 * above n_orb (the flag is then cleared). */
int dq_eloc_plan_check(dq_eloc_plan* plan, void* stream);
int dq_eloc_plan_free(dq_eloc_plan* plan);

/* ---- serving loop ----------------------------------------------------------------
 * No counterpart in the reference.  A server owns a loop thread, two streams
 * and an HBM arena of hbm_budget_bytes.  Requests (one compute_variational_energy
 * instance each) are admitted in arrival order while their device memory fits
 * in the arena, staged on a copy stream, and computed by multi-instance
 * launches of at most rows_per_launch rows in which every running request gets
 * an equal share of rows (continuous batching).  Results are bit-identical to
 * dq_variational_energy on the same request.  V sets submitted with v_key != 0
 * stay resident (determinants + DetMap) for later requests with the same
 * (ctx, v_key, n_rows) until the arena needs the space; the caller keeps a key
 * bound to one V.  Contexts must outlive the server and keep their heat-bath
 * table unchanged while requests use it. */
typedef struct dq_server dq_server;
typedef struct dq_server_config {
  int64_t hbm_budget_bytes;
  int64_t rows_per_launch;
  int max_inflight;        /* launches queued on the device at once, 1..8 */
  dq_eloc_tuning tune;
  int vv_layout;           /* 1: a V set kept under a v_key gets its lane-ordered
                              H_VV (dq_vvcache layout) in the arena, built before
                              its first request's rows, and every request on it
                              with the same n_orb, use_hb, eps1 (string method)
                              is a product over it, with the same bits; 0: never */
} dq_server_config;

typedef struct dq_server_request {
  const dq_ctx* ctx;
  const uint64_t* dets;    /* n_rows x 2 words */
  int dets_on_device;      /* 0: host memory, copied at submit */
  const double* coeffs;    /* n_rows */
  int coeffs_on_device;    /* device inputs must stay valid until the wait returns */
  int64_t n_rows;
  int n_orb;
  int use_hb;
  double eps1;
  uint64_t v_key;          /* 0: V private to this request */
  void* stream;            /* stream that produced device inputs: the server waits for
                              its work up to the submit before reading them */
  int method;              /* DQ_ELOC_STRINGS or DQ_ELOC_WALK */
} dq_server_request;

typedef struct dq_server_result {
  double energy;           /* sum_i c_i sigma_i */
  int status;
  double t_submit, t_admit, t_first_launch, t_done;  /* seconds since the server started */
} dq_server_result;

typedef struct dq_server_stats {
  int64_t submitted, completed, launches, rows;
  int64_t v_hits, v_misses, v_evictions, v_cached;
  int64_t arena_bytes, arena_used, arena_peak;
  int64_t queued, running;
  int64_t vv_builds;       /* lane-ordered H_VV layouts built */
  int64_t vv_rows;         /* rows computed as a product over a layout */
} dq_server_stats;

int dq_server_create(const dq_server_config* cfg, dq_server** out);
/* Arena bytes one request of n_rows takes while it runs (V set, DetMap, string
 * index and its build scratch, psi copy unless on the device, sigma). */
int dq_server_request_bytes(int64_t n_rows, int coeffs_on_device, int64_t* bytes);
/* DQ_ERANGE if the request could never fit in the arena. */
int dq_server_submit(dq_server* s, const dq_server_request* req, int64_t* ticket);
/* Blocks until the request is done (timeout_s < 0: no limit; DQ_ETIMEDOUT
 * otherwise), fills res and releases the ticket (one wait per ticket).
 * DQ_EINVAL if a determinant has orbitals at or above n_orb. */
int dq_server_wait(dq_server* s, int64_t ticket, double timeout_s, dq_server_result* res);
int dq_server_stats_get(dq_server* s, dq_server_stats* out);
/* Finishes every submitted request, then stops the loop and frees the arena. */
int dq_server_destroy(dq_server* s);

/* ---- device operator and energy step ------------------------------------------
 * Replaces the host SciPy products of operator/kernel.py:48-145 and the real
 * energy steps of operator/functional.py:31-268.  Blocks are device COO
 * (H_VV n_v x n_v, H_VP n_v x n_p, any order); H_PV is H_VP^T.  Product
 * entries are sequential column-order sums (SciPy's order). */
typedef struct dq_operator dq_operator;
/* Synchronises; DQ_ERANGE if an index is outside the block shape. */
int dq_operator_create(int64_t n_v, int64_t n_p, const uint32_t* d_vv_row, const uint32_t* d_vv_col,
                       const double* d_vv_val, int64_t nnz_vv, const uint32_t* d_vp_row,
                       const uint32_t* d_vp_col, const double* d_vp_val, int64_t nnz_vp,
                       const double* d_h_pp, void* stream, dq_operator** out);
/* y_V = H_VV psi_V + H_VP psi_P; y_P = H_PV psi_V + diag(H_PP) psi_P (y_P may be NULL).
 * Asynchronous. */
int dq_operator_apply(const dq_operator* op, const double* d_psi_v, const double* d_psi_p,
                      double* d_y_v, double* d_y_p, void* stream);
/* mode 0 variational (n = n_v), 1 proxy, 2 asymmetric (n = n_v + n_p; V then P):
 * psi = sign exp(logabs - max logabs), H psi, E = num / max(den, eps),
 * cot_logabs = 2 psi (H psi - E psi) / max(den, eps) (zero on P in mode 2).
 * d_psi, d_hpsi, d_cot_logabs have n entries.  Synchronises. */
int dq_energy_step(const dq_operator* op, int mode, const double* d_sign, const double* d_logabs,
                   double eps, double* d_psi, double* d_hpsi, double* d_cot_logabs,
                   double* h_energy, void* stream);
int dq_operator_free(dq_operator* op);

/* ---- batched slogdet -------------------------------------------------------
 * Replaces the XLA-FFI targets detnqs_fused_logdet_{f64,f32}_{fwd,bwd}
 * (csrc/ffi/logdet_cuda.cu:500-586, ffi_module.cpp:28-45): A is row-major
 * (batch, n, n), 2 <= n <= 64.  Backward: grad = cot_logabs * A^{-T}
 * (cot_sign is ignored, logdet_cuda.cu:528).  Asynchronous. */
int dq_slogdet_fwd_f64(const double* d_A, int64_t batch, int n, double* d_sign, double* d_logabs,
                       void* stream);
int dq_slogdet_bwd_f64(const double* d_A, const double* d_cot_sign, const double* d_cot_logabs,
                       int64_t batch, int n, double* d_grad, void* stream);
/* The FP32 targets detnqs_fused_logdet_f32_{fwd,bwd} (logdet_cuda.cu:554-586):
 * FP32 in and out, computed in FP64 (the reference computes in FP32; these are
 * at least as accurate).  Asynchronous. */
int dq_slogdet_fwd_f32(const float* d_A, int64_t batch, int n, float* d_sign, float* d_logabs,
                       void* stream);
int dq_slogdet_bwd_f32(const float* d_A, const float* d_cot_sign, const float* d_cot_logabs,
                       int64_t batch, int n, float* d_grad, void* stream);
/* Forward that also keeps the LU factors (row-major, in place, unit-lower L) and
 * the row permutation (perm[k] = original row at position k), and the backward
 * that reuses them instead of refactorising (no counterpart in the reference,
 * whose backward refactorises A: logdet_cuda.cu:293-386; for custom_vjp). */
int dq_slogdet_fwd_lu_f64(const double* d_A, int64_t batch, int n, double* d_sign,
                          double* d_logabs, double* d_lu, int32_t* d_perm, void* stream);
int dq_slogdet_bwd_lu_f64(const double* d_lu, const int32_t* d_perm, const double* d_cot_logabs,
                          int64_t batch, int n, double* d_grad, void* stream);
/* Tile of the 8 < n <= 64 kernels: lanes per matrix, each lane holding every
 * lanes-th row, | 0x100 for the kernel that reads the bounds at run time also
 * at n equal to its bucket's column count (12, 16, 24, 32; without the flag
 * that n takes the kernel with the bounds fixed; both do the same
 * arithmetic).  No counterpart in the reference, whose cuSolverDx launch is
 * fixed per size bucket: logdet_cuda.cu:75-85,402-428.  kind: 0 forward
 * (f64 and f32), 1 forward keeping the LU, 2 backward refactorising A (f64
 * and f32), 3 backward from the LU.  _choices writes the supported codes for
 * n (at most cap) and their count; _set picks one for every later launch of
 * that kind at that n in the process (0 restores the default); _get reads
 * the code in use.  Host-only, thread-safe. */
int dq_slogdet_tile_choices(int n, int* lanes, int cap, int* count);
int dq_slogdet_tile_set(int kind, int n, int lanes);
int dq_slogdet_tile_get(int kind, int n, int* lanes);

/* ---- determinant batch features --------------------------------------------
 * prepare_det_batch (bridge.cpp:347-444; det_batch.cpp:189-319) for d_dets[n][2]
 * against the reference (ref_alpha, ref_beta), block spin-orbital order
 * (beta = n_orb + mo):
 *   d_occ[n][n_alpha + n_beta]  occupied SOs, alpha ascending then beta ascending;
 *   d_k[n]                      excitation rank (capped at 127), DQ_PREP_K;
 *   d_phase[n]                  +-1, DQ_PREP_PHASE;
 *   d_holes / d_parts / d_hp_mask [n][kmax]  holes descending and particles
 *                               ascending (alpha then beta), -1 padded; mask =
 *                               position < k; DQ_PREP_HP (implies DQ_PREP_K);
 *   d_holes_pos / d_parts_pos [n][kmax]  positions in the reference's occupied /
 *                               virtual lists, DQ_PREP_HP_POS (needs DQ_PREP_HP).
 * Errors follow the reference's checks and messages (DQ_EINVAL): the first
 * failing determinant in batch order decides the message.  Synchronises. */
#define DQ_PREP_K 1
#define DQ_PREP_PHASE 2
#define DQ_PREP_HP 4
#define DQ_PREP_HP_POS 8
int dq_prepare_det_batch(const uint64_t* d_dets, int64_t n, uint64_t ref_alpha, uint64_t ref_beta,
                         int n_orb, int n_alpha, int n_beta, int kmax, int flags, int32_t* d_occ,
                         int8_t* d_k, int8_t* d_phase, int32_t* d_holes, int32_t* d_parts,
                         uint8_t* d_hp_mask, int32_t* d_holes_pos, int32_t* d_parts_pos,
                         void* stream);

/* ---- full-CI enumeration -----------------------------------------------------
 * gen_fci_dets (bridge.cpp:253-260; FCISpace, det_enum.cpp:87-107): beta words
 * in the outer loop, each spin's words in ascending numeric order.
 * dq_fci_size gives the count; dq_gen_fci_dets writes d_out[n][2] (n must be
 * that count).  Asynchronous. */
int dq_fci_size(int n_orb, int n_alpha, int n_beta, int64_t* h_n);
int dq_gen_fci_dets(int n_orb, int n_alpha, int n_beta, uint64_t* d_out, int64_t n, void* stream);

/* ---- connected and perturbative spaces, EN-PT2 -------------------------------
 * Results are device-resident sorted determinant sets (ascending (alpha, beta)),
 * optionally with one FP64 value per determinant.  chunk_items bounds the
 * candidates generated and sorted per pass (0: 2^28); it changes memory use,
 * not results (except the order in which PT2 partial sums of several passes
 * are added).  All synchronise `stream`. */
typedef struct dq_detset dq_detset;
#define DQ_SCREEN_NONE 0
#define DQ_SCREEN_STATIC 1
#define DQ_SCREEN_DYNAMIC 2
/* gen_connected_dets (bridge.cpp:262-271; det_space.cpp:219-227, 94-98): every
 * single and double excitation of d_dets, sorted unique. */
int dq_gen_connected(const uint64_t* d_dets, int64_t n, int n_orb, int64_t chunk_items, void* stream,
                     dq_detset** out);
/* gen_perturbative_dets (bridge.cpp:278-338): P = C \ V sorted, with C from
 * DQ_SCREEN_NONE (det_space.cpp:233-255; ctx may be NULL), DQ_SCREEN_STATIC or
 * DQ_SCREEN_DYNAMIC (ham_utils.cpp:546-701; heat-bath table required, dynamic
 * needs d_psi[n]).  V must hold no duplicates (DetMap::from_ordered strict). */
int dq_gen_perturbative(const dq_ctx* ctx, const uint64_t* d_dets, int64_t n, int n_orb,
                        const double* d_psi, int mode, double eps1, int64_t chunk_items,
                        void* stream, dq_detset** out);
/* compute_pt2 (bridge.cpp:721-771; local_conn.cpp:247-418): the internal and
 * external EN-PT2 terms and the number of external determinants; ext_out (may
 * be NULL) receives the external set with v_a = sum_i H_ai c_i. */
int dq_compute_pt2(const dq_ctx* ctx, const uint64_t* d_dets, const double* d_coeffs, int64_t n,
                   int n_orb, double e_ref, int use_hb, double eps1, int64_t chunk_items,
                   double* h_internal, double* h_external, int64_t* h_n_ext, void* stream,
                   dq_detset** ext_out);
int dq_detset_size(const dq_detset* set, int64_t* n, int* has_values);
/* Copies the set (dets and/or values; NULL skips) to device (1) or host (0). */
int dq_detset_copy(const dq_detset* set, uint64_t* dets, double* vals, int dst_on_device,
                   void* stream);
int dq_detset_free(dq_detset* set);

#ifdef __cplusplus
}
#endif
#endif /* DETNQS_HIP_H */
