// Determinant batch features for the neural ansatz and full-CI enumeration,
// on gfx950.
//
//  * prepare_det_batch (det_batch.cpp:189-319): per determinant the occupied
//    spin orbitals (alpha ascending, then beta ascending + n_orb), the
//    excitation rank k against a reference, the fermionic phase of reaching it
//    (holes annihilated in descending order, particles created in ascending
//    order), and the hole / particle lists (alpha then beta; holes descending,
//    particles ascending) with their positions in the reference's occupied /
//    virtual lists.  Note the block spin-orbital order (beta = n_orb + mo) of
//    this layer, not the interleaved order of the Hamiltonian.
//  * FCISpace (det_enum.cpp:87-107): every (alpha, beta) pair, beta words in
//    the outer loop, each spin's words in ascending numeric order (Gosper's
//    order, det_enum.cpp:44-66).
//
// Both are HBM-bound integer work.  prepare_det_batch works on tiles of 256
// determinants per block: each thread builds its determinant's rows once, as
// bytes in LDS laid out like the output (prep_staged_kernel), and the block
// then writes every array of the tile with consecutive 16-byte stores, so the
// variable-width rows leave as full coalesced lines.  When the rows do not
// fit LDS (n_alpha + n_beta + 5 kmax > 240) prep_kernel computes each output
// element on its own (element e of the tile -> row e / width, the j-th set bit
// by six popcounts).
#include <algorithm>

#include "dq_common.hpp"

namespace dq {

namespace {

constexpr int kTile = 256;

// Index of the j-th set bit (j >= 0, counted from bit 0) of w; w must have
// more than j set bits.
__device__ __forceinline__ int select_bit(u64 w, int j) {
  int pos = 0;
#pragma unroll
  for (int width = 32; width >= 1; width >>= 1) {
    const u64 lo = w & ((1ULL << width) - 1);
    const int c = popc(lo);
    if (j >= c) {
      j -= c;
      w >>= width;
      pos += width;
    } else {
      w = lo;
    }
  }
  return pos;
}

// phase_one_spin (det_batch.cpp:61-97): sign of annihilating the holes of
// `det` in `ref` from the highest down, then creating its particles from the
// lowest up, each operator passing the occupied orbitals below it.  Inputs are
// already restricted to the first n_orb orbitals and have equal hole and
// particle counts.
__device__ __forceinline__ int phase_one_spin(u64 ref, u64 det) {
  u64 holes = ref & ~det, parts = det & ~ref, occ = ref;
  int n = 0;
  while (holes) {
    const int h = 63 - __clzll(static_cast<long long>(holes));
    n += popc(occ & low_mask(h));
    occ ^= 1ULL << h;
    holes ^= 1ULL << h;
  }
  while (parts) {
    const int p = ctz(parts);
    n += popc(occ & low_mask(p));
    occ ^= 1ULL << p;
    parts &= parts - 1;
  }
  return (n & 1) ? -1 : 1;
}

struct PrepArgs {
  const u64* dets;
  int64_t n;
  u64 ra, rb;  // reference words, restricted to n_orb orbitals
  int n_orb, na, nb, kmax, flags;
  int32_t* occ;
  int8_t* k;
  int8_t* phase;
  int32_t* holes;
  int32_t* parts;
  uint8_t* mask;
  int32_t* hpos;
  int32_t* ppos;
  unsigned long long* err;  // min over failing rows of (row << 2 | code)
};

// Position of a reference-occupied spin orbital in the reference's occupied
// list (alpha ascending, then beta ascending), and of a reference-virtual one
// in the complement of that list over [0, 2 n_orb) (build_ref_maps,
// det_batch.cpp:105-143).
__device__ __forceinline__ int occ_pos(const PrepArgs& A, int so) {
  return so < A.n_orb ? popc(A.ra & low_mask(so)) : A.na + popc(A.rb & low_mask(so - A.n_orb));
}
__device__ __forceinline__ int virt_pos(const PrepArgs& A, int so) {
  if (so < A.n_orb) return so - popc(A.ra & low_mask(so));
  const int s = so - A.n_orb;
  return (A.n_orb - A.na) + s - popc(A.rb & low_mask(s));
}

__global__ __launch_bounds__(kTile) void prep_kernel(PrepArgs A) {
  __shared__ u64 sa[kTile], sb[kTile];
  const int t = threadIdx.x;
  const u64 m = low_mask(A.n_orb);
  const int ne = A.na + A.nb;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * kTile; base < A.n;
       base += static_cast<int64_t>(gridDim.x) * kTile) {
    const int rows = static_cast<int>(A.n - base < kTile ? A.n - base : kTile);
    __syncthreads();  // the previous tile's readers are done
    if (t < rows) {
      sa[t] = A.dets[2 * (base + t)];
      sb[t] = A.dets[2 * (base + t) + 1];
    }
    __syncthreads();

    // per determinant: checks (in the reference's order), k, phase
    if (t < rows) {
      const int64_t i = base + t;
      const u64 a = sa[t], b = sb[t];
      int code = 0;
      if ((a | b) & ~m)
        code = 1;  // det bits outside n_orb range
      else if (popc(a) != A.na || popc(b) != A.nb)
        code = 2;  // electron count mismatch
      const u64 ha = A.ra & ~a, hb = A.rb & ~b;
      const int kk = popc(ha) + popc(hb);
      if (A.flags & DQ_PREP_K) A.k[i] = static_cast<int8_t>(kk > 127 ? 127 : kk);
      if (code == 0 && (A.flags & DQ_PREP_PHASE)) {
        const u64 pa = a & ~A.ra, pb = b & ~A.rb;
        if (popc(ha) != popc(pa) || popc(hb) != popc(pb))
          code = 3;  // holes != parts in spin channel
        else
          A.phase[i] = static_cast<int8_t>(phase_one_spin(A.ra, a) * phase_one_spin(A.rb, b));
      }
      if (code) atomicMin(A.err, (static_cast<unsigned long long>(i) << 2) | static_cast<unsigned>(code));
    }

    // occupied spin orbitals: width ne
    if (ne > 0) {
      int32_t* out = A.occ + base * ne;
      const int total = rows * ne;
      for (int e = t; e < total; e += kTile) {
        const int r = e / ne, j = e - r * ne;
        out[e] = j < A.na ? select_bit(sa[r] & m, j) : select_bit(sb[r] & m, j - A.na) + A.n_orb;
      }
    }

    // hole / particle lists: width kmax
    if ((A.flags & DQ_PREP_HP) && A.kmax > 0) {
      const int K = A.kmax;
      const int64_t o0 = base * K;
      const int total = rows * K;
      for (int e = t; e < total; e += kTile) {
        const int r = e / K, j = e - r * K;
        const u64 a = sa[r] & m, b = sb[r] & m;
        const u64 ha = A.ra & ~a, hb = A.rb & ~b, pa = a & ~A.ra, pb = b & ~A.rb;
        const int nha = popc(ha), nhb = popc(hb), npa = popc(pa), npb = popc(pb);
        int h = -1, p = -1;
        if (j < nha)
          h = select_bit(ha, nha - 1 - j);
        else if (j - nha < nhb)
          h = select_bit(hb, nhb - 1 - (j - nha)) + A.n_orb;
        if (j < npa)
          p = select_bit(pa, j);
        else if (j - npa < npb)
          p = select_bit(pb, j - npa) + A.n_orb;
        A.holes[o0 + e] = h;
        A.parts[o0 + e] = p;
        A.mask[o0 + e] = j < nha + nhb ? 1 : 0;
        if (A.flags & DQ_PREP_HP_POS) {
          A.hpos[o0 + e] = h < 0 ? -1 : occ_pos(A, h);
          A.ppos[o0 + e] = p < 0 ? -1 : virt_pos(A, p);
        }
      }
    }
  }
}

// The same outputs, staged: each thread builds its determinant's lists once
// (bit scans) as bytes in LDS laid out like the output rows (spin orbitals and
// positions are < 128, -1 pads), then the block copies every array of the tile
// out with 16-byte stores (VEC) or element stores.  LDS: kTile (ne + 5 kmax)
// bytes.  Output pointers of the VEC form are 16-byte (int32 arrays) and
// 4-byte (hp_mask) aligned.
template <bool VEC>
__device__ __forceinline__ void copy_i8_i32(const int8_t* __restrict__ src, int32_t* __restrict__ dst, int count) {
  const int t = threadIdx.x;
  int done = 0;
  if (VEC) {
    const int q4 = count >> 2;
    const uint32_t* s4 = reinterpret_cast<const uint32_t*>(src);
    int4* d4 = reinterpret_cast<int4*>(dst);
    for (int q = t; q < q4; q += kTile) {
      const uint32_t w = s4[q];
      d4[q] = make_int4(static_cast<int8_t>(w & 0xFF), static_cast<int8_t>((w >> 8) & 0xFF),
                        static_cast<int8_t>((w >> 16) & 0xFF), static_cast<int8_t>(w >> 24));
    }
    done = q4 << 2;
  }
  for (int e = done + t; e < count; e += kTile) dst[e] = src[e];
}

template <bool VEC>
__device__ __forceinline__ void copy_u8(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int count) {
  const int t = threadIdx.x;
  int done = 0;
  if (VEC) {
    const int q4 = count >> 2;
    for (int q = t; q < q4; q += kTile)
      reinterpret_cast<uint32_t*>(dst)[q] = reinterpret_cast<const uint32_t*>(src)[q];
    done = q4 << 2;
  }
  for (int e = done + t; e < count; e += kTile) dst[e] = src[e];
}

template <bool VEC>
__global__ __launch_bounds__(kTile) void prep_staged_kernel(PrepArgs A) {
  extern __shared__ unsigned char lds[];
  const int t = threadIdx.x;
  const u64 m = low_mask(A.n_orb);
  const int ne = A.na + A.nb;
  const bool hp = (A.flags & DQ_PREP_HP) && A.kmax > 0, pos = A.flags & DQ_PREP_HP_POS;
  const int K = hp ? A.kmax : 0;
  int8_t* Locc = reinterpret_cast<int8_t*>(lds);
  int8_t* Lh = Locc + kTile * ne;
  int8_t* Lp = Lh + kTile * K;
  uint8_t* Lm = reinterpret_cast<uint8_t*>(Lp + kTile * K);
  int8_t* Lhp = reinterpret_cast<int8_t*>(Lm + kTile * K);
  int8_t* Lpp = Lhp + kTile * K;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * kTile; base < A.n;
       base += static_cast<int64_t>(gridDim.x) * kTile) {
    const int rows = static_cast<int>(A.n - base < kTile ? A.n - base : kTile);
    __syncthreads();  // the previous tile's copy-out is done
    if (t < rows) {
      const int64_t i = base + t;
      const u64 a = A.dets[2 * i], b = A.dets[2 * i + 1];
      int code = 0;
      if ((a | b) & ~m)
        code = 1;  // det bits outside n_orb range
      else if (popc(a) != A.na || popc(b) != A.nb)
        code = 2;  // electron count mismatch
      const u64 am = a & m, bm = b & m;
      const u64 ha = A.ra & ~am, hb = A.rb & ~bm, pa = am & ~A.ra, pb = bm & ~A.rb;
      const int kk = popc(ha) + popc(hb);
      if (A.flags & DQ_PREP_K) A.k[i] = static_cast<int8_t>(kk > 127 ? 127 : kk);
      if (code == 0 && (A.flags & DQ_PREP_PHASE)) {
        if (popc(ha) != popc(pa) || popc(hb) != popc(pb))
          code = 3;  // holes != parts in spin channel
        else
          A.phase[i] = static_cast<int8_t>(phase_one_spin(A.ra, am) * phase_one_spin(A.rb, bm));
      }
      if (code) atomicMin(A.err, (static_cast<unsigned long long>(i) << 2) | static_cast<unsigned>(code));
      // occupied list (a row with the wrong counts fails the call; its row
      // stays in bounds)
      int8_t* o = Locc + t * ne;
      int j = 0;
      for (u64 w = am; w && j < ne; w &= w - 1) o[j++] = static_cast<int8_t>(ctz(w));
      for (u64 w = bm; w && j < ne; w &= w - 1) o[j++] = static_cast<int8_t>(ctz(w) + A.n_orb);
      if (hp) {
        int8_t* h = Lh + t * K;
        int8_t* hpos = Lhp + t * K;
        int c = 0;
        for (u64 w = ha; w && c < K; ++c) {
          const int x = 63 - __clzll(static_cast<long long>(w));
          w ^= 1ULL << x;
          h[c] = static_cast<int8_t>(x);
          hpos[c] = static_cast<int8_t>(popc(A.ra & low_mask(x)));
        }
        for (u64 w = hb; w && c < K; ++c) {
          const int x = 63 - __clzll(static_cast<long long>(w));
          w ^= 1ULL << x;
          h[c] = static_cast<int8_t>(x + A.n_orb);
          hpos[c] = static_cast<int8_t>(A.na + popc(A.rb & low_mask(x)));
        }
        for (; c < K; ++c) h[c] = hpos[c] = -1;
        int8_t* p = Lp + t * K;
        int8_t* ppos = Lpp + t * K;
        c = 0;
        for (u64 w = pa; w && c < K; w &= w - 1, ++c) {
          const int x = ctz(w);
          p[c] = static_cast<int8_t>(x);
          ppos[c] = static_cast<int8_t>(x - popc(A.ra & low_mask(x)));
        }
        for (u64 w = pb; w && c < K; w &= w - 1, ++c) {
          const int x = ctz(w);
          p[c] = static_cast<int8_t>(x + A.n_orb);
          ppos[c] = static_cast<int8_t>((A.n_orb - A.na) + x - popc(A.rb & low_mask(x)));
        }
        for (; c < K; ++c) p[c] = ppos[c] = -1;
        uint8_t* mk = Lm + t * K;
        for (int q = 0; q < K; ++q) mk[q] = q < kk ? 1 : 0;
      }
    }
    __syncthreads();
    if (ne > 0) copy_i8_i32<VEC>(Locc, A.occ + base * ne, rows * ne);
    if (hp) {
      const int64_t o0 = base * K;
      copy_i8_i32<VEC>(Lh, A.holes + o0, rows * K);
      copy_i8_i32<VEC>(Lp, A.parts + o0, rows * K);
      copy_u8<VEC>(Lm, A.mask + o0, rows * K);
      if (pos) {
        copy_i8_i32<VEC>(Lhp, A.hpos + o0, rows * K);
        copy_i8_i32<VEC>(Lpp, A.ppos + o0, rows * K);
      }
    }
  }
}

__global__ void set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// One device word per calling thread and device for the first-failure report
// of a synchronous call, allocated on first use and kept for the life of the
// process (calls of one thread never overlap: each synchronises before it
// returns).
unsigned long long* error_word() {
  constexpr int kMaxDev = 64;
  thread_local unsigned long long* words[kMaxDev] = {};
  int dev = 0;
  DQ_HIP(hipGetDevice(&dev));
  DQ_REQUIRE(dev >= 0 && dev < kMaxDev, DQ_ERUNTIME, "device index out of range");
  if (!words[dev]) DQ_HIP(hipMalloc(reinterpret_cast<void**>(&words[dev]), sizeof(unsigned long long)));
  return words[dev];
}

// ------------------------------------------------------------------ FCI space
struct Binom {
  u64 v[65][65];
  constexpr Binom() : v() {
    for (int n = 0; n <= 64; ++n) {
      v[n][0] = 1;
      for (int k = 1; k <= n; ++k) v[n][k] = v[n - 1][k - 1] + v[n - 1][k];
    }
  }
};
__constant__ Binom kBinom;  // C(n, k) for n, k <= 64 (C(64, 32) < 2^61)

// The r-th smallest n-bit word with k bits set: scanning from the top bit, a
// clear bit p leaves C(p, k) words below it.
__device__ __forceinline__ u64 unrank(u64 r, int n, int k) {
  u64 w = 0;
  for (int p = n - 1; p >= 0 && k > 0; --p) {
    const u64 c = kBinom.v[p][k];
    if (r >= c) {
      w |= 1ULL << p;
      r -= c;
      --k;
    }
  }
  return w;
}

__global__ void fci_kernel(int n_orb, int na, int nb, u64 n_alpha_words, int64_t n, u64* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 bi = static_cast<u64>(i) / n_alpha_words, ai = static_cast<u64>(i) - bi * n_alpha_words;
  const u64 a = unrank(ai, n_orb, na), b = unrank(bi, n_orb, nb);
  reinterpret_cast<ulonglong2*>(out)[i] = make_ulonglong2(a, b);
}

u64 binom_host(int n, int k) {
  static const Binom table;
  return table.v[n][k];
}

void fci_check(int n_orb, int na, int nb) {
  if (n_orb <= 0 || n_orb > 64)
    throw Error(DQ_EINVAL, "FCISpace: n_orb=" + std::to_string(n_orb) + " out of range [1,64]");
  if (na < 0 || na > n_orb || nb < 0 || nb > n_orb)
    throw Error(DQ_EINVAL, "FCISpace: invalid electrons n_α=" + std::to_string(na) + ", n_β=" + std::to_string(nb) +
                               " for " + std::to_string(n_orb) + " orbitals");
}

}  // namespace
}  // namespace dq

extern "C" {

int dq_prepare_det_batch(const uint64_t* d_dets, int64_t n, uint64_t ref_alpha, uint64_t ref_beta, int n_orb,
                         int n_alpha, int n_beta, int kmax, int flags, int32_t* d_occ, int8_t* d_k, int8_t* d_phase,
                         int32_t* d_holes, int32_t* d_parts, uint8_t* d_hp_mask, int32_t* d_holes_pos,
                         int32_t* d_parts_pos, void* stream) {
  return dq::guarded([&] {
    using dq::u64;
    // argument checks in the order of det_batch.cpp:206-226
    DQ_REQUIRE(n_orb >= 0 && n_orb <= 64, DQ_EINVAL, "n_orb out of range [0,64]");
    DQ_REQUIRE(n >= 0, DQ_EINVAL, "negative batch size");
    DQ_REQUIRE(n_alpha >= 0 && n_beta >= 0, DQ_EINVAL, "invalid electron counts");
    const bool need_hp = flags & DQ_PREP_HP, need_pos = flags & DQ_PREP_HP_POS;
    if (need_hp) flags |= DQ_PREP_K;  // bridge.cpp:369: need_k = need_k || need_hp
    DQ_REQUIRE(n == 0 || (d_dets && (n_alpha + n_beta == 0 || d_occ)), DQ_EINVAL, "null pointer in required outputs");
    DQ_REQUIRE(!need_pos || need_hp, DQ_EINVAL, "need_hp_pos requires need_hp");
    DQ_REQUIRE(!need_hp || kmax >= 0, DQ_EINVAL, "kmax must be non-negative");
    const u64 m = dq::low_mask(n_orb);
    const u64 ra = ref_alpha & m, rb = ref_beta & m;
    if (need_pos)
      DQ_REQUIRE(__builtin_popcountll(ra) == n_alpha && __builtin_popcountll(rb) == n_beta, DQ_EINVAL,
                 "reference electron count mismatch");
    if (n == 0) return;
    DQ_REQUIRE(!(flags & DQ_PREP_K) || d_k, DQ_EINVAL, "k_out is null");
    DQ_REQUIRE(!(flags & DQ_PREP_PHASE) || d_phase, DQ_EINVAL, "phase_out is null");
    DQ_REQUIRE(!need_hp || kmax == 0 || (d_holes && d_parts && d_hp_mask), DQ_EINVAL, "hp outputs are null");
    DQ_REQUIRE(!need_pos || kmax == 0 || (d_holes_pos && d_parts_pos), DQ_EINVAL, "pos outputs are null");
    DQ_REQUIRE(static_cast<int64_t>(dq::kTile) * std::max(n_alpha + n_beta, kmax) < (1LL << 31), DQ_EINVAL,
               "row width too large");
    hipStream_t s = dq::as_stream(stream);
    unsigned long long* err = dq::error_word();
    dq::set_u64<<<1, 1, 0, s>>>(err, ~0ULL);
    dq::PrepArgs A{d_dets,  n,       ra,   rb,      n_orb,   n_alpha,     n_beta,      kmax,  flags,
                   d_occ,   d_k,     d_phase, d_holes, d_parts, d_hp_mask, d_holes_pos, d_parts_pos, err};
    const int64_t tiles = (n + dq::kTile - 1) / dq::kTile;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(tiles, 8192));
    const int K = need_hp ? kmax : 0;
    const size_t lds = static_cast<size_t>(dq::kTile) * (n_alpha + n_beta + 5 * K);
    if (lds <= 60 * 1024) {
      // 16-byte stores need 16-byte aligned int32 arrays and a 4-byte aligned mask
      auto al = [](const void* p, uintptr_t a) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) % a) == 0; };
      const bool vec = al(d_occ, 16) && al(d_holes, 16) && al(d_parts, 16) && al(d_holes_pos, 16) &&
                       al(d_parts_pos, 16) && al(d_hp_mask, 4);
      if (vec)
        dq::prep_staged_kernel<true><<<grid, dq::kTile, lds, s>>>(A);
      else
        dq::prep_staged_kernel<false><<<grid, dq::kTile, lds, s>>>(A);
    } else {
      dq::prep_kernel<<<grid, dq::kTile, 0, s>>>(A);
    }
    DQ_HIP(hipGetLastError());
    unsigned long long e = 0;
    DQ_HIP(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    if (e != ~0ULL) {
      const int code = static_cast<int>(e & 3);
      throw dq::Error(DQ_EINVAL, code == 1   ? "det bits outside n_orb range"
                                 : code == 2 ? "electron count mismatch"
                                             : "holes != parts in spin channel");
    }
  });
}

int dq_fci_size(int n_orb, int n_alpha, int n_beta, int64_t* h_n) {
  return dq::guarded([&] {
    dq::fci_check(n_orb, n_alpha, n_beta);
    const dq::u64 ca = dq::binom_host(n_orb, n_alpha), cb = dq::binom_host(n_orb, n_beta);
    DQ_REQUIRE(cb == 0 || ca <= static_cast<dq::u64>(INT64_MAX) / cb, DQ_EINVAL, "FCISpace: too many determinants");
    *h_n = static_cast<int64_t>(ca * cb);
  });
}

int dq_gen_fci_dets(int n_orb, int n_alpha, int n_beta, uint64_t* d_out, int64_t n, void* stream) {
  return dq::guarded([&] {
    dq::fci_check(n_orb, n_alpha, n_beta);
    const dq::u64 ca = dq::binom_host(n_orb, n_alpha), cb = dq::binom_host(n_orb, n_beta);
    DQ_REQUIRE(cb == 0 || ca <= static_cast<dq::u64>(INT64_MAX) / cb, DQ_EINVAL, "FCISpace: too many determinants");
    DQ_REQUIRE(n == static_cast<int64_t>(ca * cb), DQ_EINVAL, "output size does not match the FCI space");
    if (n == 0) return;
    DQ_REQUIRE(d_out, DQ_EINVAL, "null output");
    hipStream_t s = dq::as_stream(stream);
    dq::fci_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(n_orb, n_alpha, n_beta, ca, n, d_out);
    DQ_HIP(hipGetLastError());
  });
}

}  // extern "C"
