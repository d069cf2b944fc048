// Shared host/device helpers of libdetnqs_hip.so (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "detnqs_hip.h"

namespace dq {

using u64 = std::uint64_t;
using u32 = std::uint32_t;

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_error(const std::string& msg);

#define DQ_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw ::dq::Error(DQ_ERUNTIME, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define DQ_REQUIRE(cond, code, msg)                  \
  do {                                               \
    if (!(cond)) throw ::dq::Error((code), (msg));   \
  } while (0)

// Runs f(); converts exceptions to status codes + last-error text.
template <class F>
int guarded(F&& f) {
  try {
    f();
    return DQ_OK;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("out of host memory");
    return DQ_ERUNTIME;
  } catch (const std::exception& e) {
    set_error(e.what());
    return DQ_ERUNTIME;
  }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Sets `words` 32-bit words at p to v with a kernel on stream s.  Used instead
// of hipMemsetAsync where a caller may capture the work into a HIP graph.
void fill_u32(void* p, size_t words, u32 v, hipStream_t s);

// Block cache for large device temporaries and results, per device.  A block
// released on stream s is kept with an event recorded on s; an allocation that
// reuses it makes its own stream wait for that event (s == nullptr: the host
// waits), so a block is never handed out while work queued before its release
// may still touch it.  Best fit among cached blocks of this device within 2x
// the request; otherwise hipMalloc (after returning the device's cached blocks
// if that fails).  cache_release_idle() takes a block no queued work uses.
// dq_trim_memory() returns every cached block to the driver.
void* cache_alloc(size_t bytes, hipStream_t s);
void cache_release(void* p, size_t bytes, hipStream_t s);
void cache_release_idle(void* p, size_t bytes);
void cache_trim();

// Buffers of at least this size go through the block cache.
constexpr size_t kCachedBytes = size_t(8) << 20;

// Device buffer owned by the library (RAII), in one of two modes.
//
// Stream-ordered (constructed or first allocated with a stream): for
// temporaries of a call.  Every size comes from the block cache and release is
// ordered after the work queued on that stream, so the buffer may go out of
// scope while kernels that use it are still queued, and no release waits for
// the device.  The stream must outlive the buffer.
//
// Unordered (no stream): for results and context tables that outlive the
// call's stream.  Releasing one is a hipFree: a small buffer is freed, a large
// one is kept in the block cache after the device has drained (as hipFree
// would wait).  Only the free functions of results and contexts do this.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t s = nullptr;
  bool ordered = false;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(size_t count, hipStream_t st) { alloc(count, st); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), s(o.s), ordered(o.ordered) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      s = o.s;
      ordered = o.ordered;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  // Allocates in the buffer's current mode.
  void alloc(size_t count) {
    release();
    if (count) {
      if (ordered)
        p = static_cast<T*>(cache_alloc(count * sizeof(T), s));
      else if (count * sizeof(T) >= kCachedBytes)
        p = static_cast<T*>(cache_alloc(count * sizeof(T), nullptr));
      else
        DQ_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    }
    n = count;
  }
  // Makes the buffer stream-ordered on `st`, then allocates.
  void alloc(size_t count, hipStream_t st) {
    release();
    s = st;
    ordered = true;
    alloc(count);
  }
  void release() {
    if (p) {
      if (ordered) {
        cache_release(p, n * sizeof(T), s);
      } else if (n * sizeof(T) >= kCachedBytes) {
        if (hipDeviceSynchronize() == hipSuccess)
          cache_release_idle(p, n * sizeof(T));
        else
          (void)hipFree(p);
      } else {
        (void)hipFree(p);
      }
    }
    p = nullptr;
    n = 0;
  }
};

// Device buffer from the block cache: for large temporaries and results.
// Release is ordered after the work already queued on `s`, so a buffer may go
// out of scope while kernels that use it are still queued.
template <class T>
struct PoolBuf {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t s = nullptr;
  PoolBuf() = default;
  PoolBuf(size_t count, hipStream_t st) { alloc(count, st); }
  PoolBuf(const PoolBuf&) = delete;
  PoolBuf& operator=(const PoolBuf&) = delete;
  PoolBuf(PoolBuf&& o) noexcept : p(o.p), n(o.n), s(o.s) {
    o.p = nullptr;
    o.n = 0;
  }
  PoolBuf& operator=(PoolBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      s = o.s;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~PoolBuf() { release(); }
  void alloc(size_t count, hipStream_t st) {
    release();
    s = st;
    if (count) p = static_cast<T*>(cache_alloc(count * sizeof(T), st));
    n = count;
  }
  void release() {
    if (p) cache_release(p, n * sizeof(T), s);
    p = nullptr;
    n = 0;
  }
};

// ------------------------------------------------------------------ bits (device)
__device__ __forceinline__ int popc(u64 x) { return __popcll(x); }
__device__ __forceinline__ int ctz(u64 x) { return __ffsll(static_cast<long long>(x)) - 1; }
__host__ __device__ __forceinline__ u64 low_mask(int n) {
  return n <= 0 ? 0ULL : (n >= 64 ? ~0ULL : ((1ULL << n) - 1));
}
// popcount of occ strictly between positions x and y (det_ops.cpp:85-93), for
// x != y, both in [0, 63]: bits lo+1 .. hi-1 are 2^hi - 2^(lo+1).
__device__ __forceinline__ int between(u64 occ, int x, int y) {
  const int lo = x < y ? x : y, hi = x < y ? y : x;
  return popc(occ & ((1ULL << hi) - (2ULL << lo)));
}
// Index of the k-th set bit (k >= 0) of w.
__device__ __forceinline__ int nth_bit(u64 w, int k) {
  for (int i = 0; i < k; ++i) w &= w - 1;
  return ctz(w);
}

// Scemama–Giner parity of one spin channel for a same-spin double excitation
// where orbitals h0<h1 are removed from `bra` and p0<p1 are added
// (det_ops.cpp:101-127 with holes = added, particles = removed).
__device__ __forceinline__ int parity2(u64 bra, int h0, int h1, int p0, int p1) {
  int n = between(bra, p0, h0) + between(bra, p1, h1);
  int a = p0 < h0 ? p0 : h0, b = p0 < h0 ? h0 : p0;
  int c = p1 < h1 ? p1 : h1, d = p1 < h1 ? h1 : p1;
  if ((c > a && c < b && d > b) || (a > c && a < d && b > d)) ++n;
  return n;
}

// ------------------------------------------------------------------ packed keys
// integral_mo.hpp:30-49.
__host__ __device__ __forceinline__ size_t key1(int p, int q) {
  int hi = p > q ? p : q, lo = p > q ? q : p;
  return static_cast<size_t>(hi) * (hi + 1) / 2 + lo;
}
__host__ __device__ __forceinline__ size_t key2(int p, int q, int r, int s) {
  size_t x = key1(p, q), y = key1(r, s);
  if (x < y) {
    size_t t = x;
    x = y;
    y = t;
  }
  return x * (x + 1) / 2 + y;
}

constexpr double kMatThr = 1e-15;  // MAT_ELEMENT_THRESH, ham_utils.hpp:39

}  // namespace dq
