// Shared definitions for the OpenSC MI355X (gfx950) HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/opensc_hip.h"

namespace osc {

// ---------------------------------------------------------------------------
// complex helpers (interleaved re/im, same layout as Fortran complex)
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b) {
  cpx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) {
  return mk(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conjg(cpx<T> a) {
  return mk(a.x, -a.y);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) {
  return mk(a.x * s, a.y * s);
}
// multiply by +i (DIR=+1) or -i (DIR=-1)
template <int DIR, typename T>
__host__ __device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
  return DIR > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// status / errors
// ---------------------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

#define OSC_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::osc::Error{OSC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define OSC_REQUIRE(cond, code, msg)                 \
  do {                                               \
    if (!(cond)) throw ::osc::Error{(code), (msg)}; \
  } while (0)

// ---------------------------------------------------------------------------
// device memory: a size-bucketed cache so that repeated solves of the same
// shape allocate nothing after the first call
// ---------------------------------------------------------------------------
class DevPool {
 public:
  void* get(size_t bytes);
  void put(void* p);
  void release_free();
  void release_all();
  size_t bytes_cached() const { return cached_; }
  size_t bytes_live() const { return live_bytes_; }
  ~DevPool() { release_all(); }

 private:
  std::multimap<size_t, void*> free_;  // by size: best fit
  std::unordered_map<void*, size_t> live_;
  size_t cached_ = 0, live_bytes_ = 0;
};

// RAII handle on a pooled device buffer
struct DevBuf {
  DevPool* pool = nullptr;
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(DevPool& pl, size_t bytes) : pool(&pl), p(bytes ? pl.get(bytes) : nullptr) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : pool(o.pool), p(o.p) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    reset();
    pool = o.pool;
    p = o.p;
    o.p = nullptr;
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p && pool) pool->put(p);
    p = nullptr;
  }
  template <typename U>
  U* as() const {
    return static_cast<U*>(p);
  }
};

}  // namespace osc
