// Shared definitions for the OpenSC MI355X (gfx950) HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../include/opensc_hip.h"

namespace osc {

// ---------------------------------------------------------------------------
// kernel launches, recordable into a HIP graph
// ---------------------------------------------------------------------------
// While a host thread records a solve (t_recorder set), launch() adds each
// kernel as a node of the recorder's graph, after the previous one (the
// stream order of the eager launches), instead of launching it.  Explicit
// nodes need no stream capture, which would make other threads' legacy-stream
// copies fail while it lasts.
struct GraphRecorder {
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  bool failed = false;
};
extern thread_local GraphRecorder* t_recorder;

template <typename... KArgs, typename... Args>
inline void launch(void (*k)(KArgs...), dim3 grid, dim3 block, size_t shm, hipStream_t s,
                   Args&&... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  // the arguments converted to the kernel's parameter types, one address each
  std::tuple<std::decay_t<KArgs>...> vals(std::forward<Args>(args)...);
  void* ptrs[sizeof...(KArgs) > 0 ? sizeof...(KArgs) : 1] = {};
  std::apply(
      [&](auto&... v) {
        size_t i = 0;
        ((ptrs[i++] = static_cast<void*>(&v)), ...);
      },
      vals);
  if (GraphRecorder* r = t_recorder) {
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void*>(k);
    kp.gridDim = grid;
    kp.blockDim = block;
    kp.sharedMemBytes = unsigned(shm);
    kp.kernelParams = ptrs;
    kp.extra = nullptr;
    hipGraphNode_t node = nullptr;
    if (hipGraphAddKernelNode(&node, r->graph, r->last ? &r->last : nullptr, r->last ? 1 : 0,
                              &kp) != hipSuccess) {
      (void)hipGetLastError();
      r->failed = true;
      return;
    }
    r->last = node;
    return;
  }
  (void)hipLaunchKernel(reinterpret_cast<const void*>(k), grid, block, ptrs, shm, s);
}

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
// Complex products (FFT twiddles, spectral products) may contract into FMAs:
// one rounding fewer per part, against the -ffp-contract=off that keeps the
// Green and index arithmetic in upstream's operation order.
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b) {
#pragma clang fp contract(fast)
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
  // frees cached buffers, largest first, until at most `target` bytes are
  // cached or live; returns the bytes freed
  size_t trim(size_t target);
  size_t bytes_cached() const { return cached_; }
  size_t bytes_live() const { return live_bytes_; }
  // device allocations made, and the host time they took (ns)
  uint64_t allocs() const { return allocs_; }
  uint64_t alloc_ns() const { return alloc_ns_; }
  // while set, buffers given back are moved here instead of the free list
  // (the buffers a captured HIP graph keeps using), and get() takes from
  // here first; adopt() returns one to the free list
  std::vector<std::pair<void*, size_t>>* retain = nullptr;
  void adopt(void* p, size_t bytes);
  ~DevPool() { release_all(); }

 private:
  std::multimap<size_t, void*> free_;  // by size: best fit
  std::unordered_map<void*, size_t> live_;
  size_t cached_ = 0, live_bytes_ = 0;
  uint64_t allocs_ = 0, alloc_ns_ = 0;
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
