// Context plumbing: device buffer pool, twiddle tables, pointer staging,
// FFT length support.
#include <cmath>

#include "osc_internal.h"

namespace osc {

// ---------------------------------------------------------------------------
// DevPool
// ---------------------------------------------------------------------------
static size_t bucket(size_t bytes) {
  // round to 1/8 of the next power of two (>= 256 B) so shapes reuse buffers
  size_t b = 256;
  while (b < bytes) b <<= 1;
  size_t step = b >> 3;
  return step ? (bytes + step - 1) / step * step : b;
}

void* DevPool::get(size_t bytes) {
  size_t bb = bucket(bytes);
  // best fit among cached buffers of at most twice the request, so that
  // requests of varying sizes (a stream of bunches) reuse memory
  auto it = free_.lower_bound(bb);
  if (it != free_.end() && it->first <= 2 * bb) {
    void* p = it->second;
    const size_t have = it->first;
    free_.erase(it);
    cached_ -= have;
    live_[p] = have;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bb);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // release cached buffers and retry once
    release_free();
    e = hipMalloc(&p, bb);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error{OSC_ERR_NOMEM, "device allocation of " + std::to_string(bb) + " bytes failed"};
    }
  }
  live_[p] = bb;
  return p;
}

void DevPool::put(void* p) {
  auto it = live_.find(p);
  if (it == live_.end()) return;
  size_t bb = it->second;
  live_.erase(it);
  free_.emplace(bb, p);
  cached_ += bb;
}

void DevPool::release_free() {
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
  cached_ = 0;
}

void DevPool::release_all() {
  release_free();
  for (auto& kv : live_) (void)hipFree(kv.first);
  live_.clear();
}

// ---------------------------------------------------------------------------
// twiddles: tw[m] = exp(-2 pi i m / L), computed in long double
// ---------------------------------------------------------------------------
template <typename T>
const cpx<T>* twiddles(osc_ctx* c, int L) {
  auto& tab = sizeof(T) == 8 ? c->tw64 : c->tw32;
  auto it = tab.find(L);
  if (it != tab.end()) return static_cast<const cpx<T>*>(it->second);
  std::vector<cpx<T>> h(L);
  const long double twopi = 6.28318530717958647692528676655900577L;
  for (int m = 0; m < L; ++m) {
    // reduce by symmetry so the argument stays in [0, pi/4] where possible
    long double ang = twopi * (long double)m / (long double)L;
    h[m].x = (T)cosl(ang);
    h[m].y = (T)(-sinl(ang));
  }
  // exact values at the quarter points
  for (int m = 0; m < L; ++m) {
    if ((4 * m) % L == 0) {
      int q = (4 * m) / L;
      const T cs[4] = {1, 0, -1, 0}, sn[4] = {0, -1, 0, 1};
      h[m].x = cs[q];
      h[m].y = sn[q];
    }
  }
  void* d = nullptr;
  OSC_HIP(hipMalloc(&d, sizeof(cpx<T>) * L));
  OSC_HIP(hipMemcpy(d, h.data(), sizeof(cpx<T>) * L, hipMemcpyHostToDevice));
  tab[L] = d;
  return static_cast<const cpx<T>*>(d);
}
template const cpx<double>* twiddles<double>(osc_ctx*, int);
template const cpx<float>* twiddles<float>(osc_ctx*, int);

// ---------------------------------------------------------------------------
// FFT lengths
// ---------------------------------------------------------------------------
bool fft_has_plan(int L) {
  switch (L) {
#define OSC_CASE(LL) case LL:
    OSC_FOR_EACH_L(OSC_CASE)
#undef OSC_CASE
    return true;
    default:
      return false;
  }
}

GenericPlan generic_plan(int L) {
  GenericPlan p;
  p.L = L;
  p.ns = 0;
  int n = L;
  for (int f = 2; f <= 23 && n > 1; ++f)
    while (n % f == 0 && p.ns < 24) {
      p.r[p.ns++] = f;
      n /= f;
    }
  if (n != 1) p.ns = -1;
  return p;
}

bool fft_len_supported(int L) {
  if (L < 1) return false;
  if (L == 1) return true;
  if (fft_has_plan(L)) return true;
  GenericPlan p = generic_plan(L);
  if (p.ns < 0) return false;
  // generic engine keeps 2 lines of L complex doubles in LDS
  return size_t(LdsPad<double>::size(2 * kGenBX * L)) * sizeof(cpx<double>) <= 65536;
}

// ---------------------------------------------------------------------------
// pointers and staging
// ---------------------------------------------------------------------------
bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice;
}

template <typename U>
const U* Staging::in(const U* p, size_t n) {
  if (!p || n == 0) return p;
  if (is_device_ptr(p)) return p;
  any_host_ = true;
  bufs_.emplace_back(c_->pool, n * sizeof(U));
  void* d = bufs_.back().p;
  OSC_HIP(hipMemcpyAsync(d, p, n * sizeof(U), hipMemcpyHostToDevice, c_->stream));
  return static_cast<const U*>(d);
}

template <typename U>
U* Staging::out(U* p, size_t n, bool copy_in) {
  if (!p || n == 0) return p;
  if (is_device_ptr(p)) return p;
  any_host_ = true;
  bufs_.emplace_back(c_->pool, n * sizeof(U));
  void* d = bufs_.back().p;
  if (copy_in) OSC_HIP(hipMemcpyAsync(d, p, n * sizeof(U), hipMemcpyHostToDevice, c_->stream));
  outs_.push_back({p, d, n * sizeof(U)});
  return static_cast<U*>(d);
}

void Staging::finish() {
  for (auto& o : outs_)
    OSC_HIP(hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, c_->stream));
  if (any_host_ || !c_->async) OSC_HIP(hipStreamSynchronize(c_->stream));
  outs_.clear();
}

#define OSC_INST_STAGE(U)                              \
  template const U* Staging::in<U>(const U*, size_t); \
  template U* Staging::out<U>(U*, size_t, bool);
OSC_INST_STAGE(double)
OSC_INST_STAGE(int8_t)
#undef OSC_INST_STAGE

GreenSpec make_green_spec() {
  GreenSpec s;
  std::memset(&s, 0, sizeof(s));
  return s;
}

void GreenCache::evict_to(size_t budget) {
  while (bytes > budget && !map.empty()) {
    auto victim = map.begin();
    for (auto it = map.begin(); it != map.end(); ++it)
      if (it->second->last < victim->second->last) victim = it;
    bytes -= victim->second->bytes;
    map.erase(victim);
  }
}

}  // namespace osc
