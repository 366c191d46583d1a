// Context plumbing: device buffer pool, twiddle tables, pointer staging,
// FFT length support.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <condition_variable>
#include <thread>

#include <chrono>

#include "osc_internal.h"

namespace osc {

// ---------------------------------------------------------------------------
// DevPool
// ---------------------------------------------------------------------------
constexpr size_t kReuse = 4;
static size_t bucket(size_t bytes) {
  // round up to a power of two (>= 256 B): buffers of a few size classes
  // serve every shape of a mixed stream; above 1 GiB to a quarter of one
  // (at most a quarter more memory for the largest buffers)
  size_t b = 256;
  while (b < bytes) b <<= 1;
  if (b <= (size_t(1) << 30)) return b;
  const size_t step = b >> 2;
  return (bytes + step - 1) / step * step;
}

void* DevPool::get(size_t bytes) {
  size_t bb = bucket(bytes);
  if (retain) {
    // while a graph is captured, its own freed temporaries are reused first:
    // its launches run in stream order, as eager ones do
    auto best = retain->end();
    for (auto it = retain->begin(); it != retain->end(); ++it)
      if (it->second >= bb && it->second <= 2 * bb &&
          (best == retain->end() || it->second < best->second))
        best = it;
    if (best != retain->end()) {
      void* p = best->first;
      const size_t have = best->second;
      retain->erase(best);
      live_[p] = have;
      live_bytes_ += have;
      return p;
    }
  }
  // best fit among cached buffers of at most kReuse times the request, so
  // that requests of varying sizes (a stream of bunches of mixed grids) reuse
  // memory: a device allocation costs tens of microseconds to several
  // milliseconds (the runtime's state decides), a larger buffer nothing
  auto it = free_.lower_bound(bb);
  if (it != free_.end() && it->first <= kReuse * bb) {
    void* p = it->second;
    const size_t have = it->first;
    free_.erase(it);
    cached_ -= have;
    live_[p] = have;
    live_bytes_ += have;
    return p;
  }
  void* p = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
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
  ++allocs_;
  alloc_ns_ += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                            std::chrono::steady_clock::now() - t0)
                            .count());
  live_[p] = bb;
  live_bytes_ += bb;
  return p;
}

void DevPool::put(void* p) {
  auto it = live_.find(p);
  if (it == live_.end()) return;
  size_t bb = it->second;
  live_.erase(it);
  live_bytes_ -= bb;
  if (retain) {
    retain->emplace_back(p, bb);
    return;
  }
  free_.emplace(bb, p);
  cached_ += bb;
}

thread_local GraphRecorder* t_recorder = nullptr;

// ---------------------------------------------------------------------------
// FFT tile-width tuner (FftTune)
// ---------------------------------------------------------------------------
namespace {
constexpr int kTuneSamples = 2;
struct TuneEntry {
  int prec, L, kind, bucket;
  int decided = -1;
  bool distinct[3] = {false, true, false};
  float best[3] = {1e30f, 1e30f, 1e30f};
  int count[3] = {0, 0, 0};
  int pending_n[3] = {0, 0, 0};
  struct Pending {
    int var;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune;  // few keys: linear search
int g_tune_mode = 1;
bool g_tune_loaded = false;

int tune_bucket(int64_t lines) {
  int b = 0;
  while ((int64_t(1) << (b + 1)) <= lines) ++b;
  return b;
}

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = std::getenv("OSC_FFT_TUNE_FILE");
  if (!path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  int prec, L, kind, bucket, var;
  while (std::fscanf(f, "%d %d %d %d %d", &prec, &L, &kind, &bucket, &var) == 5) {
    if (var < 0 || var > 2) continue;
    TuneEntry e;
    e.prec = prec;
    e.L = L;
    e.kind = kind;
    e.bucket = bucket;
    e.decided = var;
    g_tune.push_back(e);
  }
  std::fclose(f);
}

void tune_save(const TuneEntry& e) {
  const char* path = std::getenv("OSC_FFT_TUNE_FILE");
  if (!path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%d %d %d %d %d\n", e.prec, e.L, e.kind, e.bucket, e.decided);
  std::fclose(f);
}

// reads the completed timings of e without waiting; decides when it can
void tune_poll_locked(TuneEntry& e) {
  for (size_t i = 0; i < e.pending.size();) {
    auto& p = e.pending[i];
    if (hipEventQuery(p.b) != hipSuccess) {
      (void)hipGetLastError();
      ++i;
      continue;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      e.best[p.var] = std::min(e.best[p.var], ms);
      ++e.count[p.var];
    } else {
      (void)hipGetLastError();
    }
    --e.pending_n[p.var];
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
    e.pending.erase(e.pending.begin() + long(i));
  }
  bool all = true;
  for (int v = 0; v < 3; ++v) all = all && (!e.distinct[v] || e.count[v] >= kTuneSamples);
  if (!all) return;
  int best = 1;
  for (int v = 0; v < 3; ++v)
    if (e.distinct[v] && e.best[v] < e.best[best]) best = v;
  e.decided = best;
  tune_save(e);
}
}  // namespace

FftTune::FftTune(int prec, int L, int kind, int64_t lines, const bool distinct[3], hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_tune_mode == 0) return;
  if (g_tune_mode >= 2) {
    const int v = g_tune_mode - 2;
    var_ = distinct[v] ? v : 1;
    return;
  }
  tune_load_locked();
  const int bucket = tune_bucket(lines);
  int k = -1;
  for (size_t i = 0; i < g_tune.size(); ++i) {
    const TuneEntry& e = g_tune[i];
    if (e.prec == prec && e.L == L && e.kind == kind && e.bucket == bucket) k = int(i);
  }
  if (k < 0) {
    TuneEntry e;
    e.prec = prec;
    e.L = L;
    e.kind = kind;
    e.bucket = bucket;
    for (int v = 0; v < 3; ++v) e.distinct[v] = distinct[v];
    g_tune.push_back(e);
    k = int(g_tune.size()) - 1;
  }
  TuneEntry& e = g_tune[size_t(k)];
  if (e.decided >= 0) {
    var_ = e.decided;
    return;
  }
  tune_poll_locked(e);
  if (e.decided >= 0) {
    var_ = e.decided;
    return;
  }
  // a graph being recorded cannot hold timing events: the default width
  if (t_recorder) return;
  int pick = -1;
  for (int v = 0; v < 3; ++v) {
    if (!e.distinct[v] || e.count[v] + e.pending_n[v] >= kTuneSamples) continue;
    if (pick < 0 || e.count[v] + e.pending_n[v] < e.count[pick] + e.pending_n[pick]) pick = v;
  }
  if (pick < 0) {  // every sample is in flight: the default width meanwhile
    return;
  }
  if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess ||
      hipEventRecord(a_, s) != hipSuccess) {
    (void)hipGetLastError();
    if (a_) (void)hipEventDestroy(a_);
    if (b_) (void)hipEventDestroy(b_);
    a_ = b_ = nullptr;
    return;
  }
  var_ = pick;
  key_ = k;
  ++e.pending_n[pick];
}

void FftTune::done(hipStream_t s) {
  if (key_ < 0) return;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  TuneEntry& e = g_tune[size_t(key_)];
  if (hipEventRecord(b_, s) != hipSuccess) {
    (void)hipGetLastError();
    --e.pending_n[var_];
    (void)hipEventDestroy(a_);
    (void)hipEventDestroy(b_);
    return;
  }
  e.pending.push_back({var_, a_, b_});
}

void fft_tune_mode(int mode) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_mode = mode;
}

void fft_tune_stats(int64_t* decided, int64_t* measuring) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  int64_t d = 0, m = 0;
  for (auto& e : g_tune) {
    if (e.decided < 0) tune_poll_locked(e);
    if (e.decided >= 0) ++d;
    else ++m;
  }
  if (decided) *decided = d;
  if (measuring) *measuring = m;
}

void DevPool::adopt(void* p, size_t bytes) {
  free_.emplace(bytes, p);
  cached_ += bytes;
}

void DevPool::release_free() {
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
  cached_ = 0;
}

size_t DevPool::trim(size_t target) {
  size_t freed = 0;
  while (!free_.empty() && cached_ + live_bytes_ > target) {
    auto it = std::prev(free_.end());
    (void)hipFree(it->second);
    cached_ -= it->first;
    freed += it->first;
    free_.erase(it);
  }
  return freed;
}

void DevPool::release_all() {
  release_free();
  for (auto& kv : live_) (void)hipFree(kv.first);
  live_.clear();
  live_bytes_ = 0;
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
  // on the context's stream (a legacy-stream copy fails while another thread
  // captures a stream), waited for before h goes away
  OSC_HIP(hipMemcpyAsync(d, h.data(), sizeof(cpx<T>) * L, hipMemcpyHostToDevice, c->stream));
  OSC_HIP(hipStreamSynchronize(c->stream));
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
// host copies split over a few persistent threads
// ---------------------------------------------------------------------------
class CopyThreads {
 public:
  explicit CopyThreads(int n) {
    for (int i = 1; i < n; ++i) th_.emplace_back([this, i] { run(i); });
  }
  ~CopyThreads() {
    {
      std::lock_guard<std::mutex> l(m_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  // dst[0, bytes) = src[0, bytes); returns when every part is done
  void copy(void* dst, const void* src, size_t bytes) {
    const int parts = int(th_.size()) + 1;
    if (parts == 1 || bytes < (size_t(1) << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> l(m_);
      dst_ = static_cast<char*>(dst);
      src_ = static_cast<const char*>(src);
      bytes_ = bytes;
      parts_ = parts;
      left_ = parts - 1;
      ++gen_;
    }
    cv_.notify_all();
    part(0);
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return left_ == 0; });
  }

 private:
  void part(int i) {
    // 4 KiB-aligned parts, so no two threads touch one page
    const size_t per = ((bytes_ + parts_ - 1) / parts_ + 4095) & ~size_t(4095);
    const size_t off = per * size_t(i);
    if (off < bytes_) std::memcpy(dst_ + off, src_ + off, std::min(per, bytes_ - off));
  }
  void run(int i) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      part(i);
      std::lock_guard<std::mutex> l(m_);
      if (--left_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  bool stop_ = false;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t bytes_ = 0;
  int parts_ = 1, left_ = 0;
};

// ---------------------------------------------------------------------------
// pinned ring for pageable host transfers
// ---------------------------------------------------------------------------
HostRing::HostRing() {
  for (auto& sl : slot_) {
    OSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.buf), kChunk, hipHostMallocDefault));
    OSC_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
  }
  const unsigned hw = std::thread::hardware_concurrency();
  threads_ = new CopyThreads(int(std::max(1u, std::min(hw, 8u))));
}

HostRing::~HostRing() {
  for (auto& sl : slot_) {
    if (sl.busy) (void)hipEventSynchronize(sl.ev);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.buf) (void)hipHostFree(sl.buf);
  }
  delete threads_;
}

// finishes a slot's last copy: waits for its DMA and copies out a to_host chunk
void HostRing::settle(Slot& sl) {
  if (!sl.busy) return;
  OSC_HIP(hipEventSynchronize(sl.ev));
  sl.busy = false;
  if (sl.out) threads_->copy(sl.out, sl.buf, sl.len);
  sl.out = nullptr;
}

HostRing::Slot& HostRing::take() {
  Slot& sl = slot_[next_];
  next_ = (next_ + 1) % kBufs;
  settle(sl);
  return sl;
}

void HostRing::to_device(hipStream_t s, void* dev, const void* host, size_t bytes) {
  const char* h = static_cast<const char*>(host);
  char* d = static_cast<char*>(dev);
  for (size_t off = 0; off < bytes; off += kChunk) {
    const size_t len = std::min(kChunk, bytes - off);
    Slot& sl = take();
    threads_->copy(sl.buf, h + off, len);
    OSC_HIP(hipMemcpyAsync(d + off, sl.buf, len, hipMemcpyHostToDevice, s));
    OSC_HIP(hipEventRecord(sl.ev, s));
    sl.busy = true;
    sl.seq = ++seq_;
  }
}

void HostRing::to_host(hipStream_t s, void* host, const void* dev, size_t bytes) {
  char* h = static_cast<char*>(host);
  const char* d = static_cast<const char*>(dev);
  for (size_t off = 0; off < bytes; off += kChunk) {
    const size_t len = std::min(kChunk, bytes - off);
    Slot& sl = take();
    OSC_HIP(hipMemcpyAsync(sl.buf, d + off, len, hipMemcpyDeviceToHost, s));
    OSC_HIP(hipEventRecord(sl.ev, s));
    sl.busy = true;
    sl.out = h + off;
    sl.len = len;
    sl.seq = ++seq_;
  }
}

void HostRing::drain() {
  // oldest first, in the order the copies were queued
  Slot* order[kBufs];
  for (int i = 0; i < kBufs; ++i) order[i] = &slot_[i];
  std::sort(order, order + kBufs, [](const Slot* a, const Slot* b) { return a->seq < b->seq; });
  for (Slot* sl : order) settle(*sl);
}

void HostRing::discard() {
  for (auto& sl : slot_) {
    if (sl.busy) (void)hipEventSynchronize(sl.ev);
    sl.busy = false;
    sl.out = nullptr;
  }
}

bool is_pageable(const void* p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return at.type == hipMemoryTypeUnregistered;
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

// large pageable arrays go through the context's pinned ring
static HostRing* ring_for(osc_ctx* c, const void* host, size_t bytes) {
  if (bytes < HostRing::kMin || !is_pageable(host)) return nullptr;
  if (!c->ring) c->ring = std::make_unique<HostRing>();
  return c->ring.get();
}

static void host_to_device(osc_ctx* c, void* d, const void* p, size_t bytes) {
  if (HostRing* r = ring_for(c, p, bytes))
    r->to_device(c->stream, d, p, bytes);
  else
    OSC_HIP(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, c->stream));
}

template <typename U>
const U* Staging::in(const U* p, size_t n) {
  if (!p || n == 0) return p;
  if (is_device_ptr(p)) return p;
  any_host_ = true;
  bufs_.emplace_back(c_->pool, n * sizeof(U));
  void* d = bufs_.back().p;
  host_to_device(c_, d, p, n * sizeof(U));
  return static_cast<const U*>(d);
}

template <typename U>
U* Staging::out(U* p, size_t n, bool copy_in) {
  if (!p || n == 0) return p;
  if (is_device_ptr(p)) return p;
  any_host_ = true;
  bufs_.emplace_back(c_->pool, n * sizeof(U));
  void* d = bufs_.back().p;
  if (copy_in) host_to_device(c_, d, p, n * sizeof(U));
  outs_.push_back({p, d, n * sizeof(U)});
  return static_cast<U*>(d);
}

Staging::~Staging() {
  if (ring_pending_ && c_->ring) c_->ring->discard();
}

void Staging::finish() {
  for (auto& o : outs_) {
    if (HostRing* r = ring_for(c_, o.host, o.bytes)) {
      ring_pending_ = true;
      r->to_host(c_->stream, o.host, o.dev, o.bytes);
    } else {
      OSC_HIP(hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, c_->stream));
    }
  }
  if (ring_pending_) {
    c_->ring->drain();
    ring_pending_ = false;
  }
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
  while (bytes > budget) {
    auto victim = map.end();
    for (auto it = map.begin(); it != map.end(); ++it)
      if (it->second->last < pin_from &&
          (victim == map.end() || it->second->last < victim->second->last))
        victim = it;
    if (victim == map.end()) return;  // everything left is in use
    bytes -= victim->second->bytes;
    map.erase(victim);
  }
}

}  // namespace osc
