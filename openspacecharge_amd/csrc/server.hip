// Request loop for a stream of mixed space-charge solves on one GPU (no
// upstream counterpart: upstream solves one bunch per call on the CPU).
//
// Workers: each owns an osc_ctx with its own HIP stream, buffer pool and
// Green-spectrum cache, and a host thread that runs one request at a time
// through the C-ABI (deposit -> solve -> gather), so requests on different
// workers overlap on the device.  Admission is work-conserving: requests wait
// in one queue, oldest first, and an idle worker takes the oldest request
// that fits the memory budget, preferring one (among the first 2 x nworkers)
// whose geometry key it served recently, since its context holds that Green
// spectrum.  A request that does not fit may be passed over by smaller ones
// at most kMaxBypass times; after that nothing starts before it.  Memory
// accounting: the budget (free HBM at creation less an eighth, or the
// caller's limit) covers the estimated bytes of the running requests plus
// what every worker's context holds between requests (Green caches, the pipe
// Green functions, captured solves, pooled buffers); each worker's Green cache is capped at
// budget / (8 nworkers).  A worker holding more than budget / (2 nworkers)
// after a request gives back its largest free pooled buffers until it holds
// three quarters of that: freeing stalls the whole device and the buffers
// are allocated again by later requests, so it gives back only what it must.
// A request larger than the budget runs alone.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "osc_internal.h"

#define OSC_API extern "C" __attribute__((visibility("default")))

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
  int64_t ticket;
  osc_request r;
  int64_t est;
  uint64_t key;
  int bypassed = 0;
  Clock::time_point t_submit;
};

// a request that does not fit may be passed over this many times
constexpr int kMaxBypass = 8;
// route keys a worker remembers as served (its likely cached Green spectra)
constexpr size_t kAffinityKeys = 16;

struct Done {
  int status = OSC_OK;
  double queued_ms = 0, run_ms = 0, total_ms = 0;
  int64_t n_dropped = 0, n_clamped = 0;
};

uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
  return h;
}
uint64_t dbits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, sizeof(u));
  return u;
}

// geometry key: requests that can share Green spectra
uint64_t route_key(const osc_request& r) {
  uint64_t h = mix(0, uint64_t(r.solver));
  for (int a = 0; a < 3; ++a) h = mix(h, uint64_t(r.nhi[a]));
  h = mix(h, dbits(r.gamma));
  if (!r.resize_mesh)
    for (int a = 0; a < 3; ++a) h = mix(mix(h, dbits(r.min[a])), dbits(r.max[a]));
  if (r.solver == OSC_SOLVE_PIPE) h = mix(mix(h, dbits(r.apipe)), dbits(r.bpipe));
  return h;
}

// device bytes of one request: mesh arrays, doubled-grid half spectra (rho,
// up to four Green terms, the inverse work buffer), the real Green build,
// node records, and the particle sort
int64_t estimate_bytes(const osc_request& r) {
  const int64_t n = int64_t(r.nhi[0]) * r.nhi[1] * r.nhi[2];
  const int64_t half = (int64_t(r.nhi[0]) + 1) * 4 * int64_t(r.nhi[1]) * r.nhi[2] * 16;
  return 64 * n + 10 * half + 64 * n + 64 * n + 96 * r.n + (int64_t(64) << 20);
}

}  // namespace

struct osc_server {
  int device = 0;
  int green_cache = 1;
  int64_t budget = 0, reserved = 0;
  int running = 0;
  std::vector<osc_ctx*> ctxs;
  std::vector<std::thread> threads;
  std::deque<Job> queue;                    // admission order, oldest first
  std::vector<std::deque<uint64_t>> affinity;  // per worker: keys served recently
  std::vector<int64_t> held;                // per worker: bytes its context holds when idle
  int64_t held_total = 0;
  int64_t n_done = 0, max_running = 0, n_affine = 0, n_bypass = 0;
  int64_t n_trims = 0, trimmed_bytes = 0;
  std::unordered_map<int64_t, Done> done;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  int64_t next_ticket = 1;
  int64_t in_flight = 0;
  bool stop = false;
  std::string last_error;

  void run_worker(int w);
  int run_one(osc_ctx* c, const osc_request& r, Done& d);
  // index in `queue` of the request worker w should start now, or -1
  int pick(int w);
  void set_held(int w, int64_t bytes);
};

int osc_server::pick(int w) {
  if (queue.empty()) return -1;
  auto fits = [&](const Job& j) {
    return running == 0 || held_total + reserved + j.est <= budget;
  };
  const auto& aff = affinity[w];
  auto mine = [&](uint64_t k) {
    for (uint64_t a : aff)
      if (a == k) return true;
    return false;
  };
  const size_t look = std::min(queue.size(), 2 * ctxs.size());
  for (size_t i = 0; i < look; ++i)
    if (mine(queue[i].key) && fits(queue[i]) && (i == 0 || fits(queue[0]) ||
                                                  queue[0].bypassed < kMaxBypass))
      return int(i);
  if (fits(queue[0])) return 0;
  if (queue[0].bypassed >= kMaxBypass) return -1;  // the head waits for memory
  for (size_t i = 1; i < queue.size(); ++i)
    if (fits(queue[i])) return int(i);
  return -1;
}

void osc_server::set_held(int w, int64_t bytes) {
  held_total += bytes - held[w];
  held[w] = bytes;
}

int osc_server::run_one(osc_ctx* c, const osc_request& r, Done& d) {
  using namespace osc;
  int rc = osc_ctx_set_precision(c, r.precision);
  if (rc) return rc;
  osc_mesh m{};
  for (int a = 0; a < 3; ++a) {
    m.nlo[a] = 1;
    m.nhi[a] = r.nhi[a];
    m.npad[a] = 1;
    m.min[a] = r.min[a];
    m.max[a] = r.max[a];
    m.delta[a] = r.nhi[a] > 1 ? (r.max[a] - r.min[a]) / double(r.nhi[a] - 1) : 1.0;
  }
  m.gamma = r.gamma;
  const int64_t n = int64_t(r.nhi[0]) * r.nhi[1] * r.nhi[2];
  const int want_b = r.B != nullptr;
  DevBuf rho(c->pool, 8 * n), ef(c->pool, 24 * n), bf(c->pool, 24 * n);
  const double tc = r.total_charge;
  if (r.solver == OSC_SOLVE_FREE || r.solver == OSC_SOLVE_CATHODE) {
    rc = osc_space_charge_step(c, r.n, r.x, r.y, r.z, nullptr, &tc, r.resize_mesh, nullptr,
                               r.solver == OSC_SOLVE_CATHODE, want_b, &m, rho.as<double>(),
                               ef.as<double>(), want_b ? bf.as<double>() : nullptr, r.E, r.B,
                               &d.n_dropped, &d.n_clamped);
    if (rc) return rc;
  } else if (r.solver == OSC_SOLVE_PIPE) {
    if (r.resize_mesh) {
      c->last_error = "pipe requests need a fixed mesh (the pipe cross-section)";
      return OSC_ERR_ARG;
    }
    rc = osc_deposit_particles(c, r.n, r.x, r.y, r.z, nullptr, &tc, 0, &m, rho.as<double>(),
                               &d.n_dropped);
    if (rc) return rc;
    // the pipe Green functions stay in the context while the geometry repeats
    const PipeState& ps = c->pipe;
    bool same = ps.built && ps.a == r.apipe && ps.b == r.bpipe && ps.gamma == r.gamma;
    for (int a = 0; a < 3 && same; ++a)
      same = ps.nlo[a] == 1 && ps.nhi[a] == r.nhi[a] && ps.npad[a] == 1 &&
             ps.delta[a] == m.delta[a] && ps.umin[a] == m.min[a];
    if (!same) {
      const int32_t one[3] = {1, 1, 1};
      rc = osc_alloc_rectpipe_arrays(c, one, r.nhi, one);
      if (!rc) rc = osc_getgrnpipe(c, r.gamma, r.apipe, r.bpipe, m.delta, m.min, one);
      if (rc) return rc;
    }
    DevBuf phi(c->pool, 8 * n);
    rc = osc_space_charge_rectpipe(c, &m, r.apipe, r.bpipe, 0, 1, rho.as<double>(),
                                   phi.as<double>(), ef.as<double>(), bf.as<double>());
    if (rc) return rc;
    rc = osc_interpolate_field(c, &m, ef.as<double>(), want_b ? bf.as<double>() : nullptr, r.n,
                               r.x, r.y, r.z, r.E, r.B, nullptr, &d.n_clamped);
    // the deposit's kept cell order has served its gather: back to the pool
    c->kept_ok = false;
    c->kept = SortedParticles{};
    if (rc) return rc;
  } else {
    c->last_error = "unknown solver";
    return OSC_ERR_ARG;
  }
  return osc_ctx_synchronize(c);
}

void osc_server::run_worker(int w) {
  osc_ctx* c = ctxs[w];
  // freeing pooled buffers (hipFree) stalls the device and the next request
  // then allocates again: only a context holding more than this gives back
  const int64_t share = budget / int64_t(2 * ctxs.size());
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(mu);
      int at = -1;
      // after stop, the queued requests are still served
      cv_work.wait(lk, [&] {
        if (stop && queue.empty()) return true;
        at = pick(w);
        return at >= 0;
      });
      if (at < 0) return;
      job = queue[size_t(at)];
      if (at > 0) {
        ++queue[0].bypassed;
        ++n_bypass;
      }
      for (uint64_t a : affinity[w]) n_affine += a == job.key ? 1 : 0;
      queue.erase(queue.begin() + at);
      reserved += job.est;
      ++running;
      max_running = std::max<int64_t>(max_running, running);
    }
    const auto t_start = Clock::now();
    Done d;
    d.status = run_one(c, job.r, d);
    const auto t_end = Clock::now();
    d.queued_ms = std::chrono::duration<double, std::milli>(t_start - job.t_submit).count();
    d.run_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
    d.total_ms = std::chrono::duration<double, std::milli>(t_end - job.t_submit).count();
    // what the context keeps between requests: Green cache, pipe Green
    // functions and pooled buffers; past the share, free pooled buffers go
    // back, largest first, down to three quarters of it
    size_t trimmed = 0;
    if (int64_t(c->pool.bytes_cached() + c->pool.bytes_live()) > share)
      trimmed = c->pool.trim(size_t(share / 4 * 3));
    int64_t keep = int64_t(c->pool.bytes_cached() + c->pool.bytes_live());
    for (auto& kv : c->sgraphs)  // the temporaries of captured solves
      for (auto& pb : kv.second->kept) keep += int64_t(pb.second);
    {
      std::lock_guard<std::mutex> lk(mu);
      if (d.status != OSC_OK) last_error = c->last_error;
      reserved -= job.est;
      --running;
      if (trimmed) {
        ++n_trims;
        trimmed_bytes += int64_t(trimmed);
      }
      set_held(w, keep);
      auto& aff = affinity[w];
      bool have = false;
      for (uint64_t a : aff) have = have || a == job.key;
      if (!have) {
        aff.push_back(job.key);
        if (aff.size() > kAffinityKeys) aff.pop_front();
      }
      --in_flight;
      ++n_done;
      done[job.ticket] = d;
    }
    cv_work.notify_all();
    cv_done.notify_all();
  }
}

OSC_API int osc_server_create(int device, int nworkers, int64_t max_bytes, int green_cache,
                              osc_server** out) {
  if (!out || nworkers < 1 || nworkers > 16) return OSC_ERR_ARG;
  *out = nullptr;
  auto* s = new osc_server();
  s->device = device;
  s->green_cache = green_cache;
  for (int w = 0; w < nworkers; ++w) {
    osc_ctx* c = nullptr;
    const int rc = osc_ctx_create(device, &c);
    if (rc) {
      for (auto* cc : s->ctxs) osc_ctx_destroy(cc);
      delete s;
      return rc;
    }
    s->ctxs.push_back(c);
  }
  size_t fr = 0, tot = 0;
  {
    osc::DeviceGuard g(device);
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      (void)hipGetLastError();
      fr = size_t(16) << 30;
    }
  }
  const int64_t avail = int64_t(fr) - int64_t(fr / 8);  // keep an eighth free
  s->budget = max_bytes > 0 ? std::min<int64_t>(max_bytes, avail) : avail;
  for (auto* c : s->ctxs)
    osc_ctx_set_green_cache(c, green_cache, std::max<int64_t>(1, s->budget / (8 * nworkers)));
  s->affinity.resize(nworkers);
  s->held.assign(nworkers, 0);
  for (int w = 0; w < nworkers; ++w) s->threads.emplace_back([s, w] { s->run_worker(w); });
  *out = s;
  return OSC_OK;
}

OSC_API int osc_server_submit(osc_server* s, const osc_request* r, int64_t* ticket) {
  if (!s || !r || !ticket) return OSC_ERR_ARG;
  if (r->n < 1 || !r->x || !r->y || !r->z || !r->E || r->nhi[0] < 2 || r->nhi[1] < 2 ||
      r->nhi[2] < 2)
    return OSC_ERR_ARG;
  Job j;
  j.r = *r;
  j.est = estimate_bytes(*r);
  j.key = route_key(*r);
  j.t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    j.ticket = s->next_ticket++;
    *ticket = j.ticket;
    ++s->in_flight;
    s->queue.push_back(j);
  }
  s->cv_work.notify_all();
  return OSC_OK;
}

OSC_API int osc_server_wait(osc_server* s, int64_t ticket, double* times_ms, int64_t* counts) {
  if (!s) return OSC_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (ticket < 1 || ticket >= s->next_ticket) return OSC_ERR_ARG;
  s->cv_done.wait(lk, [&] { return s->done.count(ticket) > 0; });
  const Done d = s->done[ticket];
  s->done.erase(ticket);
  if (times_ms) {
    times_ms[0] = d.queued_ms;
    times_ms[1] = d.run_ms;
    times_ms[2] = d.total_ms;
  }
  if (counts) {
    counts[0] = d.n_dropped;
    counts[1] = d.n_clamped;
  }
  return d.status;
}

OSC_API int osc_server_drain(osc_server* s) {
  if (!s) return OSC_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->in_flight == 0; });
  return OSC_OK;
}

// Once no request is in flight (the workers' streams then idle), every
// worker's cached Green spectra, its keys seen once and its pipe Green
// functions go back to its buffer pool: no device memory is freed.
OSC_API int osc_server_clear_caches(osc_server* s) {
  if (!s) return OSC_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->in_flight == 0; });
  for (auto* c : s->ctxs) {
    c->gcache.clear();
    c->pipe = osc::PipeState{};
  }
  return OSC_OK;
}

OSC_API int osc_server_stats(osc_server* s, int64_t* out, int n) {
  if (!s || !out || n < 1) return OSC_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  int64_t hits = 0, misses = 0, allocs = 0, alloc_us = 0;
  for (auto* c : s->ctxs) {
    hits += int64_t(c->gcache.hits);
    misses += int64_t(c->gcache.misses);
    allocs += int64_t(c->pool.allocs());
    alloc_us += int64_t(c->pool.alloc_ns() / 1000);
  }
  const int64_t v[OSC_SERVER_NSTATS] = {s->n_done,  s->max_running, s->n_affine,   s->n_bypass,
                                        hits,       misses,         s->held_total, s->budget,
                                        s->n_trims, s->trimmed_bytes, allocs, alloc_us};
  for (int i = 0; i < n && i < OSC_SERVER_NSTATS; ++i) out[i] = v[i];
  return OSC_OK;
}

OSC_API const char* osc_server_last_error(const osc_server* s) {
  return s ? s->last_error.c_str() : "null server";
}

OSC_API int osc_server_destroy(osc_server* s) {
  if (!s) return OSC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->threads) t.join();
  for (auto* c : s->ctxs) osc_ctx_destroy(c);
  delete s;
  return OSC_OK;
}
