// extern "C" entry points (include/opensc_hip.h).
#include <cmath>
#include <cstdlib>
#include <exception>

#include "solvers.h"

using namespace osc;

#define OSC_API extern "C" __attribute__((visibility("default")))

namespace {

template <class F>
int guarded(osc_ctx* c, F&& f) {
  if (!c) return OSC_ERR_ARG;
  try {
    DeviceGuard g(c->device);
    int rc = f();
    if (rc == OSC_OK) c->last_error.clear();
    return rc;
  } catch (const Error& e) {
    c->last_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    c->last_error = e.what();
    return OSC_ERR_ARG;
  }
}

int fail(osc_ctx* c, int code, const std::string& msg) {
  c->last_error = msg;
  return code;
}

void mesh_dims(const osc_mesh* m, int n[3]) {
  for (int a = 0; a < 3; ++a) n[a] = m->nhi[a] - m->nlo[a] + 1;
}

bool dims_ok(const int n[3]) { return n[0] >= 1 && n[1] >= 1 && n[2] >= 1; }

int64_t prod3(const int n[3]) { return int64_t(n[0]) * n[1] * n[2]; }

int check_len(osc_ctx* c, const int L[3]) {
  for (int a = 0; a < 3; ++a)
    if (!fft_len_supported(L[a]))
      return fail(c, OSC_ERR_FFT_LENGTH,
                  "Error return from FFT package due to transform length " + std::to_string(L[a]));
  return OSC_OK;
}

__global__ void k_real_to_cplx(const double* __restrict__ g, double* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    out[2 * i] = g[i];
    out[2 * i + 1] = 0.0;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
OSC_API int osc_abi_version(void) { return OSC_ABI_VERSION; }

OSC_API int osc_fft_length_supported(int n) { return fft_len_supported(n) ? 1 : 0; }

OSC_API int osc_fft_length_planned(int n) { return fft_has_plan(n) ? 1 : 0; }

OSC_API int osc_fft_tuning(int mode) {
  if (mode < 0 || mode > 4) return OSC_ERR_ARG;
  fft_tune_mode(mode);
  return OSC_OK;
}

OSC_API int osc_fft_tune_stats(int64_t* out2) {
  if (!out2) return OSC_ERR_ARG;
  fft_tune_stats(&out2[0], &out2[1]);
  return OSC_OK;
}

OSC_API int osc_ctx_create(int device, osc_ctx** out) {
  if (!out) return OSC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    (void)hipGetLastError();
    return OSC_ERR_NO_DEVICE;
  }
  osc_ctx* c = new osc_ctx();
  c->device = device;
  DeviceGuard g(device);
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->counters, kCounterBytes) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->mesh_host), kPinnedBytes,
                    hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->mesh_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->hash_ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    if (c->mesh_ev) (void)hipEventDestroy(c->mesh_ev);
    if (c->mesh_host) (void)hipHostFree(c->mesh_host);
    if (c->counters) (void)hipFree(c->counters);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return OSC_ERR_HIP;
  }
  c->hash_host = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(c->mesh_host) +
                                                       kHashHost);
  c->stream = c->own;
  if (const char* td = std::getenv("OSC_TILED_DEPOSIT")) c->tiled_deposit = td[0] != '0';
  if (const char* fg = std::getenv("OSC_FAR_GREEN")) c->far_green = fg[0] != '0';
  if (const char* ga = std::getenv("OSC_GREEN_ADMIT")) c->gcache.admit_second = ga[0] != '0';
  // the device address of the no-allocation gather's scratch on this device
  const int rc = noalloc_setup(device, &c->last_error);
  if (rc != OSC_OK) {
    osc_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return OSC_OK;
}

OSC_API int osc_ctx_destroy(osc_ctx* c) {
  if (!c) return OSC_OK;
  {
    DeviceGuard g(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->stream != c->own) (void)hipStreamSynchronize(c->own);
    c->ring.reset();
    c->kept = SortedParticles{};
    c->sgraphs.clear();
    c->gcache.clear();
    for (int t = 0; t < 4; ++t) {
      c->pipe.spec64[t].reset();
      c->pipe.spec32[t].reset();
    }
    c->pool.release_all();
    for (auto& r : c->prof_recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : c->prof_free) (void)hipEventDestroy(e);
    for (auto& kv : c->tw64) (void)hipFree(kv.second);
    for (auto& kv : c->tw32) (void)hipFree(kv.second);
    if (c->counters) (void)hipFree(c->counters);
    if (c->mesh_host) (void)hipHostFree(c->mesh_host);
    if (c->mesh_ev) (void)hipEventDestroy(c->mesh_ev);
    if (c->hash_ev) (void)hipEventDestroy(c->hash_ev);
    if (c->own) (void)hipStreamDestroy(c->own);
  }
  delete c;
  return OSC_OK;
}

OSC_API int osc_ctx_set_stream(osc_ctx* c, void* s) {
  return guarded(c, [&] {
    OSC_HIP(hipStreamSynchronize(c->stream));
    c->stream = s ? static_cast<hipStream_t>(s) : c->own;
    return OSC_OK;
  });
}

OSC_API int osc_ctx_set_async(osc_ctx* c, int async) {
  if (!c) return OSC_ERR_ARG;
  c->async = async != 0;
  return OSC_OK;
}

OSC_API int osc_ctx_set_precision(osc_ctx* c, int prec) {
  if (!c || (prec != OSC_PREC_FP64 && prec != OSC_PREC_FP32)) return OSC_ERR_ARG;
  c->prec = prec;
  return OSC_OK;
}

OSC_API int osc_ctx_set_green_cache(osc_ctx* c, int enable, int64_t max_bytes) {
  return guarded(c, [&] {
    OSC_HIP(hipStreamSynchronize(c->stream));
    c->gcache.enable = enable != 0;
    if (max_bytes > 0) c->gcache.max_bytes = size_t(max_bytes);
    if (!c->gcache.enable) c->gcache.clear();
    else c->gcache.evict_to(c->gcache.max_bytes);
    return OSC_OK;
  });
}

OSC_API int osc_ctx_set_graphs(osc_ctx* c, int enable) {
  return guarded(c, [&] {
    OSC_HIP(hipStreamSynchronize(c->stream));
    c->graphs = enable != 0;
    if (!c->graphs) c->sgraphs.clear();
    return OSC_OK;
  });
}

OSC_API int osc_ctx_graph_stats(osc_ctx* c, int64_t* out3) {
  if (!c || !out3) return OSC_ERR_ARG;
  out3[0] = int64_t(c->sgraphs.size());
  out3[1] = c->graph_builds;
  out3[2] = c->graph_replays;
  return OSC_OK;
}

OSC_API int osc_ctx_clear_caches(osc_ctx* c) {
  return guarded(c, [&] {
    OSC_HIP(hipStreamSynchronize(c->stream));
    c->kept = SortedParticles{};
    c->kept_ok = false;
    c->sgraphs.clear();
    c->gcache.clear();
    c->pool.release_free();
    return OSC_OK;
  });
}

OSC_API int osc_ctx_synchronize(osc_ctx* c) {
  return guarded(c, [&] {
    OSC_HIP(hipStreamSynchronize(c->stream));
    return OSC_OK;
  });
}

OSC_API const char* osc_ctx_last_error(const osc_ctx* c) {
  return c ? c->last_error.c_str() : "null context";
}

static_assert(OSC_NSTAGE == ST_COUNT, "stage table");

OSC_API const char* osc_stage_name(int s) {
  static const char* names[ST_COUNT] = {
      "deposit.bbox", "deposit.keys", "deposit.sort", "deposit.cells", "deposit.nodes",
      "green.eval",   "green.fft",    "fft.fwd_x",    "fft.fwd_y",     "fft.fwd_z",
      "fft.inv_zmul", "fft.inv_y",    "fft.inv_x",    "gather",        "fd_bfield",
      "other",        "gather.pack",  "gather.sort",  "gather.out",    "deposit.fix"};
  return (s >= 0 && s < ST_COUNT) ? names[s] : "?";
}

OSC_API int osc_ctx_set_profiling(osc_ctx* c, int on) {
  return guarded(c, [&] {
    c->prof = on != 0;
    return OSC_OK;
  });
}

OSC_API int osc_ctx_profile_read(osc_ctx* c, double* ms, int64_t* launches) {
  return guarded(c, [&] {
    OSC_HIP(hipStreamSynchronize(c->stream));
    for (auto& r : c->prof_recs) {
      float t = 0.f;
      OSC_HIP(hipEventElapsedTime(&t, r.a, r.b));
      c->prof_ms[r.stage] += t;
      c->prof_n[r.stage] += 1;
      c->prof_free.push_back(r.a);
      c->prof_free.push_back(r.b);
    }
    c->prof_recs.clear();
    for (int s = 0; s < ST_COUNT; ++s) {
      if (ms) ms[s] = c->prof_ms[s];
      if (launches) launches[s] = c->prof_n[s];
      c->prof_ms[s] = 0;
      c->prof_n[s] = 0;
    }
    return OSC_OK;
  });
}

// ---------------------------------------------------------------------------
// deposition
// ---------------------------------------------------------------------------
// deposit_particles body on staged (device) coordinates; rho is device memory
// ---------------------------------------------------------------------------
// captured solves of the fused step
// ---------------------------------------------------------------------------
namespace {
MeshDev* mesh_dev(osc_ctx* c) {
  return reinterpret_cast<MeshDev*>(static_cast<char*>(c->counters) + kMeshDev);
}

// graphs per context (least recently used evicted) and the largest mesh
// captured: beyond it the launches are a small part of a step, and each graph
// keeps its own buffers (about 0.8 GB at 128^3 in FP64)
constexpr size_t kMaxGraphs = 8;
constexpr int64_t kGraphMaxNodes = int64_t(1) << 21;

std::string graph_key(osc_ctx* c, const int nn[3], double gamma, int cath, int calcb) {
  struct {
    int32_t n[3], prec, cath, calcb;
    double gamma;
  } k{{nn[0], nn[1], nn[2]}, c->prec, cath, calcb, gamma};
  return std::string(reinterpret_cast<const char*>(&k), sizeof(k));
}

SolveGraph* step_graph(osc_ctx* c, const int nn[3], double gamma, int cath, int calcb) {
  auto it = c->sgraphs.find(graph_key(c, nn, gamma, cath, calcb));
  if (it == c->sgraphs.end()) return nullptr;
  it->second->last = ++c->graph_tick;
  return it->second.get();
}

// After an eager step of a new (shape, precision, solver, B, gamma): record
// the same solve as graph nodes for the following steps (osc::launch adds
// each kernel as a node while the recorder is set; nothing runs).  The eager
// step has created the twiddle tables and left its temporaries in the pool,
// so recording allocates (almost) nothing; a recording that fails leaves the
// eager path in place for this shape.
void note_step_shape(osc_ctx* c, const int nn[3], double gamma, int cath, int calcb,
                     const osc_mesh& m) {
  const int64_t nt = prod3(nn);
  if (nt > kGraphMaxNodes || c->stream == nullptr) return;
  const std::string key = graph_key(c, nn, gamma, cath, calcb);
  if (c->sgraphs.count(key)) return;
  if (c->sgraphs.size() >= kMaxGraphs) {
    auto victim = c->sgraphs.begin();
    for (auto it = c->sgraphs.begin(); it != c->sgraphs.end(); ++it)
      if (it->second->last < victim->second->last) victim = it;
    OSC_HIP(hipStreamSynchronize(c->stream));
    c->sgraphs.erase(victim);
  }
  auto g = std::make_unique<SolveGraph>();
  g->pool = &c->pool;
  g->rho = DevBuf(c->pool, sizeof(double) * nt);
  g->ef = DevBuf(c->pool, sizeof(double) * 3 * nt);
  if (calcb) g->bf = DevBuf(c->pool, sizeof(double) * 3 * nt);
  if (hipGraphCreate(&g->graph, 0) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  GraphRecorder rec;
  rec.graph = g->graph;
  c->green_md = mesh_dev(c);
  c->pool.retain = &g->kept;
  t_recorder = &rec;
  bool ok = true;
  try {
    space_charge_3d(c, m, g->rho.as<double>(), nullptr, cath, calcb, g->ef.as<double>(),
                    calcb ? g->bf.as<double>() : nullptr);
  } catch (const Error&) {
    ok = false;
  }
  t_recorder = nullptr;
  c->green_md = nullptr;
  c->pool.retain = nullptr;
  if (!ok || rec.failed || !rec.last ||
      hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    return;  // g's destructor gives its buffers back
  }
  g->last = ++c->graph_tick;
  ++c->graph_builds;
  c->sgraphs.emplace(key, std::move(g));
}

}  // namespace

// finish_deposit_mesh: the host side of a deposit queued with defer_mesh
static void finish_deposit_mesh(osc_ctx* c, bool resize_mesh, bool qa, osc_mesh* m) {
  OSC_HIP(hipEventSynchronize(c->mesh_ev));
  const MeshDev& h = *c->mesh_host;
  if (resize_mesh)
    for (int d = 0; d < 3; ++d) {
      m->min[d] = h.mn[d];
      m->max[d] = h.mx[d];
      m->delta[d] = h.d[d];
    }
  if (qa) m->charge = h.charge;
}

// defer_mesh (with *deferred): a resized mesh or charge sum is not waited
// for here; *deferred tells the caller to call finish_deposit_mesh
static int deposit_staged(osc_ctx* c, Staging& st, int64_t n, const double* x, const double* y,
                          const double* z, const double* qa, const double* total_charge,
                          int resize_mesh, osc_mesh* m, double* rho, SortedParticles* keep,
                          unsigned long long* hash = nullptr, bool* deferred = nullptr,
                          bool f32 = false) {
  int nn[3];
  mesh_dims(m, nn);
  // The resized mesh (:176-189) and the sum of the charges (:244-245) are
  // computed on the device; the deposit kernels read the mesh there, and the
  // host waits for its copy only after they are queued.
  MeshDev* md = reinterpret_cast<MeshDev*>(static_cast<char*>(c->counters) + kMeshDev);
  DepositArgs a{};
  a.n = n;
  a.x = x;
  a.y = y;
  a.z = z;
  a.hash = hash;
  a.f32 = f32;
  bool readback = false;
  if (resize_mesh) {
    const int cells[3] = {m->nhi[0] - m->nlo[0], m->nhi[1] - m->nlo[1], m->nhi[2] - m->nlo[2]};
    bbox_resize(c, n, x, y, z, cells, md);
    a.md = md;
    readback = true;
  }
  if (qa) {  // :244-245
    a.q = st.in(qa, size_t(n));
    if (n > 0) {
      device_sum_to(c, a.q, n, &md->charge);
      readback = true;
    } else {
      m->charge = 0.0;
    }
  } else if (total_charge) {  // :246-248
    m->charge = *total_charge;
    a.q_each = *total_charge / double(n);
  }
  if (readback) {
    OSC_HIP(hipMemcpyAsync(c->mesh_host, md, sizeof(MeshDev), hipMemcpyDeviceToHost, c->stream));
    OSC_HIP(hipEventRecord(c->mesh_ev, c->stream));
  }
  for (int d = 0; d < 3; ++d) {
    a.nlo[d] = m->nlo[d];
    a.nn[d] = nn[d];
    a.mn[d] = m->min[d];
    a.delta[d] = m->delta[d];
  }
  if (qa || total_charge) deposit_cic(c, a, rho, keep);
  if (deferred) *deferred = readback;
  if (readback && !deferred) finish_deposit_mesh(c, resize_mesh, qa != nullptr, m);
  // upstream resizes the mesh before it finds no charge (:249-252)
  if (!qa && !total_charge)
    return fail(c, OSC_ERR_NO_CHARGE, "Error, must specify qa or total_charge");
  return OSC_OK;
}
OSC_API int osc_deposit_particles(osc_ctx* c, int64_t n, const double* xa, const double* ya,
                                  const double* za, const double* qa, const double* total_charge,
                                  int resize_mesh, osc_mesh* m, double* rho, int64_t* n_dropped) {
  return guarded(c, [&]() -> int {
    if (!m || !rho || n < 0 || (n > 0 && (!xa || !ya || !za)))
      return fail(c, OSC_ERR_ARG, "deposit_particles: bad arguments");
    int nn[3];
    mesh_dims(m, nn);
    if (!dims_ok(nn) || nn[0] < 2 || nn[1] < 2 || nn[2] < 2)
      return fail(c, OSC_ERR_ARG, "deposit_particles: mesh needs >= 2 nodes per axis");
    Staging st(c);
    const double* x = st.in(xa, size_t(n));
    const double* y = st.in(ya, size_t(n));
    const double* z = st.in(za, size_t(n));
    // upstream resizes the mesh, then returns without touching rho (:249-252)
    double* r = (qa || total_charge) ? st.out(rho, size_t(prod3(nn))) : nullptr;
    // a large deposit keeps its cell order and the content hash of the points
    // for an interpolate_field at the same points (the usual next call but one)
    // the previous kept order goes back to the pool first, so that a new
    // large deposit never holds two copies
    c->kept_ok = false;
    c->kept = SortedParticles{};
    const bool keep = r && n >= kGatherSortMin;
    unsigned long long* hash = reinterpret_cast<unsigned long long*>(
        static_cast<char*>(c->counters) + kHashDeposit);
    const int rc = deposit_staged(c, st, n, x, y, z, qa, total_charge, resize_mesh, m, r,
                                  keep ? &c->kept : nullptr, keep ? hash : nullptr);
    c->kept_ok = keep && rc == OSC_OK;
    if (rc == OSC_OK && n_dropped) *n_dropped = read_counter(c, 0);
    st.finish();
    return rc;
  });
}

OSC_API int osc_space_charge_step(osc_ctx* c, int64_t n, const double* xa, const double* ya,
                                  const double* za, const double* qa, const double* total_charge,
                                  int resize_mesh, const double* offset, int at_cathode,
                                  int calc_bfield, osc_mesh* m, double* rho, double* efield,
                                  double* bfield, double* E, double* B, int64_t* n_dropped,
                                  int64_t* n_clamped) {
  return guarded(c, [&]() -> int {
    if (!m || !rho || !efield || n < 1 || !xa || !ya || !za || (calc_bfield && !bfield) ||
        (B && !calc_bfield))
      return fail(c, OSC_ERR_ARG, "space_charge_step: bad arguments");
    if (!qa && !total_charge)
      return fail(c, OSC_ERR_NO_CHARGE, "Error, must specify qa or total_charge");
    int nn[3], L[3];
    mesh_dims(m, nn);
    if (nn[0] < 2 || nn[1] < 2 || nn[2] < 2)
      return fail(c, OSC_ERR_ARG, "space_charge_step: mesh needs >= 2 nodes per axis");
    for (int a = 0; a < 3; ++a) L[a] = 2 * nn[a];
    if (int rc = check_len(c, L)) return rc;
    const int64_t nt = prod3(nn);
    Staging st(c);
    const double* x = st.in(xa, size_t(n));
    const double* y = st.in(ya, size_t(n));
    const double* z = st.in(za, size_t(n));
    double* r = st.out(rho, size_t(nt));
    double* e = st.out(efield, size_t(3 * nt));
    double* b = calc_bfield ? st.out(bfield, size_t(3 * nt)) : nullptr;
    double* Eo = st.out(E, size_t(3 * n));
    double* Bo = st.out(B, size_t(3 * n));
    SortedParticles sp;
    // a mesh that follows the bunch is known only on the device until the
    // deposit is done: its solve is replayed from a graph that builds the
    // Green functions from the device mesh, with no host round trip.  A
    // fixed mesh runs eagerly, its Green spectra cached for the next step
    // of the same geometry.
    const bool zero_off = !offset || (offset[0] == 0.0 && offset[1] == 0.0 && offset[2] == 0.0);
    const bool graphable = c->graphs && !c->prof && zero_off && resize_mesh;
    SolveGraph* g = graphable ? step_graph(c, nn, m->gamma, at_cathode, calc_bfield) : nullptr;
    if (g) {
      // the captured solve: deposit into its rho, the graph (reading the
      // mesh on the device), the mesh for the host while it runs
      bool deferred = false;
      int rc = deposit_staged(c, st, n, x, y, z, qa, total_charge, resize_mesh, m,
                              g->rho.as<double>(), &sp, nullptr, &deferred,
                              c->prec == OSC_PREC_FP32);
      if (rc != OSC_OK) return rc;
      OSC_HIP(hipGraphLaunch(g->exec, c->stream));
      ++g->replays;
      ++c->graph_replays;
      if (deferred) finish_deposit_mesh(c, resize_mesh != 0, qa != nullptr, m);
      OSC_HIP(hipMemcpyAsync(r, g->rho.p, sizeof(double) * nt, hipMemcpyDeviceToDevice, c->stream));
      OSC_HIP(hipMemcpyAsync(e, g->ef.p, sizeof(double) * 3 * nt, hipMemcpyDeviceToDevice,
                             c->stream));
      if (b)
        OSC_HIP(hipMemcpyAsync(b, g->bf.p, sizeof(double) * 3 * nt, hipMemcpyDeviceToDevice,
                               c->stream));
      // the gather reads the graph's fields (the copies run beside it in order)
      e = g->ef.as<double>();
      b = calc_bfield ? g->bf.as<double>() : nullptr;
    } else {
      // an FP32 context deposits in FP32 here (the step's own gather uses the
      // weights); the standalone deposit_particles stays FP64
      int rc = deposit_staged(c, st, n, x, y, z, qa, total_charge, resize_mesh, m, r, &sp,
                              nullptr, nullptr, c->prec == OSC_PREC_FP32);
      if (rc != OSC_OK) return rc;
      space_charge_3d(c, *m, r, offset, at_cathode, calc_bfield, e, b);
      if (graphable) note_step_shape(c, nn, m->gamma, at_cathode, calc_bfield, *m);
    }
    if (Eo || Bo) {
      InterpArgs ia{};
      for (int d = 0; d < 3; ++d) {
        ia.nlo[d] = m->nlo[d];
        ia.nhi[d] = m->nhi[d];
        ia.nn[d] = nn[d];
        ia.mn[d] = m->min[d];
        ia.delta[d] = m->delta[d];
      }
      interpolate_sorted(c, ia, e, b, sp, Eo, Bo, nullptr, nullptr, x, y, z);
    }
    if (n_dropped) *n_dropped = read_counter(c, 0);
    if (n_clamped) *n_clamped = read_counter(c, 2);
    st.finish();
    return OSC_OK;
  });
}

OSC_API int osc_get_mesh_quantities(osc_ctx* c, int64_t n, const double* xa, const double* ya,
                                    const double* za, const int32_t nlo[3], const int32_t nhi[3],
                                    int isetzmin, double delta[3], double umin[3]) {
  return guarded(c, [&]() -> int {
    if (!nlo || !nhi || !delta || !umin || n < 0) return fail(c, OSC_ERR_ARG, "bad arguments");
    double sml[3], big[3];
    if (n > 0) {
      Staging st(c);
      bbox(c, n, st.in(xa, size_t(n)), st.in(ya, size_t(n)), st.in(za, size_t(n)), sml, big);
      st.finish();
    } else {
      for (int a = 0; a < 3; ++a) {  // getbeamboundingbox sentinel (test_mod.f90:379-380)
        sml[a] = 9999999.;
        big[a] = -9999999.;
      }
    }
    const double eps = 1.e-13;  // test_mod.f90:304
    for (int a = 0; a < 3; ++a) {
      big[a] = big[a] * (1.0 + std::copysign(1.0, big[a]) * eps);
      sml[a] = sml[a] * (1.0 - std::copysign(1.0, sml[a]) * eps);
    }
    for (int a = 0; a < 3; ++a) {
      const int na = nhi[a] - nlo[a] + 1;
      delta[a] = (big[a] - sml[a]) / double(na - 3);
      umin[a] = sml[a] - delta[a];
    }
    if (isetzmin == 1) {
      const int nz = nhi[2] - nlo[2] + 1;
      const double zsml = 1.e-9;
      delta[2] = (big[2] - zsml) / double(nz - 1);
      umin[2] = zsml;
    }
    return OSC_OK;
  });
}

OSC_API int osc_depose_rho_scalar(osc_ctx* c, int64_t n, const double* xa, const double* ya,
                                  const double* za, const int32_t nlo[3], const int32_t nhi[3],
                                  double chrgpermacro, const double delta[3],
                                  const double umin[3], double* rho, int64_t* ifail) {
  return guarded(c, [&]() -> int {
    if (!nlo || !nhi || !delta || !umin || !rho || n < 0)
      return fail(c, OSC_ERR_ARG, "bad arguments");
    int nn[3] = {nhi[0] - nlo[0] + 1, nhi[1] - nlo[1] + 1, nhi[2] - nlo[2] + 1};
    if (nn[0] < 2 || nn[1] < 2 || nn[2] < 2) return fail(c, OSC_ERR_ARG, "mesh too small");
    Staging st(c);
    DepositArgs a{};
    a.n = n;
    a.x = st.in(xa, size_t(n));
    a.y = st.in(ya, size_t(n));
    a.z = st.in(za, size_t(n));
    a.harness = 1;
    a.scale_after = chrgpermacro;
    for (int d = 0; d < 3; ++d) {
      a.nlo[d] = nlo[d];
      a.nn[d] = nn[d];
      a.mn[d] = umin[d];
      a.delta[d] = delta[d];
    }
    double* r = st.out(rho, size_t(prod3(nn)));
    deposit_cic(c, a, r);
    if (ifail) *ifail = read_counter(c, 0);
    st.finish();
    return OSC_OK;
  });
}

// ---------------------------------------------------------------------------
// free space (offset-grid IGF)
// ---------------------------------------------------------------------------
OSC_API int osc_space_charge_3d(osc_ctx* c, const osc_mesh* m, const double* rho,
                                const double* offset, int at_cathode, int calc_bfield,
                                double* efield, double* bfield) {
  return guarded(c, [&]() -> int {
    if (!m || !rho || !efield || (calc_bfield && !bfield))
      return fail(c, OSC_ERR_ARG, "space_charge_3d: bad arguments");
    int n[3], L[3];
    mesh_dims(m, n);
    if (!dims_ok(n)) return fail(c, OSC_ERR_ARG, "bad mesh");
    for (int a = 0; a < 3; ++a) L[a] = 2 * n[a];
    if (int rc = check_len(c, L)) return rc;
    const int64_t nt = prod3(n);
    Staging st(c);
    const double* r = st.in(rho, size_t(nt));
    double* e = st.out(efield, size_t(3 * nt));
    double* b = calc_bfield ? st.out(bfield, size_t(3 * nt)) : nullptr;
    space_charge_3d(c, *m, r, offset, at_cathode, calc_bfield, e, b);
    st.finish();
    return OSC_OK;
  });
}

OSC_API int osc_freespace_solver2(osc_ctx* c, int nx, int ny, int nz, const double* rho,
                                  double gamma, const double delta[3], double* efield,
                                  double* phi, const double* offset) {
  return guarded(c, [&]() -> int {
    if (!rho || !delta || nx < 1 || ny < 1 || nz < 1)
      return fail(c, OSC_ERR_ARG, "osc_freespace_solver2: bad arguments");
    const int n[3] = {nx, ny, nz}, L[3] = {2 * nx, 2 * ny, 2 * nz};
    if (int rc = check_len(c, L)) return rc;
    const int64_t nt = prod3(n);
    Staging st(c);
    const double* r = st.in(rho, size_t(nt));
    double* e = st.out(efield, size_t(3 * nt));
    double* p = st.out(phi, size_t(nt));
    solver2(c, n, r, gamma, delta, e, p, offset);
    st.finish();
    return OSC_OK;
  });
}

OSC_API int osc_get_cgrn_freespace(osc_ctx* c, double* cgrn, int isize, int jsize, int ksize,
                                   const double delta[3], double gamma, int icomp,
                                   const double* offset) {
  return guarded(c, [&]() -> int {
    if (!cgrn || !delta || isize < 1 || jsize < 1 || ksize < 1 || icomp < 0 || icomp > 3)
      return fail(c, OSC_ERR_ARG, "osc_get_cgrn_freespace: bad arguments");
    GreenSpec s = make_green_spec();
    s.kind = GK_OFFSET;
    s.icomp = icomp;
    s.L[0] = isize;
    s.L[1] = jsize;
    s.L[2] = ksize;
    for (int a = 0; a < 3; ++a) {
      s.delta[a] = delta[a];
      s.offset[a] = offset ? offset[a] : 0.0;
    }
    s.gamma = gamma;
    const int64_t N = int64_t(isize) * jsize * ksize;
    Staging st(c);
    double* out = st.out(cgrn, size_t(2 * N));
    DevBuf g(c->pool, sizeof(double) * N);
    build_green_real(c, s, g.as<double>());
    hipLaunchKernelGGL(k_real_to_cplx, dim3(unsigned(std::min<int64_t>((N + 255) / 256, 1 << 19))),
                       dim3(256), 0, c->stream, g.as<const double>(), out, N);
    OSC_HIP(hipGetLastError());
    st.finish();
    return OSC_OK;
  });
}

OSC_API int osc_ccfft3d(osc_ctx* c, const double* a, double* b, const int32_t idir[3], int n1,
                        int n2, int n3) {
  return guarded(c, [&]() -> int {
    if (!a || !b || !idir || n1 < 1 || n2 < 1 || n3 < 1)
      return fail(c, OSC_ERR_ARG, "ccfft3d: bad arguments");
    const bool ident = idir[0] == 0 && idir[1] == 0 && idir[2] == 0;
    const int nd[3] = {n1, n2, n3};
    if (!ident)
      for (int d = 0; d < 3; ++d)
        if ((std::abs(idir[d]) != 1 && nd[d] > 1) || !fft_len_supported(nd[d]))
          return fail(c, OSC_ERR_FFT_LENGTH,
                      "Error return from FFT package due to transform length.");
    const int64_t N = int64_t(n1) * n2 * n3;
    Staging st(c);
    const double* ain = st.in(a, size_t(2 * N));
    double* bo = st.out(b, size_t(2 * N));
    if (ain != bo)
      OSC_HIP(hipMemcpyAsync(bo, ain, sizeof(double) * 2 * N, hipMemcpyDeviceToDevice, c->stream));
    if (!ident) {
      cpx<double>* v = reinterpret_cast<cpx<double>*>(bo);
      if (n1 > 1) {  // x: lines indexed by (y, z)
        AxisIO x{};
        x.s_b = x.d_b = n1;
        x.s_p = x.d_p = 1;
        x.s_o1 = x.d_o1 = int64_t(n1) * n2;
        x.nb = n2;
        x.n_in = n1;
        x.n_out = n1;
        launch_axis<double>(n1, idir[0] > 0 ? 1 : -1, v, v, x, n3, 1, twiddles<double>(c, n1),
                            c->stream);
      }
      if (n2 > 1) {
        AxisIO y{};
        y.s_b = y.d_b = 1;
        y.s_p = y.d_p = n1;
        y.s_o1 = y.d_o1 = int64_t(n1) * n2;
        y.nb = n1;
        y.n_in = n2;
        y.n_out = n2;
        launch_axis<double>(n2, idir[1] > 0 ? 1 : -1, v, v, y, n3, 1, twiddles<double>(c, n2),
                            c->stream);
      }
      if (n3 > 1) {
        AxisIO z{};
        z.s_b = z.d_b = 1;
        z.s_p = z.d_p = int64_t(n1) * n2;
        z.s_o1 = z.d_o1 = n1;
        z.nb = n1;
        z.n_in = n3;
        z.n_out = n3;
        launch_axis<double>(n3, idir[2] > 0 ? 1 : -1, v, v, z, n2, 1, twiddles<double>(c, n3),
                            c->stream);
      }
    }
    st.finish();
    return OSC_OK;
  });
}

// ---------------------------------------------------------------------------
// legacy solvers
// ---------------------------------------------------------------------------
static int legacy_lengths(osc_ctx* c, const osc_mesh* m) {
  int L[3];
  for (int a = 0; a < 3; ++a) L[a] = 2 * (m->nhi[a] - m->nlo[a]) + m->npad[a] + 1;
  return check_len(c, L);
}

OSC_API int osc_space_charge_freespace(osc_ctx* c, const osc_mesh* m, const double* rho,
                                       int direct, int igf, double* phi, double* efield,
                                       double* bfield) {
  return guarded(c, [&]() -> int {
    if (!m || !rho || !efield || !bfield || (!direct && !phi))
      return fail(c, OSC_ERR_ARG, "space_charge_freespace: bad arguments");
    if (igf == 0 && direct)
      return fail(c, OSC_ERR_DIRECT_NEEDS_IGF,
                  "Error: Direct field calc must use integrated Green function");
    int n[3];
    mesh_dims(m, n);
    if (!dims_ok(n)) return fail(c, OSC_ERR_ARG, "bad mesh");
    if (int rc = legacy_lengths(c, m)) return rc;
    const int64_t nt = prod3(n);
    Staging st(c);
    const double* r = st.in(rho, size_t(nt));
    double* p = direct ? nullptr : st.out(phi, size_t(nt));
    double* e = st.out(efield, size_t(3 * nt));
    double* b = st.out(bfield, size_t(3 * nt));
    int rc = legacy_free(c, *m, r, direct ? 1 : 0, igf ? 1 : 0, p, e, b);
    st.finish();
    return rc;
  });
}

OSC_API int osc_space_charge_cathodeimages(osc_ctx* c, const osc_mesh* m, double* rho, int direct,
                                           int igf, int image_method, double* phi, double* efield,
                                           double* bfield) {
  return guarded(c, [&]() -> int {
    if (!m || !rho || !efield || !bfield || (!direct && !phi))
      return fail(c, OSC_ERR_ARG, "space_charge_cathodeimages: bad arguments");
    if (igf == 0 && direct)
      return fail(c, OSC_ERR_DIRECT_NEEDS_IGF,
                  "Error: Direct field calc must use integrated Green function");
    const int method = image_method == 2 ? 2 : 1;  // :97-98
    int n[3];
    mesh_dims(m, n);
    if (!dims_ok(n)) return fail(c, OSC_ERR_ARG, "bad mesh");
    if (int rc = legacy_lengths(c, m)) return rc;
    const int64_t nt = prod3(n);
    Staging st(c);
    double* r = method == 2 ? st.out(rho, size_t(nt), true) : const_cast<double*>(st.in(rho, size_t(nt)));
    double* p = direct ? nullptr : st.out(phi, size_t(nt));
    double* e = st.out(efield, size_t(3 * nt));
    double* b = st.out(bfield, size_t(3 * nt));
    int rc = cathode_legacy(c, *m, r, direct ? 1 : 0, igf ? 1 : 0, method, p, e, b);
    st.finish();
    return rc;
  });
}

template <class Fill>
static int upstream_green(osc_ctx* c, double* cgrn, const GreenSpec& s, const int sgn[3]) {
  const SpecDims d = SpecDims::of(s.L);
  if (int rc = check_len(c, s.L)) return rc;
  const int64_t N = int64_t(s.L[0]) * s.L[1] * s.L[2];
  DevBuf g(c->pool, sizeof(double) * N), sp(c->pool, sizeof(cpx<double>) * d.size());
  build_green_real(c, s, g.as<double>());
  spectrum_forward_full(c, g.as<const double>(), sp.as<cpx<double>>(), d);
  const bool dev = is_device_ptr(cgrn);
  std::vector<double> h(dev ? size_t(2 * N) : 0);
  double* dst = dev ? h.data() : cgrn;
  half_to_upstream(c, sp.as<const cpx<double>>(), d, sgn, dst);
  if (dev) {
    OSC_HIP(hipMemcpyAsync(cgrn, h.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice, c->stream));
    OSC_HIP(hipStreamSynchronize(c->stream));
  }
  return OSC_OK;
}

static GreenSpec legacy_spec_c(const int32_t nlo[3], const int32_t nhi[3], const int32_t npad[3],
                               const double delta[3], double gamma, int icomp, int igf) {
  GreenSpec s = make_green_spec();
  s.kind = GK_LEGACY;
  s.icomp = icomp;
  s.igf = igf;
  for (int a = 0; a < 3; ++a) {
    s.lo[a] = nlo[a] - nhi[a];
    s.L[a] = (nhi[a] - nlo[a] + npad[a]) - s.lo[a] + 1;
    s.shift[a] = s.L[a] / 2 - (npad[a] + 1) / 2;
    s.delta[a] = delta[a];
  }
  s.gamma = gamma;
  return s;
}

OSC_API int osc_getgrnfree(osc_ctx* c, double* cgrn, const int32_t nlo[3], const int32_t nhi[3],
                           const int32_t npad[3], const double delta[3], double gamma, int icomp,
                           int igfflag) {
  return guarded(c, [&]() -> int {
    if (!cgrn || !nlo || !nhi || !npad || !delta) return fail(c, OSC_ERR_ARG, "bad arguments");
    const int sgn[3] = {1, 1, 1};
    return upstream_green<int>(c, cgrn, legacy_spec_c(nlo, nhi, npad, delta, gamma, icomp, igfflag),
                               sgn);
  });
}

OSC_API int osc_getgrnimageshift(osc_ctx* c, double* cgrn, const int32_t nlo[3],
                                 const int32_t nhi[3], const int32_t npad[3],
                                 const double delta[3], double gamma, double zshft, int icomp,
                                 int igfflag) {
  return guarded(c, [&]() -> int {
    if (!cgrn || !nlo || !nhi || !npad || !delta) return fail(c, OSC_ERR_ARG, "bad arguments");
    GreenSpec s = legacy_spec_c(nlo, nhi, npad, delta, gamma, icomp, igfflag);
    s.kind = GK_IMG_SHIFT;
    s.zshft = zshft;
    const int sgn[3] = {1, 1, 1};
    return upstream_green<int>(c, cgrn, s, sgn);
  });
}

OSC_API int osc_getgrnimageconvcorr(osc_ctx* c, double* cgrn, const int32_t nlo[3],
                                    const int32_t nhi[3], const int32_t npad[3],
                                    const double umin[3], const double delta[3], double gamma,
                                    int icomp, int igfflag) {
  return guarded(c, [&]() -> int {
    if (!cgrn || !nlo || !nhi || !npad || !delta || !umin)
      return fail(c, OSC_ERR_ARG, "bad arguments");
    GreenSpec s = legacy_spec_c(nlo, nhi, npad, delta, gamma, icomp, igfflag);
    s.kind = GK_IMG_CONVCORR;
    s.lo[2] = 2 * nlo[2];
    s.L[2] = 2 * nhi[2] + npad[2] - s.lo[2] + 1;
    s.nlo3 = nlo[2];
    s.umin[2] = umin[2];
    const int sgn[3] = {1, 1, -1};
    return upstream_green<int>(c, cgrn, s, sgn);
  });
}

// ---------------------------------------------------------------------------
// rectangular pipe
// ---------------------------------------------------------------------------
OSC_API int osc_alloc_rectpipe_arrays(osc_ctx* c, const int32_t nlo[3], const int32_t nhi[3],
                                      const int32_t npad[3]) {
  return guarded(c, [&]() -> int {
    if (!nlo || !nhi || !npad) return fail(c, OSC_ERR_ARG, "bad arguments");
    PipeState& p = c->pipe;
    bool same = p.alloc;
    for (int a = 0; a < 3; ++a) {
      const int L = 2 * (nhi[a] - nlo[a]) + npad[a] + 1;  // core:888-896
      same = same && p.nlo[a] == nlo[a] && p.nhi[a] == nhi[a] && p.npad[a] == npad[a];
      p.nlo[a] = nlo[a];
      p.nhi[a] = nhi[a];
      p.npad[a] = npad[a];
      p.L[a] = L;
    }
    if (int rc = check_len(c, p.L)) return rc;
    if (!same) {
      p.built = false;
      for (int t = 0; t < 4; ++t) {
        p.spec64[t].reset();
        p.spec32[t].reset();
      }
    }
    p.alloc = true;
    return OSC_OK;
  });
}

OSC_API int osc_getgrnpipe(osc_ctx* c, double gamma, double apipe, double bpipe,
                           const double delta[3], const double umin[3], const int32_t npad[3]) {
  return guarded(c, [&]() -> int {
    if (!delta || !umin || !npad) return fail(c, OSC_ERR_ARG, "bad arguments");
    if (!c->pipe.alloc)
      return fail(c, OSC_ERR_NO_GREEN, "osc_alloc_rectpipe_arrays has not been called");
    const int np[3] = {npad[0], npad[1], npad[2]};
    pipe_build(c, gamma, apipe, bpipe, delta, umin, np);
    OSC_HIP(hipStreamSynchronize(c->stream));
    return OSC_OK;
  });
}

OSC_API int osc_write_rectpipe_grn(osc_ctx* c, const char* path, double apipe, double bpipe,
                                   const double delta[3], const double umin[3],
                                   const double umax[3], const int32_t nlo[3],
                                   const int32_t nhi[3], double gamma) {
  return guarded(c, [&]() -> int {
    if (!path || !delta || !umin || !umax || !nlo || !nhi)
      return fail(c, OSC_ERR_ARG, "bad arguments");
    PipeState& p = c->pipe;
    if (!p.built) return fail(c, OSC_ERR_NO_GREEN, "pipe Green functions not built");
    const SpecDims d = SpecDims::of(p.L);
    const int64_t N = int64_t(p.L[0]) * p.L[1] * p.L[2];
    const int64_t rec2 = 4 * N * 16;
    if (rec2 > INT32_MAX) return fail(c, OSC_ERR_IO, "Green-function record exceeds 2 GiB");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, OSC_ERR_IO, std::string("cannot open ") + path);
    // record 1 (core:864-865)
    const int32_t r1 = 11 * 8 + 9 * 4 + 8;
    const double rv[11] = {apipe,   bpipe,   delta[0], delta[1], delta[2], umin[0],
                           umin[1], umin[2], umax[0],  umax[1],  umax[2]};
    const int32_t iv[9] = {nlo[0], nlo[1], nlo[2], nhi[0], nhi[1], nhi[2], p.L[0], p.L[1], p.L[2]};
    bool ok = std::fwrite(&r1, 4, 1, f) == 1 && std::fwrite(rv, 8, 11, f) == 11 &&
              std::fwrite(iv, 4, 9, f) == 9 && std::fwrite(&gamma, 8, 1, f) == 1 &&
              std::fwrite(&r1, 4, 1, f) == 1;
    // record 2: cgrn1..4 in the upstream transform convention (core:866)
    const int32_t r2 = int32_t(rec2);
    ok = ok && std::fwrite(&r2, 4, 1, f) == 1;
    std::vector<double> full(size_t(2 * N));
    for (int t = 0; ok && t < 4; ++t) {
      half_to_upstream(c, p.spec64[t].as<const cpx<double>>(), d, kPipeSigns[t], full.data());
      ok = std::fwrite(full.data(), 16, size_t(N), f) == size_t(N);
    }
    ok = ok && std::fwrite(&r2, 4, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? OSC_OK : fail(c, OSC_ERR_IO, std::string("write failed: ") + path);
  });
}

OSC_API int osc_read_rectpipe_grn(osc_ctx* c, const char* path, double* hdr_real,
                                  int32_t* hdr_int) {
  return guarded(c, [&]() -> int {
    if (!path) return fail(c, OSC_ERR_ARG, "bad arguments");
    PipeState& p = c->pipe;
    if (!p.alloc) return fail(c, OSC_ERR_NO_GREEN, "osc_alloc_rectpipe_arrays has not been called");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, OSC_ERR_IO, std::string("cannot open ") + path);
    int32_t m1 = 0, m2 = 0;
    double rv[12];
    int32_t iv[9];
    bool ok = std::fread(&m1, 4, 1, f) == 1 && m1 == 132 && std::fread(rv, 8, 11, f) == 11 &&
              std::fread(iv, 4, 9, f) == 9 && std::fread(rv + 11, 8, 1, f) == 1 &&
              std::fread(&m2, 4, 1, f) == 1 && m2 == 132;
    if (ok)
      for (int a = 0; a < 3; ++a) ok = ok && iv[6 + a] == p.L[a];
    const int64_t N = int64_t(p.L[0]) * p.L[1] * p.L[2];
    ok = ok && std::fread(&m1, 4, 1, f) == 1 && int64_t(m1) == 4 * N * 16;
    const SpecDims d = SpecDims::of(p.L);
    std::vector<double> full(size_t(2 * N));
    for (int t = 0; ok && t < 4; ++t) {
      ok = std::fread(full.data(), 16, size_t(N), f) == size_t(N);
      if (!ok) break;
      if (!p.spec64[t].p) p.spec64[t] = DevBuf(c->pool, sizeof(cpx<double>) * d.size());
      upstream_to_half(c, full.data(), d, kPipeSigns[t], p.spec64[t].as<cpx<double>>());
      p.spec32[t].reset();
    }
    std::fclose(f);
    if (!ok) return fail(c, OSC_ERR_IO, std::string("bad rectpipe Green file: ") + path);
    p.built = true;
    p.a = rv[0];
    p.b = rv[1];
    p.gamma = rv[11];
    if (hdr_real) std::memcpy(hdr_real, rv, sizeof(rv));
    if (hdr_int) std::memcpy(hdr_int, iv, sizeof(iv));
    return OSC_OK;
  });
}

OSC_API int osc_space_charge_rectpipe(osc_ctx* c, const osc_mesh* m, double apipe, double bpipe,
                                      int direct, int igf, const double* rho, double* phi,
                                      double* efield, double* bfield) {
  (void)apipe;
  (void)bpipe;
  (void)igf;
  return guarded(c, [&]() -> int {
    if (!m || !rho || !efield || !bfield || (!direct && !phi))
      return fail(c, OSC_ERR_ARG, "space_charge_rectpipe: bad arguments");
    int n[3];
    mesh_dims(m, n);
    if (!dims_ok(n)) return fail(c, OSC_ERR_ARG, "bad mesh");
    if (!c->pipe.built)
      return fail(c, OSC_ERR_NO_GREEN,
                  "something wrong. rectpipe green functions should have been computed during "
                  "initialization");
    const int64_t nt = prod3(n);
    Staging st(c);
    const double* r = st.in(rho, size_t(nt));
    double* p = direct ? nullptr : st.out(phi, size_t(nt));
    double* e = st.out(efield, size_t(3 * nt), direct != 0);
    double* b = st.out(bfield, size_t(3 * nt));
    const int rc = rectpipe(c, *m, direct ? 1 : 0, r, p, e, b);
    st.finish();
    if (rc == OSC_ERR_PIPE_DIRECT)
      return fail(c, rc, "**********ERROR********* rectpipe implementation presently requires "
                         "directfieldcalc=0");
    return rc;
  });
}

// ---------------------------------------------------------------------------
// gather
// ---------------------------------------------------------------------------
OSC_API int osc_interpolate_field(osc_ctx* c, const osc_mesh* m, const double* efield,
                                  const double* bfield, int64_t n, const double* x,
                                  const double* y, const double* z, double* E, double* B,
                                  int8_t* clamped, int64_t* n_clamped) {
  return guarded(c, [&]() -> int {
    if (!m || n < 0 || (n > 0 && (!x || !y || !z)))
      return fail(c, OSC_ERR_ARG, "interpolate_field: bad arguments");
    InterpArgs a{};
    for (int d = 0; d < 3; ++d) {
      a.nlo[d] = m->nlo[d];
      a.nhi[d] = m->nhi[d];
      a.nn[d] = m->nhi[d] - m->nlo[d] + 1;
      a.mn[d] = m->min[d];
      a.delta[d] = m->delta[d];
    }
    const int64_t nt = int64_t(a.nn[0]) * a.nn[1] * a.nn[2];
    Staging st(c);
    const bool wantE = efield && E, wantB = bfield && B;
    const double* ef = wantE ? st.in(efield, size_t(3 * nt)) : nullptr;
    const double* bf = wantB ? st.in(bfield, size_t(3 * nt)) : nullptr;
    const double* xd = st.in(x, size_t(n));
    const double* yd = st.in(y, size_t(n));
    const double* zd = st.in(z, size_t(n));
    double* Eo = wantE ? st.out(E, size_t(3 * n)) : nullptr;
    double* Bo = wantB ? st.out(B, size_t(3 * n)) : nullptr;
    int8_t* cl = st.out(clamped, size_t(n));
    interpolate(c, a, ef, bf, n, xd, yd, zd, Eo, Bo, cl);
    st.finish();
    if (n_clamped) {
      uint32_t v = 0;
      OSC_HIP(hipMemcpyAsync(&v, static_cast<uint32_t*>(c->counters) + 2, sizeof(v),
                             hipMemcpyDeviceToHost, c->stream));
      OSC_HIP(hipStreamSynchronize(c->stream));
      *n_clamped = v;
    }
    return OSC_OK;
  });
}
