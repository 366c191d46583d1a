// Context, caches and internal entry points of the library.
#pragma once
#include <list>
#include <map>
#include <memory>
#include <unordered_set>

#include "fft_launch.h"

namespace osc {

// dimensions of a doubled-grid (or period-L) half spectrum
struct SpecDims {
  int L[3];
  int KX;   // L0/2 + 1
  int KXP;  // padded row length
  static SpecDims of(const int L_[3]) {
    SpecDims d;
    for (int i = 0; i < 3; ++i) d.L[i] = L_[i];
    d.KX = L_[0] / 2 + 1;
    d.KXP = (d.KX + 15) / 16 * 16;
    return d;
  }
  int64_t size() const { return int64_t(KXP) * L[1] * L[2]; }
};

// Everything that determines one real-space Green array (all fields are part
// of the cache key; construct with make_green_spec so padding is zero).
enum GreenKind : int {
  GK_OFFSET = 0,      // osc_get_cgrn_freespace (open_spacecharge_mod.f90:633-707)
  GK_LEGACY = 1,      // osc_getgrnfree loop (core:454-497)
  GK_IMG_SHIFT = 2,   // osc_getgrnimageshift loop (core:1135-1178)
  GK_IMG_CONVCORR = 3,// osc_getgrnimageconvcorr loop (core:1185-1230)
  GK_PIPE = 4         // osc_getgrnpipe loops (core:652-713)
};
struct GreenSpec {
  int32_t kind, icomp, igf, term;
  int32_t L[3];
  int32_t lo[3];      // lower index bound of the Green array (legacy / pipe)
  int32_t shift[3];   // iperiod/2-(ipad+1)/2 (legacy / pipe)
  int32_t nlo3;
  int32_t out_shift_z;  // circular shift applied to the output z index
  int32_t zero_raw;     // offset kind: zero the raw last planes (never extracted)
  double delta[3], gamma, offset[3], zshft, umin[3], a, b;
};
GreenSpec make_green_spec();

// A Green function as conv_solve consumes it: a half spectrum, or (zr) an
// array transformed along x and y only, whose z transform conv_solve takes
// per line; then zmh > 0 when only z < zmh is stored, the lines being
// mirror-symmetric about L/2 - 1 with sign zms.
struct GreenRef {
  const void* g = nullptr;
  int zr = 0, zmh = 0, zms = 0;
};

class GreenCache {
 public:
  struct Entry {
    DevBuf f64, f32;
    int zr = 0, zmh = 0, zms = 0;
    size_t bytes = 0;
    uint64_t last = 0;
  };
  bool enable = true;
  size_t max_bytes = size_t(8) << 30;
  size_t bytes = 0;
  uint64_t tick = 0;
  std::unordered_map<std::string, std::unique_ptr<Entry>> map;
  uint64_t hits = 0, misses = 0;
  // entries used since pin() (last >= pin_from) are not evicted until the
  // matching unpin(): a solve holds references to all its spectra until its
  // kernels are queued, and an evicted buffer could be reused for the next
  // spectrum before that
  uint64_t pin_from = ~uint64_t(0);
  int pin_depth = 0;
  void pin() {
    if (pin_depth++ == 0) pin_from = tick + 1;
  }
  void unpin() {
    if (--pin_depth == 0) pin_from = ~uint64_t(0);
  }
  // admission: a spectrum is kept from its key's second miss on.  One whose
  // key is new (a bunch whose spacing does not recur) is built into a pooled
  // buffer that goes back to the pool after its solve, where the next request
  // of that size finds it, instead of a new device allocation held by the
  // cache; `seen` holds the keys missed once (at most kSeenMax, then
  // cleared).  admit_second = false keeps every spectrum built
  static constexpr size_t kSeenMax = 4096;
  bool admit_second = true;
  std::unordered_set<std::string> seen;
  bool admit(const std::string& key) {
    if (!admit_second || seen.erase(key)) return true;
    if (seen.size() >= kSeenMax) seen.clear();
    seen.insert(key);
    return false;
  }
  void clear() {
    map.clear();
    seen.clear();
    bytes = 0;
  }
  void evict_to(size_t budget);
};

// pipe Green spectra (the module globals cgrn1..4 of core:22-23)
struct PipeState {
  bool alloc = false, built = false;
  int nlo[3], nhi[3], npad[3], L[3];
  double a = 0, b = 0, gamma = 0, delta[3] = {0, 0, 0}, umin[3] = {0, 0, 0};
  DevBuf spec64[4], spec32[4];
};

// timed stages (osc_stage_name)
enum Stage : int {
  ST_DEP_BBOX = 0,
  ST_DEP_KEYS,
  ST_DEP_SORT,
  ST_DEP_CELLS,
  ST_DEP_NODES,
  ST_GREEN_EVAL,
  ST_GREEN_FFT,
  ST_FWD_X,
  ST_FWD_Y,
  ST_FWD_Z,
  ST_INV_ZMUL,
  ST_INV_Y,
  ST_INV_X,
  ST_GATHER,
  ST_FD_B,
  ST_OTHER,
  ST_GATHER_PACK,
  ST_GATHER_SORT,
  ST_GATHER_OUT,
  ST_DEP_FIX,
  ST_COUNT
};

struct ProfRec {
  int stage;
  hipEvent_t a, b;
};


// a gather of at least this many points runs in cell order
constexpr int64_t kGatherSortMin = int64_t(1) << 16;
// byte offsets in the context's counter area of the particle-content hashes:
// of the last osc_deposit_particles, and of a gather's points
constexpr int kHashDeposit = 64, kHashQuery = 72;

// The mesh of a resizing deposit, computed on the device from the bounding
// box (and the total of the charges when they are given), so that the
// deposit kernels are queued without a round trip to the host.  It lives at
// byte offset kMeshDev of the counter area (kCounterBytes long); a copy in
// pinned host memory fills the caller's mesh once it is needed.
struct MeshDev {
  double mn[3], mx[3], d[3], dinv[3];
  double charge;
};
// Tiles of the tiled deposit and its gather: a tile is 2^tb[0] x 2^tb[1] x
// 2^tb[2] cells and owns the (2^tb[a] + 1) nodes per axis its cells touch;
// the particles are binned by tile (bin ntiles: off the mesh).
constexpr int kTileMaxBins = 2048;
// LDS of the tile gather's node records (E and B of every node of a tile, 7
// values per node: FP64 and FP32)
constexpr int kTileGatherBytes64 = 155648, kTileGatherBytes32 = 79872;
// particles per work item of the tile gather
constexpr int kTileGatherChunk = 4096;
struct TileParams {
  double mn[3], dinv[3], d[3];
  int nlo[3], nn[3];
  int tb[3], nt[3], tn[3];  // log2 tile cells, tiles and tile nodes per axis
  int ntiles, dbits, nodes, words;
  const MeshDev* md;  // if set: mn, d and dinv come from here
};
// the bins' first slots, the work items (chunks of at most `chunk` particles
// of one tile) and the deposit's fixed-point scale, built on the device
struct TilePlan {
  uint32_t bstart[kTileMaxBins + 1];  // first slot of every bin
  uint32_t wl[kTileMaxBins + 1];      // first work item of every tile; wl[ntiles] = all
  uint32_t gwl[kTileMaxBins + 1];     // the same for the gather's items of kTileGatherChunk
  double scale, unscale;              // 2^E, 2^-E
  int bad;                            // a charge or the bound n max|q| is not finite
  int pad;
  unsigned long long qmax;            // bits of max |q| (non-negative: ordered as integers)
};

// particles in cell-sorted order (or, `tiled`, binned by tile), as left by a
// deposit (for a following gather) or by sort_for_gather
struct SortedParticles {
  int64_t n = 0;
  DevBuf sv;    // original particle index of each sorted slot (u32)
  DevBuf srec;  // 32-byte records (x, y, z, q) in sorted order (double4), or with
                // f32 16-byte records (ab, de, gh, q) of the FP32 deposit (float4)
  bool f32 = false;
  DevBuf sk;    // f32: the sorted cell keys (u32; the sentinel for dropped particles)
  // tiled: the slots are in bin order (deposit_tiled), not cell order; the
  // gather then stages each tile's nodes in LDS (tile geometry, plan and bin
  // counts of the deposit)
  bool tiled = false;
  TileParams tile{};
  DevBuf plan, total;
  int chunk = 0;
  int64_t nwork = 0;
};

constexpr int kMeshDev = 256, kCounterBytes = 512;
static_assert(kMeshDev + sizeof(MeshDev) <= kCounterBytes, "counter area too small");
// the context's pinned host block: the mesh copy at 0, the two hashes at
// kHashHost, the clamped count of the no-allocation gather at kNoallocHost
constexpr int kHashHost = 128, kNoallocHost = 192, kPinnedBytes = 256;
static_assert(sizeof(MeshDev) <= kHashHost, "pinned block too small");
static_assert(kHashHost + 16 <= kNoallocHost && kNoallocHost + 4 <= kPinnedBytes,
              "pinned block layout");

}  // namespace osc

namespace osc {
class HostRing;

// The grid part of a fused step (Green build, spectral solve, B) for one
// (shape, precision, solver, B, gamma), captured once as a HIP graph and
// replayed for every later mesh of that shape: the Green primitive reads the
// spacing from the device mesh record, so the host needs no mesh first.  The
// graph owns its buffers (rho in, fields out, every temporary of the solve).
struct SolveGraph {
  DevPool* pool = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf rho, ef, bf;
  std::vector<std::pair<void*, size_t>> kept;  // temporaries the graph uses
  uint64_t last = 0;
  int64_t replays = 0;
  ~SolveGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& pb : kept) pool->adopt(pb.first, pb.second);
  }
};
}

struct osc_ctx {
  int device = 0;
  hipStream_t own = nullptr, stream = nullptr;
  bool async = false;
  int prec = OSC_PREC_FP64;
  osc::DevPool pool;
  std::string last_error;
  std::map<int, void*> tw64, tw32;
  osc::GreenCache gcache;
  osc::PipeState pipe;
  // the cell order of the last osc_deposit_particles (when large), kept for a
  // following interpolate_field at the same points (content hash checked)
  osc::SortedParticles kept;
  bool kept_ok = false;
  std::unique_ptr<osc::HostRing> ring;  // pageable host transfers (created on first use)
  void* counters = nullptr;  // small device scratch for counters (persistent)
  osc::MeshDev* mesh_host = nullptr;  // pinned copy of the device mesh
  hipEvent_t mesh_ev = nullptr;        // recorded after the copy is queued
  // pinned copies of the deposit's and the query's content hashes (in the
  // mesh_host allocation), and the event recorded after their copy
  unsigned long long* hash_host = nullptr;
  hipEvent_t hash_ev = nullptr;
  bool prof = false;
  std::vector<osc::ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_free;
  double prof_ms[osc::ST_COUNT] = {0};
  int64_t prof_n[osc::ST_COUNT] = {0};
  // captured solves of the fused step (osc_ctx_set_graphs), and the device
  // mesh record the Green primitive reads while one is being captured
  bool graphs = false;
  // charge deposits through tiles of cells and fixed-point node sums in LDS
  // (deposit_tiled) where the mesh allows; off (OSC_TILED_DEPOSIT=0): the
  // cell-sorted deposit everywhere
  bool tiled_deposit = true;
  // FP32 Green builds take far cells from the field kernel's expansion
  // (far_field_setup); off (OSC_FAR_GREEN=0): every cell from the primitive
  bool far_green = true;
  std::map<std::string, std::unique_ptr<osc::SolveGraph>> sgraphs;
  uint64_t graph_tick = 0;
  int64_t graph_builds = 0, graph_replays = 0;
  const osc::MeshDev* green_md = nullptr;
};

namespace osc {
// brackets the launches of one stage with events when profiling is on
class ProfScope {
 public:
  ProfScope(osc_ctx* c, int stage) : c_(c->prof ? c : nullptr), stage_(stage) {
    if (!c_) return;
    a_ = take();
    (void)hipEventRecord(a_, c_->stream);
  }
  ~ProfScope() { stop(); }
  void stop() {
    if (!c_) return;
    hipEvent_t b = take();
    (void)hipEventRecord(b, c_->stream);
    c_->prof_recs.push_back({stage_, a_, b});
    c_ = nullptr;
  }

 private:
  hipEvent_t take() {
    if (!c_->prof_free.empty()) {
      hipEvent_t e = c_->prof_free.back();
      c_->prof_free.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  osc_ctx* c_;
  int stage_;
  hipEvent_t a_ = nullptr;
};
}  // namespace osc

namespace osc {

// current device guard
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
  }
};

template <typename T>
const cpx<T>* twiddles(osc_ctx* c, int L);

// Staging of caller arrays that may be host or device memory.
class Staging {
 public:
  explicit Staging(osc_ctx* c) : c_(c) {}
  ~Staging();  // drops ring copy-outs left by a call that failed before finish()
  Staging(const Staging&) = delete;
  Staging& operator=(const Staging&) = delete;
  template <typename U>
  const U* in(const U* p, size_t n);
  template <typename U>
  U* out(U* p, size_t n, bool copy_in = false);
  void finish();  // copies outputs back; synchronises unless fully device-resident + async
  bool any_host() const { return any_host_; }

 private:
  struct Pending {
    void* host;
    void* dev;
    size_t bytes;
  };
  osc_ctx* c_;
  bool any_host_ = false;
  bool ring_pending_ = false;
  std::vector<Pending> outs_;
  std::vector<DevBuf> bufs_;
};

bool is_device_ptr(const void* p);

// Copies between pageable caller memory and the device through a ring of
// pinned buffers: the host side of each chunk is copied by a few threads,
// and a chunk's DMA overlaps the host copy of the next.  (HIP's own path
// for pageable memory copies on one thread: 29 GB/s in, 17 GB/s out on the
// MI355X host, against 57 GB/s for pinned DMA either way.)
class CopyThreads;
class HostRing {
 public:
  static constexpr size_t kChunk = size_t(16) << 20;  // bytes per pinned buffer
  static constexpr int kBufs = 4;
  static constexpr size_t kMin = size_t(4) << 20;    // smaller arrays go through HIP
  HostRing();
  ~HostRing();
  HostRing(const HostRing&) = delete;
  HostRing& operator=(const HostRing&) = delete;
  // queues the copy of host[0, bytes) to dev on s; returns once host has been read
  void to_device(hipStream_t s, void* dev, const void* host, size_t bytes);
  // queues the copy of dev[0, bytes) to host on s; finished by drain()
  void to_host(hipStream_t s, void* host, const void* dev, size_t bytes);
  // waits for every queued to_host and copies it out
  void drain();
  // waits for every queued copy and drops the pending copy-outs (error paths)
  void discard();

 private:
  struct Slot {
    char* buf = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    char* out = nullptr;  // pending copy-out of a to_host chunk
    size_t len = 0;
    uint64_t seq = 0;
  };
  Slot& take();  // the next slot, idle (its last copy finished and copied out)
  void settle(Slot& sl);
  Slot slot_[kBufs];
  int next_ = 0;
  uint64_t seq_ = 0;
  CopyThreads* threads_ = nullptr;
};
// true for pageable host memory (not known to HIP)
bool is_pageable(const void* p);

// ---- green.hip --------------------------------------------------------------
void build_green_real(osc_ctx* c, const GreenSpec& s, double* g, bool fast = false);  // (L0,L1,L2) real
// the half spectrum of one pipe Green term from its separable form
void pipe_spectrum(osc_ctx* c, const GreenSpec& s, const SpecDims& d, cpx<double>* spec);
// spectrum of a Green function in precision T (z-real when the z length has a
// compile-time plan); cached by spec
template <typename T>
GreenRef green_spectrum(osc_ctx* c, const GreenSpec& s, DevBuf& keepalive);
// forward R2C of a real (L0,L1,L2) FP64 array into a half spectrum of precision T
// (with cd: g is a primitive and the x pass takes its corner difference;
// sym: per-axis mirror symmetry of that difference, 0 / +1 even / -1 odd;
// with zr: the z pass is left out and *zr describes the stored z lines)
template <typename T>
void spectrum_forward_full(osc_ctx* c, const double* g, cpx<T>* spec, const SpecDims& d,
                           const CornerDiff* cd = nullptr, const int* sym = nullptr,
                           GreenRef* zr = nullptr);
void convert_spectrum(osc_ctx* c, const cpx<double>* in, cpx<float>* out, int64_t n);
// download a half spectrum and expand it to the upstream full-array
// convention F_s[g] (sign s[d] on axis d)
void half_to_upstream(osc_ctx* c, const cpx<double>* spec_dev, const SpecDims& d, const int s[3],
                      double* full_host);
void upstream_to_half(osc_ctx* c, const double* full_host, const SpecDims& d, const int s[3],
                      cpx<double>* spec_dev);

// ---- solve.hip ----------------------------------------------------------------
struct OutSpec {
  int nt = 0;
  struct {
    GreenRef g;
    double sgn;
    int mir;
  } t[kMaxTerm];
  Epilogue ep;
};
template <typename T>
void conv_solve(osc_ctx* c, const double* rho, const int n[3], const SpecDims& d, const int win[3],
                const std::vector<OutSpec>& outs);
void fill_f64(osc_ctx* c, double* p, int64_t n, double v);
void fill_u32(osc_ctx* c, uint32_t* p, int64_t n, uint32_t v);
void fd_gradient(osc_ctx* c, const double* phi, const int n[3], const double delta[3], double gam,
                 double* ef);
void bfield_core(osc_ctx* c, const double* ef, int64_t n, double gam, double* bf);
void flip_z(osc_ctx* c, double* rho, const int n[3]);

// ---- deposit.hip ----------------------------------------------------------------
struct DepositArgs {
  int64_t n;
  const double *x, *y, *z, *q;
  int nlo[3], nn[3];
  double mn[3], delta[3];
  double q_each;        // uniform charge per particle (q == nullptr, charge mode)
  int harness;          // 1: depose_rho_scalar rules (unit weights, scale at end)
  double scale_after;   // harness chrgpermacro
  unsigned long long* hash;  // if set: receives the content hash of (x, y, z)
  const MeshDev* md;  // if set: the kernels read mn and delta from here
  int f32;  // FP32 deposition: weights rounded to FP32 per particle, sums in FP32
};
// rho (nx,ny,nz) on device; the dropped count is left in counter 0
// (read_counter(c, 0)); keep (optional) receives the sorted particles
void deposit_cic(osc_ctx* c, const DepositArgs& a, double* rho, SortedParticles* keep = nullptr);
int64_t read_counter(osc_ctx* c, int idx);  // synchronises
// Queues the content hash of the n points (x, y, z) and the copy of it and
// of the last deposit's hash (byte offset kHashDeposit) to c->hash_host;
// deposit_hash_matches waits for that copy only (work queued after the check
// keeps the GPU busy meanwhile) and compares
void queue_deposit_hash_check(osc_ctx* c, int64_t n, const double* x, const double* y,
                              const double* z);
bool deposit_hash_matches(osc_ctx* c);
void bbox(osc_ctx* c, int64_t n, const double* x, const double* y, const double* z, double lo[3],
          double hi[3]);
double device_sum(osc_ctx* c, const double* v, int64_t n);
// Queued only (no synchronisation): the bounding box of the n points and the
// resized mesh of deposit_particles (open_spacecharge_mod.f90:176-189) for
// cells[a] = nhi - nlo mesh cells per axis, into md->mn, mx, d and dinv
void bbox_resize(osc_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                 const int cells[3], MeshDev* md);
// Queued only: the sum of v[0..n) (the order of device_sum) into *dst
void device_sum_to(osc_ctx* c, const double* v, int64_t n, double* dst);

// ---- gather.hip -------------------------------------------------------------------
struct InterpArgs {
  int nlo[3], nhi[3], nn[3];
  double mn[3], delta[3], hi[3];
};
int64_t interpolate(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf,
                    int64_t n, const double* x, const double* y, const double* z, double* E,
                    double* B, int8_t* clamped);
// the same gather with caller-provided device buffers and nothing allocated:
// the node records of (ef, bf) are packed into rec (64 bytes per node), the
// points are gathered in their own order, and the clamped count is left in
// cnt (cleared first by a fill kernel)
void interpolate_packed(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf,
                        int64_t n, const double* x, const double* y, const double* z, double* E,
                        double* B, int8_t* clamped, double* rec, uint32_t* cnt);
// per-device address of the no-allocation gather's scratch array, looked up
// once per device when a context is created on it (noalloc.hip)
int noalloc_setup(int device, std::string* err);
// same, for the particles of the preceding deposit in their cell-sorted order;
// packed: the node records of (ef, bf) already queued by pack_nodes, or null
// (sp.f32: the FP32 path, which reads the coordinates x, y, z of dropped
// particles)
void interpolate_sorted(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf,
                        const SortedParticles& sp, double* E, double* B, int8_t* clamped,
                        const double* packed = nullptr, const double* x = nullptr,
                        const double* y = nullptr, const double* z = nullptr);
// queues the 64-byte node records (E, B) of the mesh into rec (8 doubles per node)
void pack_nodes(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf, double* rec);
// stable sort of n points by the mesh cell their gather reads (the clamped
// cell of interpolate_field), for a cache-friendly gather
void sort_for_gather(osc_ctx* c, const InterpArgs& a, int64_t n, const double* x, const double* y,
                     const double* z, SortedParticles* out);

}  // namespace osc
