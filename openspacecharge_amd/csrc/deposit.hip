// Deterministic cloud-in-cell deposition (deposit_particles,
// open_spacecharge_mod.f90:156-288; harness twin depose_rho_scalar,
// test_mod.f90:410-474).  Two paths, both bit-identical run to run (no
// floating-point atomics):
//
// Tiled (charge deposits on meshes whose tile node blocks fit in LDS, the
// default): particles binned by tiles of cells in one scatter, each tile's
// nodes summed in LDS as fixed-point integers (deposit_tiled; the kernels
// and the number format are described at k_tile_count).
//
// Cell-sorted (the harness rule, larger meshes):
//   1. cell key per particle: the upstream FP64 index formula, no contraction
//   2. stable LSD radix sort of (cell, particle) pairs (wave-ballot ranking)
//   3. wavefront-segmented reduction of the 8 corner terms over runs of
//      equal cell keys in the sorted array (fixed shuffle and tree orders)
//   4. each node sums the <= 8 cell/corner partials that touch it, in a fixed
//      order.
#include "osc_internal.h"

namespace osc {

namespace {
constexpr int kThreads = 256;
inline int64_t nblk(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  return b > 65536 * 8 ? 65536 * 8 : b;
}

// ---------------------------------------------------------------------------
// reductions (fixed partition and tree: deterministic)
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 2048;

__global__ void k_bbox_partial(const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, int64_t n, double* __restrict__ part) {
  __shared__ double r[6][kThreads];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  const int64_t stride = int64_t(gridDim.x) * kThreads;
  int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x;
  // four independent loads of each coordinate in flight per thread
  for (; i + 3 * stride < n; i += 4 * stride) {
    double a[4], b[4], c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = x[i + u * stride];
      b[u] = y[i + u * stride];
      c[u] = z[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      v[0] = fmin(v[0], a[u]);
      v[1] = fmin(v[1], b[u]);
      v[2] = fmin(v[2], c[u]);
      v[3] = fmax(v[3], a[u]);
      v[4] = fmax(v[4], b[u]);
      v[5] = fmax(v[5], c[u]);
    }
  }
  for (; i < n; i += stride) {
    v[0] = fmin(v[0], x[i]);
    v[1] = fmin(v[1], y[i]);
    v[2] = fmin(v[2], z[i]);
    v[3] = fmax(v[3], x[i]);
    v[4] = fmax(v[4], y[i]);
    v[5] = fmax(v[5], z[i]);
  }
  for (int k = 0; k < 6; ++k) r[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 6; ++k)
        r[k][threadIdx.x] = k < 3 ? fmin(r[k][threadIdx.x], r[k][threadIdx.x + s])
                                  : fmax(r[k][threadIdx.x], r[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = r[threadIdx.x][0];
}

// cells of each axis of a resizing deposit (nhi - nlo)
struct ResizeCells {
  int n[3];
};

// Reduces the block partials to (min x, y, z, max x, y, z) in out; with md,
// also the resized mesh of deposit_particles (open_spacecharge_mod.f90:
// 176-189), by the same correctly rounded double operations in the same
// order as the formula there (the library is built with -ffp-contract=off),
// and the reciprocal spacing the key kernel uses (:225-227).
__global__ void k_bbox_final(const double* __restrict__ part, int nb, double* __restrict__ out,
                             MeshDev* __restrict__ md, ResizeCells rc) {
  __shared__ double r[6][kThreads];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nb; b += kThreads)
    for (int k = 0; k < 6; ++k) v[k] = k < 3 ? fmin(v[k], part[b * 6 + k]) : fmax(v[k], part[b * 6 + k]);
  for (int k = 0; k < 6; ++k) r[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 6; ++k)
        r[k][threadIdx.x] = k < 3 ? fmin(r[k][threadIdx.x], r[k][threadIdx.x + s])
                                  : fmax(r[k][threadIdx.x], r[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = r[threadIdx.x][0];
  if (md && threadIdx.x < 3) {
    const int a = threadIdx.x;
    const double cells = double(rc.n[a]);
    const double d0 = (r[3 + a][0] - r[a][0]) / cells;
    const double lo = r[a][0] - 1.0e-6 * d0, hi = r[3 + a][0] + 1.0e-6 * d0;
    const double d = (hi - lo) / cells;
    md->mn[a] = lo;
    md->mx[a] = hi;
    md->d[a] = d;
    md->dinv[a] = 1.0 / d;
  }
}

__global__ void k_sum_partial(const double* __restrict__ v, int64_t n, double* __restrict__ part) {
  __shared__ double r[kThreads];
  // contiguous chunk per block, strided within the block: fixed order
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
  double s = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kThreads) s = s + v[i];
  r[threadIdx.x] = s;
  __syncthreads();
  for (int st = kThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) r[threadIdx.x] = r[threadIdx.x] + r[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

__global__ void k_sum_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double r[kThreads];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kThreads) s = s + part[b];
  r[threadIdx.x] = s;
  __syncthreads();
  for (int st = kThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) r[threadIdx.x] = r[threadIdx.x] + r[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = r[0];
}

// tiles of the radix sort: RS_T threads, RS_I elements each, per block; the
// key kernel covers one tile with KEY_T threads
constexpr int RS_T = 512, RS_I = 10, RS_TILE = RS_T * RS_I, RS_W = RS_T / 64;
static_assert(RS_T >= 256, "one thread per digit in the scatter prologue");
constexpr int KEY_T = 1024, KEY_I = RS_TILE / KEY_T;
static_assert(RS_TILE % KEY_T == 0, "the key kernel must cover whole sort tiles");
static_assert((RS_TILE / RS_W) % 64 == 0, "each wave of the scatter ranks whole 64-pair steps");

// ---------------------------------------------------------------------------
// step 1: cell keys
// ---------------------------------------------------------------------------
struct KeyParams {
  double mn[3], dinv[3], base[3], d[3];
  int nlo[3], nn[3], nhi[3], wofs[3];
  uint32_t sentinel;
  int clamp;  // 0: deposit rule (off-mesh particles get the sentinel); 1: gather rule
  const MeshDev* md;  // if set: mn and dinv come from here (a mesh resized on the device)
  uint32_t* clear;    // if set: nclear words zeroed by the kernel (the cell flags of the
  int64_t nclear;     // deposit, cleared here instead of by a launch of their own)
};

__device__ __forceinline__ bool cell_of(double x, double mn, double dinv, double base, int nlo,
                                        int nn, int& i0) {
  // ip = floor((x - min)*(1/dx) + 1) (open_spacecharge_mod.f90:259); node ip-nlo
  const double f = floor((x - mn) * dinv + base);
  const double lo = double(nlo), hi = double(nlo + nn - 2);
  if (!(f >= lo && f <= hi)) return false;  // all 8 nodes must be on the mesh
  i0 = int(f) - nlo;
  return true;
}

// cell read by the gather: interpolate_field's clamp to [1, nhi-1] (:320-355),
// then the array cell inside a field with lower bound nlo
__device__ __forceinline__ int gather_cell(double x, double mn, double dinv, int nlo, int nhi,
                                           int nn) {
  const double f = floor((x - mn) * dinv + 1.0);
  const int ip = !(f >= 1.0) ? 1 : (f > double(nhi - 1) ? nhi - 1 : int(f));
  const int i0 = ip - nlo;
  return i0 < 0 ? 0 : (i0 > nn - 2 ? nn - 2 : i0);
}

// Content hash of a particle array: the sum mod 2^64 over particles of a mix
// of each coordinate's bits with the particle's index (the murmur3 64-bit
// finaliser), so any change of a coordinate or of the order changes it
// except with probability ~2^-64; the sum is order-free, so per-wave
// partial sums may be added atomically.
__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
__device__ __forceinline__ uint64_t particle_mix(int64_t i, double x, double y, double z) {
  const uint64_t s = uint64_t(i) * 0x9e3779b97f4a7c15ull;
  return fmix64(uint64_t(__double_as_longlong(x)) ^ s) +
         fmix64(uint64_t(__double_as_longlong(y)) ^ (s + 0x632be59bd9b4e019ull)) +
         fmix64(uint64_t(__double_as_longlong(z)) ^ (s + 0x85157af5c3a3f5bbull));
}
// the workgroup's sum of the threads' partial hashes, stored at part[blockIdx.x]
// (every thread calls it; at most 16 waves)
__device__ __forceinline__ void hash_block(uint64_t* part, uint64_t h) {
  __shared__ uint64_t ws[16];
  for (int off = 32; off > 0; off >>= 1) h += uint64_t(__shfl_xor((unsigned long long)h, off, 64));
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < int(blockDim.x + 63) / 64; ++w) t += ws[w];
    part[blockIdx.x] = t;
  }
}

__global__ void k_hash_xyz(const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ z, int64_t n, uint64_t* __restrict__ part) {
  uint64_t h = 0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    h += particle_mix(i, x[i], y[i], z[i]);
  hash_block(part, h);
}

// the sum of nparts partial hashes (one workgroup)
__global__ void k_hash_sum(const uint64_t* __restrict__ part, int nparts,
                           unsigned long long* __restrict__ out) {
  uint64_t h = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) h += part[b];
  __shared__ uint64_t ws[16];
  for (int off = 32; off > 0; off >>= 1) h += uint64_t(__shfl_xor((unsigned long long)h, off, 64));
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < int(blockDim.x + 63) / 64; ++w) t += ws[w];
    *out = t;
  }
}

// orders a wave's LDS accesses around lanes reading what other lanes wrote
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// cell key of every particle, and its 32-byte record (x, y, z, q) in input
// order (the sort's payload: one record per particle instead of three
// scattered coordinates).  Block b takes the sort tile b (RS_TILE particles)
// and leaves the tile's count of each first-pass digit (key & mask0) in
// hist[d * nb + b], the layout of k_rs_hist.  With hpart, the workgroup's
// partial content hash of (x, y, z) goes to hpart[b].
// F32 (FP32 deposition): the record is instead the 16-byte (ab, de, gh, q)
// in FP32 -- the CIC weights formed in FP64 by the upstream formula
// (open_spacecharge_mod.f90:262-264) from the FP64 cell index, then rounded
// -- stored by each lane directly (consecutive lanes, consecutive records).
template <bool F32>
__global__ void __launch_bounds__(KEY_T)
    k_cic_keys(const double* __restrict__ x, const double* __restrict__ y,
               const double* __restrict__ z, const double* __restrict__ q, int64_t n, KeyParams p,
               uint32_t* __restrict__ keys, double4* __restrict__ rec,
               uint32_t* __restrict__ ndrop, uint32_t mask0, uint32_t* __restrict__ hist, int nb,
               uint64_t* __restrict__ hpart) {
  __shared__ uint32_t h[256];
  __shared__ uint32_t cnt;
  // each wave's 64 records, staged to be stored as 2 contiguous KB
  __shared__ double stage[KEY_T / 64][64 * 5];
  for (int d = threadIdx.x; d < 256; d += KEY_T) h[d] = 0;
  if (threadIdx.x == 0) cnt = 0;
  if (p.md)
    for (int d = 0; d < 3; ++d) {
      p.mn[d] = p.md->mn[d];
      p.dinv[d] = p.md->dinv[d];
      p.d[d] = p.md->d[d];
    }
  __syncthreads();
  if (p.clear)
    for (int64_t w = blockIdx.x * int64_t(KEY_T) + threadIdx.x; w < p.nclear;
         w += int64_t(gridDim.x) * KEY_T)
      p.clear[w] = 0u;
  uint32_t mine = 0;
  uint64_t hs = 0;
  const int64_t base = int64_t(blockIdx.x) * RS_TILE;
  const int lane = threadIdx.x & 63;
  double* w = stage[threadIdx.x >> 6];
#pragma unroll
  for (int it = 0; it < KEY_I; ++it) {
    const int64_t wb = base + it * KEY_T + (threadIdx.x & ~63);  // the wave's first particle
    if (wb >= n) break;                                          // uniform per wave
    const int64_t i = wb + lane;
    double xi = 0.0, yi = 0.0, zi = 0.0, qi = 0.0;
    if (i < n) {
      xi = x[i];
      yi = y[i];
      zi = z[i];
      qi = q ? q[i] : 0.0;
      if (hpart) hs += particle_mix(i, xi, yi, zi);
      int i0, j0, k0;
      bool ok = true;
      if (p.clamp) {
        i0 = gather_cell(xi, p.mn[0], p.dinv[0], p.nlo[0], p.nhi[0], p.nn[0]);
        j0 = gather_cell(yi, p.mn[1], p.dinv[1], p.nlo[1], p.nhi[1], p.nn[1]);
        k0 = gather_cell(zi, p.mn[2], p.dinv[2], p.nlo[2], p.nhi[2], p.nn[2]);
      } else {
        ok = cell_of(xi, p.mn[0], p.dinv[0], p.base[0], p.nlo[0], p.nn[0], i0) &&
             cell_of(yi, p.mn[1], p.dinv[1], p.base[1], p.nlo[1], p.nn[1], j0) &&
             cell_of(zi, p.mn[2], p.dinv[2], p.base[2], p.nlo[2], p.nn[2], k0);
      }
      const uint32_t key =
          ok ? uint32_t(i0) + uint32_t(p.nn[0]) * (uint32_t(j0) + uint32_t(p.nn[1]) * uint32_t(k0))
             : p.sentinel;
      keys[i] = key;
      atomicAdd(&h[key & mask0], 1u);
      mine += ok ? 0u : 1u;
      if (F32) {
        float4 r = make_float4(0.f, 0.f, 0.f, float(qi));
        if (ok) {
          r.x = float(((p.mn[0] - xi) + double(i0 + p.wofs[0]) * p.d[0]) * p.dinv[0]);
          r.y = float(((p.mn[1] - yi) + double(j0 + p.wofs[1]) * p.d[1]) * p.dinv[1]);
          r.z = float(((p.mn[2] - zi) + double(k0 + p.wofs[2]) * p.d[2]) * p.dinv[2]);
        }
        reinterpret_cast<float4*>(rec)[i] = r;
      }
    }
    if (F32) continue;  // uniform: the 16-byte records are stored above
    w[lane * 5] = xi;
    w[lane * 5 + 1] = yi;
    w[lane * 5 + 2] = zi;
    w[lane * 5 + 3] = qi;
    wave_lds_sync();
    // 16-byte element e of the wave's records: record e / 2, half e % 2
    const int nrec = n - wb < 64 ? int(n - wb) : 64;
    double2* o = reinterpret_cast<double2*>(rec + wb);
#pragma unroll
    for (int qq = 0; qq < 2; ++qq) {
      const int e = qq * 64 + lane;
      if ((e >> 1) < nrec) {
        const double* src = w + (e >> 1) * 5 + 2 * (e & 1);
        o[e] = make_double2(src[0], src[1]);
      }
    }
    wave_lds_sync();  // read before the next iteration overwrites the stage
  }
  if (hpart) hash_block(hpart, hs);
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  for (int d = threadIdx.x; d <= int(mask0); d += KEY_T) hist[int64_t(d) * nb + blockIdx.x] = h[d];
  if (threadIdx.x == 0 && cnt) atomicAdd(ndrop, cnt);
}

// ---------------------------------------------------------------------------
// step 2: stable LSD radix sort of (key, value) pairs
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(RS_T)
    k_rs_hist(const uint32_t* __restrict__ keys, int64_t n, int shift, uint32_t mask,
              uint32_t* __restrict__ hist, int nb) {
  __shared__ uint32_t h[256];
  for (int i = threadIdx.x; i < 256; i += RS_T) h[i] = 0;
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * RS_TILE;
  for (int it = 0; it < RS_I; ++it) {
    const int64_t i = base + it * RS_T + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d <= int(mask); d += RS_T) hist[int64_t(d) * nb + blockIdx.x] = h[d];
}

// inclusive scan of one value per thread over the RS_T threads: within each
// wave by shuffles, then the wave totals (s holds RS_W entries)
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t v, uint32_t* s) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  if (lane == 63) s[wave] = v;
  __syncthreads();
  for (int w = 0; w < wave; ++w) v += s[w];
  __syncthreads();
  return v;
}

// Per-block digit counts hist[d * nb + b] -> their exclusive scan over the
// blocks of each digit (in place), and each digit's total.  One block per
// digit; thread t scans a contiguous slice of the row.
__global__ void __launch_bounds__(RS_T)
    k_digit_scan(uint32_t* __restrict__ hist, int nb, uint32_t* __restrict__ total) {
  __shared__ uint32_t s[RS_W];
  uint32_t* row = hist + int64_t(blockIdx.x) * nb;
  const int per = (nb + RS_T - 1) / RS_T;
  const int b0 = int(threadIdx.x) * per;
  uint32_t acc = 0;
  for (int q = 0; q < per; ++q)
    if (b0 + q < nb) acc += row[b0 + q];
  const uint32_t incl = block_scan_incl(acc, s);
  uint32_t run = incl - acc;
  for (int q = 0; q < per; ++q)
    if (b0 + q < nb) {
      const uint32_t v = row[b0 + q];
      row[b0 + q] = run;
      run += v;
    }
  if (threadIdx.x == RS_T - 1) total[blockIdx.x] = incl;
}

// payload gathered after the last pass: the particle records in sorted order
struct SortPayload {
  const double4* rec;
  double4* srec;
};


// the LDS operations of one wave run in program order; the fences keep the
// compiler from moving accesses across (other lanes read what was written)

// One pass over a tile of RS_TILE pairs per block.  Wave w ranks its own
// contiguous share of the tile (RS_TILE / RS_W pairs, 64 per step, held in
// registers) with wave-private digit counters: the rank of a pair is the
// count of earlier pairs of the share with its digit (earlier steps, then
// lower lanes).  The tile's pairs are then placed in LDS in stable digit order
// (digit, then wave, then rank) and written out so that consecutive threads
// store consecutive slots of one digit's run.  col[d * nb + b] is the
// exclusive scan over blocks of digit d's per-block counts (k_digit_scan) and
// total[d] that digit's count; the slots of digit d start after all smaller
// digits.
__global__ void __launch_bounds__(RS_T)
    k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                 uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n, int shift,
                 int bits, const uint32_t* __restrict__ col, const uint32_t* __restrict__ total,
                 int nb) {
  constexpr int SHARE = RS_TILE / RS_W;  // pairs per wave
  __shared__ uint32_t lk[RS_TILE], lv[RS_TILE];
  __shared__ uint32_t wcnt[RS_W][256];  // wave counters, then each wave's first local slot
  __shared__ uint32_t gofs[256], scan[RS_W];
  const uint32_t mask = (1u << bits) - 1u;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int j = threadIdx.x; j < RS_W * 256; j += RS_T) (&wcnt[0][0])[j] = 0u;
  __syncthreads();
  const int64_t wbase = int64_t(blockIdx.x) * RS_TILE + int64_t(wave) * SHARE;
  uint32_t* cnt = wcnt[wave];
  uint32_t key[SHARE / 64], val[SHARE / 64], rk[SHARE / 64];
#pragma unroll
  for (int it = 0; it < SHARE / 64; ++it) {
    const int64_t i = wbase + it * 64 + lane;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const uint32_t v = valid ? (vin ? vin[i] : uint32_t(i)) : 0u;
    const uint32_t d = (k >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t r = __popcll(peers & lt);
    const uint32_t before = cnt[d];
    wave_lds_sync();
    if (valid && r == 0) cnt[d] = before + uint32_t(__popcll(peers));
    wave_lds_sync();
    key[it] = k;
    val[it] = v;
    rk[it] = before + r;
  }
  __syncthreads();
  {
    const int d = threadIdx.x;  // threads 0..255 own a digit; all take part in the scans
    const bool dig = d < 256;
    uint32_t mine = 0;  // this block's count of digit d; wave prefixes in place
    if (dig)
#pragma unroll
      for (int w = 0; w < RS_W; ++w) {
        const uint32_t c = wcnt[w][d];
        wcnt[w][d] = mine;
        mine += c;
      }
    const uint32_t tot = dig && d <= int(mask) ? total[d] : 0u;
    const uint32_t dbase = block_scan_incl(tot, scan) - tot;  // first slot of digit d
    const uint32_t start = block_scan_incl(mine, scan) - mine;  // first local slot of digit d
    if (dig) {
#pragma unroll
      for (int w = 0; w < RS_W; ++w) wcnt[w][d] += start;
      const uint32_t g = d <= int(mask) ? dbase + col[int64_t(d) * nb + blockIdx.x] : 0u;
      gofs[d] = g - start;  // global slot of local slot j of digit d: gofs[d] + j
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < SHARE / 64; ++it) {
    const int64_t i = wbase + it * 64 + lane;
    if (i < n) {
      const uint32_t pos = wcnt[wave][(key[it] >> shift) & mask] + rk[it];
      lk[pos] = key[it];
      lv[pos] = val[it];
    }
  }
  __syncthreads();
  const int64_t rem = n - int64_t(blockIdx.x) * RS_TILE;
  const int cnt_t = rem < RS_TILE ? int(rem) : RS_TILE;
  for (int j = threadIdx.x; j < cnt_t; j += RS_T) {
    const uint32_t k = lk[j];
    const uint32_t g = gofs[(k >> shift) & mask] + uint32_t(j);
    kout[g] = k;
    vout[g] = lv[j];
  }
}

// records in sorted order (slot s holds particle sv[s])
__global__ void k_sort_payload(const uint32_t* __restrict__ sv, int64_t n, SortPayload pl) {
  for (int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; s < n;
       s += int64_t(gridDim.x) * blockDim.x)
    pl.srec[s] = pl.rec[sv[s]];
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// step 3: wavefront-segmented reduction of the corner terms over runs of
// equal cell keys in the sorted particle array
// ---------------------------------------------------------------------------
struct CellParams {
  double mn[3], d[3], dinv[3];
  int wofs[3];
  int nn[3];
  int hasq;      // per-particle charges
  int harness;   // unit weights (scale applied at the node)
  double q_each;
  const MeshDev* md;  // if set: mn, d and dinv come from here
};

// the 8 CIC terms of one particle, corner c = di + 2 dj + 4 dk, each in the
// upstream product order (open_spacecharge_mod.f90:274-281)
__device__ __forceinline__ void cic_terms(double x, double y, double z, double q,
                                          const CellParams& p, int i0, int j0, int k0,
                                          double t[8]) {
  const double ab = ((p.mn[0] - x) + double(i0 + p.wofs[0]) * p.d[0]) * p.dinv[0];
  const double de = ((p.mn[1] - y) + double(j0 + p.wofs[1]) * p.d[1]) * p.dinv[1];
  const double gh = ((p.mn[2] - z) + double(k0 + p.wofs[2]) * p.d[2]) * p.dinv[2];
  if (p.harness) {
    t[0] = ab * de * gh;
    t[2] = ab * (1. - de) * gh;
    t[6] = ab * (1. - de) * (1. - gh);
    t[4] = ab * de * (1. - gh);
    t[5] = (1. - ab) * de * (1. - gh);
    t[7] = (1. - ab) * (1. - de) * (1. - gh);
    t[3] = (1. - ab) * (1. - de) * gh;
    t[1] = (1. - ab) * de * gh;
  } else {
    t[0] = ab * de * gh * q;
    t[2] = ab * (1. - de) * gh * q;
    t[6] = ab * (1. - de) * (1. - gh) * q;
    t[4] = ab * de * (1. - gh) * q;
    t[5] = (1. - ab) * de * (1. - gh) * q;
    t[7] = (1. - ab) * (1. - de) * (1. - gh) * q;
    t[3] = (1. - ab) * (1. - de) * gh * q;
    t[1] = (1. - ab) * de * gh * q;
  }
}

// DPP lane moves of the 64-wide wavefront (gfx9 family): row_shr:n shifts
// within 16-lane rows, row_bcast:15 / row_bcast:31 pass lane 15 of a row / lane
// 31 to the rows selected by the row mask.  Lanes with no source keep `old`.
constexpr int kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v) {
  return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(v), CTRL, ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t b = uint64_t(__double_as_longlong(v));
  const uint32_t lo = dpp_u32<CTRL, ROWMASK>(0u, uint32_t(b));
  const uint32_t hi = dpp_u32<CTRL, ROWMASK>(0u, uint32_t(b >> 32));
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f32(float v) {
  return __uint_as_float(dpp_u32<CTRL, ROWMASK>(0u, __float_as_uint(v)));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_any(double v) {
  return dpp_f64<CTRL, ROWMASK>(v);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_any(float v) {
  return dpp_f32<CTRL, ROWMASK>(v);
}
// one step of the segmented scan: add the source lane's partial when it holds
// the same key (keys are sorted, so equal keys are one contiguous run)
template <int CTRL, int ROWMASK, typename A>
__device__ __forceinline__ void seg_step(uint32_t k, A (&t)[8]) {
  const uint32_t kd = dpp_u32<CTRL, ROWMASK>(0xFFFFFFFFu, k);
  A o[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = dpp_any<CTRL, ROWMASK>(t[c]);
  if (kd == k)
#pragma unroll
    for (int c = 0; c < 8; ++c) t[c] = o[c] + t[c];
}

// Each lane takes one sorted particle; a segmented inclusive scan over equal
// keys (fixed DPP pattern: 16-lane rows, then row broadcasts) leaves each run's
// sum within the wave at its last lane.  A run that starts and ends inside the wave is
// written to S directly.  A run crossing a wave boundary leaves partials:
// head[w] = its part in wave w when it began in an earlier wave and ends in w,
// tail[w] = its part from its start in w (or lane 0) to lane 63 when it
// continues into w+1.  k_seg_fix then adds, for each such run,
// tail[first] + tail[middle waves...] + head[last] in wave order (a fixed
// workgroup tree for runs longer than kSegSerial waves).  The summation order
// depends only on the sorted order: deterministic, no atomics on data.
constexpr int kSegSerial = 64;

//
// Slot s holds particle sv[s]; its record rec[sv[s]] is read here (and, with
// KEEP, for a following gather, stored to srec[s]) for every slot, the
// dropped ones at the end included.
template <bool KEEP>
__global__ void __launch_bounds__(kThreads)
    k_seg_reduce(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ sv,
                 const double4* __restrict__ rec, double4* __restrict__ srec,
                 int64_t n, const uint32_t* __restrict__ ndrop,
                 CellParams p, double* __restrict__ S, uint8_t* __restrict__ occ,
                 double* __restrict__ head, double* __restrict__ tail) {
  const int64_t nv = n - int64_t(*ndrop);
  const int64_t ncell = int64_t(p.nn[0]) * p.nn[1] * p.nn[2];
  if (p.md)
    for (int d = 0; d < 3; ++d) {
      p.mn[d] = p.md->mn[d];
      p.d[d] = p.md->d[d];
      p.dinv[d] = p.md->dinv[d];
    }
  const int lane = threadIdx.x & 63;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  // the loop bound is uniform per block (blockDim is a multiple of 64), so
  // every lane of a wave takes part in the DPP steps
  for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    const int64_t w = i >> 6;
    const bool valid = i < nv;
    // the wave's 64 records: two lanes read each 32-byte record (16 bytes
    // each, one line per record and load instruction) into LDS
    __shared__ double stage[kThreads / 64][64 * 5];
    double* st = stage[threadIdx.x >> 6];
    const int64_t wb = base + (threadIdx.x & ~63);
    const int nrec = n - wb < 64 ? int(n - wb > 0 ? n - wb : 0) : 64;
    const uint32_t myv = i < n ? sv[i] : 0u;
#pragma unroll
    for (int qq = 0; qq < 2; ++qq) {
      const int e = qq * 64 + lane;  // record e / 2, half e % 2
      const uint32_t vr = uint32_t(__shfl(int(myv), e >> 1));
      if ((e >> 1) < nrec) {
        const double2 h = reinterpret_cast<const double2*>(rec + vr)[e & 1];
        st[(e >> 1) * 5 + 2 * (e & 1)] = h.x;
        st[(e >> 1) * 5 + 2 * (e & 1) + 1] = h.y;
      }
    }
    wave_lds_sync();
    double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
    if (i < n) r = make_double4(st[lane * 5], st[lane * 5 + 1], st[lane * 5 + 2], st[lane * 5 + 3]);
    if (KEEP) {
      // the same records in cell order, stored as 2 contiguous KB
      double2* o = reinterpret_cast<double2*>(srec + wb);
#pragma unroll
      for (int qq = 0; qq < 2; ++qq) {
        const int e = qq * 64 + lane;  // record e / 2, half e % 2
        if ((e >> 1) < nrec) {
          const double* src = st + (e >> 1) * 5 + 2 * (e & 1);
          o[e] = make_double2(src[0], src[1]);
        }
      }
    }
    wave_lds_sync();  // read before the next chunk overwrites the stage
    const uint32_t k = valid ? ks[i] : 0xFFFFFFFFu;
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (valid) {
      const int i0 = int(k % uint32_t(p.nn[0]));
      const int j0 = int((k / uint32_t(p.nn[0])) % uint32_t(p.nn[1]));
      const int k0 = int(k / (uint32_t(p.nn[0]) * uint32_t(p.nn[1])));
      cic_terms(r.x, r.y, r.z, p.hasq ? r.w : p.q_each, p, i0, j0, k0, t);
    }
    // within 16-lane rows: shifts 1, 2, 4, 8; then lane 15 of rows 0 and 2
    // into rows 1 and 3, then lane 31 into rows 2 and 3
    seg_step<kDppRowShr + 1, 0xF>(k, t);
    seg_step<kDppRowShr + 2, 0xF>(k, t);
    seg_step<kDppRowShr + 4, 0xF>(k, t);
    seg_step<kDppRowShr + 8, 0xF>(k, t);
    seg_step<kDppRowBcast15, 0xA>(k, t);
    seg_step<kDppRowBcast31, 0xC>(k, t);
    const uint32_t kn = __shfl_down(k, 1, 64);
    const uint32_t kfirst = __shfl(k, 0, 64);
    if (!valid || (lane < 63 && kn == k)) continue;  // not the last lane of its run
    const int64_t w0 = w << 6;
    const bool started = kfirst != k || w0 == 0 || ks[w0 - 1] != k;
    const bool ends = lane < 63 || i + 1 >= nv || ks[i + 1] != k;
    if (started && ends) {
#pragma unroll
      for (int c = 0; c < 8; ++c) S[c * ncell + k] = t[c];
      occ[k] = 1;
    } else if (ends) {
#pragma unroll
      for (int c = 0; c < 8; ++c) head[w * 8 + c] = t[c];
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) tail[w * 8 + c] = t[c];
    }
  }
}

// FP32 deposition: the same run decomposition over the 16-byte weight
// records (ab, de, gh, q) of k_cic_keys<true>, the 8 terms formed and summed
// in FP32 in the upstream product order; cell sums, head and tail in FP32.
// With KEEP the records go to srec in cell order for the gather.
template <bool KEEP>
__global__ void __launch_bounds__(kThreads)
    k_seg_reduce32(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ sv,
                   const float4* __restrict__ rec, float4* __restrict__ srec, int64_t n,
                   const uint32_t* __restrict__ ndrop, float q_each, int hasq,
                   float* __restrict__ S, uint8_t* __restrict__ occ, float* __restrict__ head,
                   float* __restrict__ tail, int64_t ncell) {
  const int64_t nv = n - int64_t(*ndrop);
  const int lane = threadIdx.x & 63;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    const int64_t w = i >> 6;
    const bool valid = i < nv;
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < n) r = rec[sv[i]];
    if (KEEP && i < n) srec[i] = r;
    const uint32_t k = valid ? ks[i] : 0xFFFFFFFFu;
    float t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (valid) {
      const float ab = r.x, de = r.y, gh = r.z, q = hasq ? r.w : q_each;
      t[0] = ab * de * gh * q;
      t[2] = ab * (1.f - de) * gh * q;
      t[6] = ab * (1.f - de) * (1.f - gh) * q;
      t[4] = ab * de * (1.f - gh) * q;
      t[5] = (1.f - ab) * de * (1.f - gh) * q;
      t[7] = (1.f - ab) * (1.f - de) * (1.f - gh) * q;
      t[3] = (1.f - ab) * (1.f - de) * gh * q;
      t[1] = (1.f - ab) * de * gh * q;
    }
    seg_step<kDppRowShr + 1, 0xF>(k, t);
    seg_step<kDppRowShr + 2, 0xF>(k, t);
    seg_step<kDppRowShr + 4, 0xF>(k, t);
    seg_step<kDppRowShr + 8, 0xF>(k, t);
    seg_step<kDppRowBcast15, 0xA>(k, t);
    seg_step<kDppRowBcast31, 0xC>(k, t);
    const uint32_t kn = __shfl_down(k, 1, 64);
    const uint32_t kfirst = __shfl(k, 0, 64);
    if (!valid || (lane < 63 && kn == k)) continue;  // not the last lane of its run
    const int64_t w0 = w << 6;
    const bool started = kfirst != k || w0 == 0 || ks[w0 - 1] != k;
    const bool ends = lane < 63 || i + 1 >= nv || ks[i + 1] != k;
    if (started && ends) {
#pragma unroll
      for (int c = 0; c < 8; ++c) S[c * ncell + k] = t[c];
      occ[k] = 1;
    } else if (ends) {
#pragma unroll
      for (int c = 0; c < 8; ++c) head[w * 8 + c] = t[c];
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) tail[w * 8 + c] = t[c];
    }
  }
}

// Wave holding the last element of the run of key k that continues past
// the end of wave w (ks sorted, nv valid keys): the first wave ww > w whose
// last element is not k, or the last wave.  Galloping over wave ends, then
// bisection, so a run of m waves costs O(log m) reads.
__device__ __forceinline__ int64_t run_end_wave(const uint32_t* __restrict__ ks, int64_t nv,
                                                int64_t w, uint32_t k) {
  const int64_t nw = (nv + 63) >> 6;
  auto cont = [&](int64_t ww) {  // the run continues past the end of wave ww
    const int64_t e = (ww << 6) + 64;
    return e < nv && ks[e] == k;
  };
  int64_t lo = w, step = 1;  // cont(lo) holds
  int64_t hi = w + 1;
  while (hi < nw && cont(hi)) {
    lo = hi;
    step <<= 1;
    hi = lo + step;
  }
  if (hi > nw - 1) hi = nw - 1;  // !cont(hi) holds (the last wave never continues)
  while (hi - lo > 1) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (cont(mid)) lo = mid;
    else hi = mid;
  }
  return hi;
}

// One thread per wave whose last run continues into the next wave and began
// in this wave: completes that run.  Runs spanning more than kSegSerial waves
// are listed for k_seg_heavy.
template <typename A>
__global__ void k_seg_fix(const uint32_t* __restrict__ ks, int64_t n,
                          const uint32_t* __restrict__ ndrop, int64_t ncell,
                          const A* __restrict__ head, const A* __restrict__ tail,
                          A* __restrict__ S, uint8_t* __restrict__ occ,
                          uint32_t* __restrict__ heavy, uint32_t* __restrict__ nheavy) {
  const int64_t nv = n - int64_t(*ndrop);
  const int64_t nw = (nv + 63) >> 6;
  for (int64_t w = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; w < nw;
       w += int64_t(gridDim.x) * blockDim.x) {
    const int64_t last = (w << 6) + 63;
    if (last + 1 >= nv) continue;
    const uint32_t k = ks[last];
    if (ks[last + 1] != k) continue;                          // no continuation
    if (ks[w << 6] == k && w > 0 && ks[(w << 6) - 1] == k) continue;  // began earlier
    // the middle waves are summed in order as their ends are checked; past
    // kSegSerial of them the run goes to k_seg_heavy
    A acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = tail[w * 8 + c];
    int64_t ww = w + 1;
    bool long_run = false;
    for (;;) {
      const int64_t e = (ww << 6) + 64;  // first element of the next wave
      if (!(e < nv && ks[e] == k)) break;  // the run ends in wave ww
      if (ww + 1 - w > kSegSerial) {
        long_run = true;
        break;
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = acc[c] + tail[ww * 8 + c];
      ++ww;
    }
    if (long_run) {
      heavy[atomicAdd(nheavy, 1u)] = uint32_t(w);
      continue;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) S[c * ncell + k] = acc[c] + head[ww * 8 + c];
    occ[k] = 1;
  }
}

// One workgroup per long run: thread t sums the middle-wave partials
// w+1+t, w+1+t+kThreads, ... in order, then a fixed LDS tree.
template <typename A>
__global__ void __launch_bounds__(kThreads)
    k_seg_heavy(const uint32_t* __restrict__ ks, int64_t n, const uint32_t* __restrict__ ndrop,
                int64_t ncell, const A* __restrict__ head, const A* __restrict__ tail,
                A* __restrict__ S, uint8_t* __restrict__ occ,
                const uint32_t* __restrict__ heavy, const uint32_t* __restrict__ nheavy) {
  __shared__ A red[8][kThreads];
  const int64_t nv = n - int64_t(*ndrop);
  const uint32_t nh = *nheavy;
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const int64_t w = heavy[h];
    const uint32_t k = ks[(w << 6) + 63];
    const int64_t w1 = run_end_wave(ks, nv, w, k);
    A s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t ww = w + 1 + threadIdx.x; ww < w1; ww += kThreads)
#pragma unroll
      for (int c = 0; c < 8; ++c) s[c] = s[c] + tail[ww * 8 + c];
#pragma unroll
    for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int m = kThreads / 2; m > 0; m >>= 1) {
      if (int(threadIdx.x) < m)
#pragma unroll
        for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = red[c][threadIdx.x] + red[c][threadIdx.x + m];
      __syncthreads();
    }
    if (threadIdx.x < 8) {
      const int c = threadIdx.x;
      S[c * ncell + k] = (tail[w * 8 + c] + red[c][0]) + head[w1 * 8 + c];
      if (c == 0) occ[k] = 1;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// step 4: nodes
// ---------------------------------------------------------------------------
// node = sum of the 8 adjacent cells' corner sums in a fixed order; cells
// without particles (occ 0) hold no sums and add nothing
//
// The 8 flags and the 8 corner sums of a node are read together and the sums
// of unflagged cells (memory never written this time) are selected away, so
// a node costs one round of loads.  With `sparse` (fewer particles than
// cells, most cells empty) the flags are read first and only the sums of
// flagged cells are loaded: two rounds of loads, but the sums of the empty
// cells, most of the 8 planes, are never fetched.  The accumulator starts at
// +0 and a sum of doubles never yields -0 in round-to-nearest, so adding a
// selected +0 leaves it unchanged: the node values are those of the
// conditional sum either way.  Cell counts are below 2^31 (deposit_cic), so
// the index arithmetic is 32-bit.
template <typename A>
__global__ void k_node(const A* __restrict__ S, const uint8_t* __restrict__ occ, int nx,
                       int ny, int nz, int sparse, int scale_after, double scale,
                       double* __restrict__ rho) {
  const uint32_t ncell = uint32_t(nx) * uint32_t(ny) * uint32_t(nz);
  const uint32_t sxy = uint32_t(nx) * uint32_t(ny);
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += gridDim.x * blockDim.x) {
    const uint32_t k = c / sxy, r = c - k * sxy, j = r / uint32_t(nx), i = r - j * uint32_t(nx);
    uint8_t f[8];
    A v[8];
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) {
      const int di = cc & 1, dj = (cc >> 1) & 1, dk = cc >> 2;
      const bool in = i >= uint32_t(di) && j >= uint32_t(dj) && k >= uint32_t(dk);
      const uint32_t cell = c - uint32_t(di) - uint32_t(dj) * uint32_t(nx) - uint32_t(dk) * sxy;
      f[cc] = in ? occ[cell] : uint8_t(0);
      if (!sparse) v[cc] = in ? S[size_t(cc) * ncell + cell] : A(0);
    }
    if (sparse) {
#pragma unroll
      for (int cc = 0; cc < 8; ++cc) {
        const int di = cc & 1, dj = (cc >> 1) & 1, dk = cc >> 2;
        const uint32_t cell = c - uint32_t(di) - uint32_t(dj) * uint32_t(nx) - uint32_t(dk) * sxy;
        v[cc] = f[cc] ? S[size_t(cc) * ncell + cell] : A(0);
      }
    }
    A acc = A(0);
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) acc = acc + (f[cc] ? v[cc] : A(0));
    rho[c] = scale_after ? double(acc) * scale : double(acc);
  }
}

// ---------------------------------------------------------------------------
// tiled deposition (deposit_tiled): the particles binned by tiles of cells,
// each tile's nodes accumulated in LDS as fixed-point integers
// ---------------------------------------------------------------------------
// A tile is 2^tb[0] x 2^tb[1] x 2^tb[2] cells and owns the nodes its cells
// touch, (2^tb[a] + 1) per axis (neighbouring tiles share a face of nodes).
//   1. k_tile_count: the bin of every particle (its tile under the deposit's
//      cell rule; off-mesh particles go to bin ntiles), counted per block
//   2. k_digit_scan: the exclusive scan of each bin's counts over the blocks;
//      k_tile_plan: the bins' first slots, the chunks of every tile (work
//      items of at most `chunk` particles) and the fixed-point scale
//   3. k_tile_scatter: every particle's record and index to a slot of its
//      bin (wave-aggregated LDS counters rank the particles of a bin, staged
//      in LDS so that the stores of a bin's run are contiguous)
//   4. k_tile_accum: one workgroup per chunk adds the 8 corner terms of each
//      particle to the tile's nodes in LDS as fixed-point integers and stores
//      the chunk's node block
//   5. k_tile_nodes: each node adds the blocks of every chunk of the <= 8
//      tiles holding it and converts the sum to FP64.
// Fixed point: every term (the upstream product of weights and charge) is
// scaled by 2^E, the power of two with n max|q| 2^E < 2^50, and split into
// its nearest integer and the remaining fraction in units of 2^-32; both are
// summed in 64-bit integers.  Integer sums do not depend on the order of the
// additions, so the order in which particles reach a tile (the scatter does
// not fix it) cannot change rho, which is bit-identical run to run: it is the
// sum of the terms, each rounded to 2^-82 n max|q|, converted once.  FP32
// deposition keeps the integer part only (terms rounded to 2^-50 n max|q|).
constexpr int TS_T = 512, TS_P = 8192;  // threads and particles per count / scatter block
constexpr int TS_R = 1024;              // particles per scatter round (staged in LDS)
constexpr int TA_T = 512;               // threads per accumulate workgroup
constexpr int kTileAccWords = 5632;     // node accumulators of one workgroup (44 KB)
static_assert(TS_P % TS_R == 0 && TS_R % TS_T == 0, "whole rounds of whole thread steps");
static_assert(kTileMaxBins % TS_T == 0 && kTileMaxBins <= 2 * 1024, "bins per thread");

__device__ __forceinline__ void tile_mesh(TileParams& p) {
  if (p.md)
    for (int d = 0; d < 3; ++d) {
      p.mn[d] = p.md->mn[d];
      p.d[d] = p.md->d[d];
      p.dinv[d] = p.md->dinv[d];
    }
}

// the deposit's cell of a particle (all 8 nodes on the mesh) and its bin
__device__ __forceinline__ uint32_t tile_bin(const TileParams& p, double x, double y, double z,
                                             int& i0, int& j0, int& k0) {
  i0 = j0 = k0 = 0;
  const bool ok = cell_of(x, p.mn[0], p.dinv[0], 1.0, p.nlo[0], p.nn[0], i0) &&
                  cell_of(y, p.mn[1], p.dinv[1], 1.0, p.nlo[1], p.nn[1], j0) &&
                  cell_of(z, p.mn[2], p.dinv[2], 1.0, p.nlo[2], p.nn[2], k0);
  if (!ok) return uint32_t(p.ntiles);
  return uint32_t(i0 >> p.tb[0]) +
         uint32_t(p.nt[0]) * (uint32_t(j0 >> p.tb[1]) + uint32_t(p.nt[1]) * uint32_t(k0 >> p.tb[2]));
}

// Adds the wave's active lanes to cnt[d] with one LDS atomic per distinct d
// (lanes grouped by a ballot per bit of d) and returns each lane's slot: the
// counter's value before the addition plus the lane's rank among the wave's
// lanes of its d.  Every lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_count(uint32_t* cnt, uint32_t d, bool active, int bits) {
  const int lane = threadIdx.x & 63;
  uint64_t peers = __ballot(active);
  for (int b = 0; b < bits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  const int leader = peers ? __ffsll((unsigned long long)peers) - 1 : 0;
  uint32_t base = 0;
  if (active && lane == leader) base = atomicAdd(&cnt[d], uint32_t(__popcll(peers)));
  base = uint32_t(__shfl(int(base), leader));
  return base + uint32_t(__popcll(peers & ((1ull << lane) - 1ull)));
}

// exclusive scan over a block of TS_T threads of one value per thread (s: TS_T / 64 words)
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t a = __shfl_up(x, o, 64);
    if (lane >= o) x += a;
  }
  if (lane == 63) s[wave] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (int w = 0; w < wave; ++w) pre += s[w];
  __syncthreads();
  return pre + x - v;
}

// step 1: per-block bin counts hist[d * nb + b] (the layout of k_digit_scan),
// the off-mesh count added to ndrop, max |q| into tp->qmax, and with hpart the
// block's partial content hash of (x, y, z)
__global__ void __launch_bounds__(TS_T)
    k_tile_count(const double* __restrict__ x, const double* __restrict__ y,
                 const double* __restrict__ z, const double* __restrict__ q, int64_t n,
                 TileParams p, uint32_t* __restrict__ hist, int nb, uint32_t* __restrict__ ndrop,
                 TilePlan* __restrict__ tp, uint64_t* __restrict__ hpart) {
  __shared__ uint32_t h[kTileMaxBins];
  __shared__ double wmax[TS_T / 64];
  tile_mesh(p);
  for (int d = threadIdx.x; d <= p.ntiles; d += TS_T) h[d] = 0u;
  __syncthreads();
  double qm = 0.0;
  uint64_t hs = 0;
  const int64_t base = int64_t(blockIdx.x) * TS_P;
  for (int it = 0; it < TS_P / TS_T; ++it) {
    const int64_t i0w = base + int64_t(it) * TS_T;
    if (i0w >= n) break;  // uniform per block
    const int64_t i = i0w + threadIdx.x;
    const bool act = i < n;
    uint32_t d = 0;
    if (act) {
      const double xi = x[i], yi = y[i], zi = z[i];
      int a, b, c;
      d = tile_bin(p, xi, yi, zi, a, b, c);
      if (q) {
        const double aq = fabs(q[i]);
        qm = aq <= 1.7976931348623157e308 ? fmax(qm, aq) : __builtin_inf();
      }
      if (hpart) hs += particle_mix(i, xi, yi, zi);
    }
    (void)wave_count(h, d, act, p.dbits);
  }
  if (hpart) hash_block(hpart, hs);
  for (int off = 32; off > 0; off >>= 1) qm = fmax(qm, __shfl_xor(qm, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = qm;
  __syncthreads();
  for (int d = threadIdx.x; d <= p.ntiles; d += TS_T) hist[int64_t(d) * nb + blockIdx.x] = h[d];
  if (threadIdx.x == 0) {
    if (h[p.ntiles]) atomicAdd(ndrop, h[p.ntiles]);
    if (q) {
      double m = 0.0;
      for (int w = 0; w < TS_T / 64; ++w) m = fmax(m, wmax[w]);
      atomicMax(&tp->qmax, (unsigned long long)__double_as_longlong(m));
    }
  }
}

// step 2b: one workgroup of 1024 threads, thread t owning bins 2t and 2t+1
__global__ void __launch_bounds__(1024)
    k_tile_plan(const uint32_t* __restrict__ total, int ntiles, int64_t n, double q_each,
                int hasq, int chunk, TilePlan* __restrict__ tp) {
  __shared__ uint32_t s[3][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // per bin: particles, deposit chunks, gather chunks; exclusive scans of the
  // three over the bins (thread t owns bins 2t and 2t+1)
  uint32_t c[2][3];
  for (int k = 0; k < 2; ++k) {
    const int d = 2 * int(threadIdx.x) + k;
    const uint32_t cnt = d <= ntiles ? total[d] : 0u;
    c[k][0] = cnt;
    c[k][1] = d < ntiles ? (cnt + uint32_t(chunk) - 1u) / uint32_t(chunk) : 0u;
    c[k][2] = d < ntiles ? (cnt + uint32_t(kTileGatherChunk) - 1u) / uint32_t(kTileGatherChunk) : 0u;
  }
  uint32_t pre[3];
  for (int q = 0; q < 3; ++q) {
    const uint32_t own = c[0][q] + c[1][q];
    uint32_t v = own;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t a = __shfl_up(v, o, 64);
      if (lane >= o) v += a;
    }
    if (lane == 63) s[q][wave] = v;
    pre[q] = v - own;
  }
  __syncthreads();
  for (int q = 0; q < 3; ++q)
    for (int w = 0; w < wave; ++w) pre[q] += s[q][w];
  for (int k = 0; k < 2; ++k) {
    const int d = 2 * int(threadIdx.x) + k;
    if (d <= ntiles) {
      tp->bstart[d] = pre[0];
      tp->wl[d] = pre[1];
      tp->gwl[d] = pre[2];
    }
    for (int q = 0; q < 3; ++q) pre[q] += c[k][q];
  }
  if (threadIdx.x == 0) {
    const double qmax = hasq ? __longlong_as_double((long long)tp->qmax) : fabs(q_each);
    const double bound = double(n) * qmax;
    int e = 0;
    tp->bad = !(bound <= 1.7976931348623157e308);
    if (tp->bad || bound == 0.0) {
      e = 50;
    } else {
      (void)frexp(bound, &e);  // bound < 2^e
    }
    int E = 50 - e;
    E = E > 1000 ? 1000 : (E < -1000 ? -1000 : E);
    tp->scale = ldexp(1.0, E);
    tp->unscale = ldexp(1.0, -E);
  }
}

// step 3: records and particle indices in bin order.  FP64: (x, y, z, q);
// F32: the 16-byte (ab, de, gh, q) of the FP32 deposit and the linear cell
// key (the sentinel nn0 nn1 nn2 off the mesh) in sk.  The block's TS_P
// particles go in rounds of R: ranked within the round per bin, placed in
// LDS in bin order and stored so that consecutive threads write consecutive
// slots of a bin.  Rounds are as large as two workgroups per CU allow (2048
// particles, 1024 with charges in FP64; with about a thousand bins shorter
// rounds leave most bins' runs at one or two records, partial lines), the
// particle's bin packed with its index in the block.  gofs[d] is the next
// global slot of bin d for this block (the block's share of the bin starts
// at bstart[d] + col[d nb + b]).
template <int T>
__device__ __forceinline__ uint32_t block_excl_scan_t(uint32_t v, uint32_t* s) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t xv = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t a2 = __shfl_up(xv, o, 64);
    if (lane >= o) xv += a2;
  }
  if (lane == 63) s[wave] = xv;
  __syncthreads();
  uint32_t pre = 0;
  for (int w = 0; w < wave; ++w) pre += s[w];
  __syncthreads();
  return pre + xv - v;
}
constexpr int TS3_T = 512;
template <bool F32, bool HASQ>
struct ScatterGeo {
  static constexpr int R = (!F32 && HASQ) ? 1024 : 2048;  // particles per round
  static constexpr int W = F32 ? 2 : (HASQ ? 4 : 3);      // staged doubles per record
};
static_assert(TS_P <= (1 << 13) && kTileMaxBins <= (1 << 19), "bin and index share 32 bits");
template <bool F32, bool HASQ>
__global__ void __launch_bounds__(TS3_T)
    k_tile_scatter(const double* __restrict__ x, const double* __restrict__ y,
                   const double* __restrict__ z, const double* __restrict__ q, int64_t n,
                   TileParams p, const uint32_t* __restrict__ col, int nb,
                   const TilePlan* __restrict__ tp, double4* __restrict__ srec,
                   uint32_t* __restrict__ sv, uint32_t* __restrict__ sk) {
  constexpr int T = TS3_T, R = ScatterGeo<F32, HASQ>::R, W = ScatterGeo<F32, HASQ>::W;
  constexpr int PER = R / T, BPT = kTileMaxBins / T;
  static_assert(TS_P % R == 0 && R % T == 0 && kTileMaxBins % T == 0, "whole rounds and bins");
  __shared__ uint32_t gofs[kTileMaxBins], rc[kTileMaxBins];
  __shared__ uint16_t rs[kTileMaxBins];
  __shared__ double stage[R * W];
  __shared__ uint32_t sbi[R], skey[F32 ? R : 1];  // bin << 13 | index in the block
  __shared__ uint32_t sc[T / 64];
  tile_mesh(p);
  for (int d = threadIdx.x; d < kTileMaxBins; d += T) {
    gofs[d] = d <= p.ntiles ? tp->bstart[d] + col[int64_t(d) * nb + blockIdx.x] : 0u;
    rc[d] = 0u;
  }
  __syncthreads();
  const uint32_t sentinel = uint32_t(p.nn[0]) * uint32_t(p.nn[1]) * uint32_t(p.nn[2]);
  const int64_t base = int64_t(blockIdx.x) * TS_P;
  for (int r = 0; r < TS_P / R; ++r) {
    const int64_t rbase = base + int64_t(r) * R;
    if (rbase >= n) break;  // uniform per block
    const int nround = n - rbase < R ? int(n - rbase) : R;
    uint32_t dk[PER], rk[PER];
    double rec[PER][4];
    uint32_t key[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int64_t i = rbase + k * T + threadIdx.x;
      const bool act = i < n;
      dk[k] = 0;
      rec[k][0] = rec[k][1] = rec[k][2] = rec[k][3] = 0.0;
      key[k] = sentinel;
      if (act) {
        const double xi = x[i], yi = y[i], zi = z[i], qi = HASQ ? q[i] : 0.0;
        int a, b, c;
        dk[k] = tile_bin(p, xi, yi, zi, a, b, c);
        if (F32) {
          float4 w = make_float4(0.f, 0.f, 0.f, float(qi));
          if (dk[k] < uint32_t(p.ntiles)) {
            w.x = float(((p.mn[0] - xi) + double(a + p.nlo[0]) * p.d[0]) * p.dinv[0]);
            w.y = float(((p.mn[1] - yi) + double(b + p.nlo[1]) * p.d[1]) * p.dinv[1]);
            w.z = float(((p.mn[2] - zi) + double(c + p.nlo[2]) * p.d[2]) * p.dinv[2]);
            key[k] = uint32_t(a) + uint32_t(p.nn[0]) * (uint32_t(b) + uint32_t(p.nn[1]) * uint32_t(c));
          }
          rec[k][0] = __hiloint2double(__float_as_int(w.y), __float_as_int(w.x));
          rec[k][1] = __hiloint2double(__float_as_int(w.w), __float_as_int(w.z));
        } else {
          rec[k][0] = xi;
          rec[k][1] = yi;
          rec[k][2] = zi;
          rec[k][3] = qi;
        }
      }
      rk[k] = wave_count(rc, dk[k], act, p.dbits);
    }
    __syncthreads();
    // round starts of the bins: exclusive scan of the round counts
    {
      uint32_t cnt[BPT], sum = 0;
#pragma unroll
      for (int k = 0; k < BPT; ++k) {
        cnt[k] = rc[BPT * threadIdx.x + k];
        sum += cnt[k];
      }
      uint32_t pre = block_excl_scan_t<T>(sum, sc);
#pragma unroll
      for (int k = 0; k < BPT; ++k) {
        rs[BPT * threadIdx.x + k] = uint16_t(pre);
        pre += cnt[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int64_t i = rbase + k * T + threadIdx.x;
      if (i < n) {
        const uint32_t ls = rs[dk[k]] + rk[k];
#pragma unroll
        for (int w = 0; w < W; ++w) stage[W * ls + w] = rec[k][w];
        if (F32) skey[ls] = key[k];
        sbi[ls] = (dk[k] << 13) | uint32_t(i - base);
      }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nround; j += T) {
      const uint32_t d = sbi[j] >> 13;
      const uint32_t g = gofs[d] + (uint32_t(j) - rs[d]);
      if (F32) {
        reinterpret_cast<double2*>(srec)[g] = make_double2(stage[2 * j], stage[2 * j + 1]);
        sk[g] = skey[j];
      } else {
        srec[g] = make_double4(stage[W * j], stage[W * j + 1], stage[W * j + 2],
                               HASQ ? stage[W * j + 3] : 0.0);
      }
      sv[g] = uint32_t(base) + (sbi[j] & 8191u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kTileMaxBins; d += T) {
      gofs[d] += rc[d];
      rc[d] = 0u;
    }
    __syncthreads();
  }
}

// a term scaled by 2^E split into its nearest integer and the remaining
// fraction in units of 2^-32 (|t scale| < 2^51; adding 1.5 * 2^52 rounds to an
// integer, which is then the difference of the bit patterns)
__device__ __forceinline__ void fx_split(double t, double scale, unsigned long long& hi,
                                         unsigned long long& lo) {
  const double M = 6755399441055744.0;
  const long long mb = __double_as_longlong(M);
  const double m = t * scale;
  const double a = m + M;
  hi = (unsigned long long)(__double_as_longlong(a) - mb);
  const double f = m - (a - M);
  lo = (unsigned long long)(__double_as_longlong(f * 4294967296.0 + M) - mb);
}
__device__ __forceinline__ unsigned long long fx_round(double t, double scale) {
  const double M = 6755399441055744.0;
  return (unsigned long long)(__double_as_longlong(t * scale + M) - __double_as_longlong(M));
}

// work item b: chunk b - wl[t] of tile t (wl[t] <= b < wl[t+1]); -1 past the last
__device__ __forceinline__ int tile_of_item(const TilePlan* __restrict__ tp, int ntiles, int b) {
  if (b >= int(tp->wl[ntiles])) return -1;
  int lo = 0, hi = ntiles - 1;  // the last tile t with wl[t] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (int(tp->wl[mid]) <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// step 4: one workgroup per work item
template <bool F32>
__global__ void __launch_bounds__(TA_T)
    k_tile_accum(const double4* __restrict__ srec, const uint32_t* __restrict__ sk,
                 const uint32_t* __restrict__ total, const TilePlan* __restrict__ tp,
                 TileParams p, CellParams cp, int chunk, unsigned long long* __restrict__ part) {
  __shared__ unsigned long long acc[kTileAccWords];
  __shared__ int s_t;
  const int b = blockIdx.x;
  if (threadIdx.x == 0) s_t = tile_of_item(tp, p.ntiles, b);
  tile_mesh(p);
  if (cp.md)
    for (int d = 0; d < 3; ++d) {
      cp.mn[d] = cp.md->mn[d];
      cp.d[d] = cp.md->d[d];
      cp.dinv[d] = cp.md->dinv[d];
    }
  const int W = F32 ? 1 : 2;
  const int nwords = p.nodes * W;
  for (int j = threadIdx.x; j < nwords; j += TA_T) acc[j] = 0ull;
  __syncthreads();
  const int t = s_t;
  if (t < 0) return;  // uniform
  const int tx = t % p.nt[0], ty = (t / p.nt[0]) % p.nt[1], tz = t / (p.nt[0] * p.nt[1]);
  const int ox = tx << p.tb[0], oy = ty << p.tb[1], oz = tz << p.tb[2];
  const int64_t s0 = int64_t(tp->bstart[t]) + int64_t(b - int(tp->wl[t])) * chunk;
  const int64_t send = int64_t(tp->bstart[t]) + total[t];
  const int64_t s1 = s0 + chunk < send ? s0 + chunk : send;
  const double scale = tp->scale;
  const int sy = p.tn[0], sz = p.tn[0] * p.tn[1];
  for (int64_t s = s0 + threadIdx.x; s < s1; s += TA_T) {
    int i0, j0, k0;
    if (F32) {
      const float4 r = reinterpret_cast<const float4*>(srec)[s];
      const uint32_t key = sk[s];
      const uint32_t nx = uint32_t(p.nn[0]), ny = uint32_t(p.nn[1]);
      i0 = int(key % nx);
      j0 = int((key / nx) % ny);
      k0 = int(key / (nx * ny));
      const float ab = r.x, de = r.y, gh = r.z, qq = cp.hasq ? r.w : float(cp.q_each);
      float tt[8];
      tt[0] = ab * de * gh * qq;
      tt[2] = ab * (1.f - de) * gh * qq;
      tt[6] = ab * (1.f - de) * (1.f - gh) * qq;
      tt[4] = ab * de * (1.f - gh) * qq;
      tt[5] = (1.f - ab) * de * (1.f - gh) * qq;
      tt[7] = (1.f - ab) * (1.f - de) * (1.f - gh) * qq;
      tt[3] = (1.f - ab) * (1.f - de) * gh * qq;
      tt[1] = (1.f - ab) * de * gh * qq;
      const int nb0 = (i0 - ox) + sy * (j0 - oy) + sz * (k0 - oz);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int node = nb0 + (c & 1) + ((c >> 1) & 1) * sy + (c >> 2) * sz;
        atomicAdd(&acc[node], fx_round(double(tt[c]), scale));
      }
    } else {
      const double4 r = srec[s];
      (void)tile_bin(p, r.x, r.y, r.z, i0, j0, k0);
      double tt[8];
      cic_terms(r.x, r.y, r.z, cp.hasq ? r.w : cp.q_each, cp, i0, j0, k0, tt);
      const int nb0 = (i0 - ox) + sy * (j0 - oy) + sz * (k0 - oz);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int node = nb0 + (c & 1) + ((c >> 1) & 1) * sy + (c >> 2) * sz;
        unsigned long long h, l;
        fx_split(tt[c], scale, h, l);
        // integer parts, then fractions: two arrays (a node's two words
        // adjacent would put random nodes on half as many LDS banks)
        atomicAdd(&acc[node], h);
        atomicAdd(&acc[p.nodes + node], l);
      }
    }
  }
  __syncthreads();
  unsigned long long* out = part + int64_t(b) * nwords;
  for (int j = threadIdx.x; j < nwords; j += TA_T) out[j] = acc[j];
}

// step 5: node (i, j, k) of the mesh lies in the tiles a = i >> tb and, on a
// tile face (i a multiple of the tile width), a - 1, per axis
template <bool F32>
__global__ void k_tile_nodes(const unsigned long long* __restrict__ part,
                             const TilePlan* __restrict__ tp, TileParams p,
                             double* __restrict__ rho) {
  const uint32_t nx = uint32_t(p.nn[0]), ny = uint32_t(p.nn[1]);
  const uint32_t ntot = nx * ny * uint32_t(p.nn[2]);
  const int W = F32 ? 1 : 2;
  const double unscale = tp->unscale;
  const bool bad = tp->bad != 0;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ntot; g += gridDim.x * blockDim.x) {
    const int k = int(g / (nx * ny)), r = int(g - uint32_t(k) * nx * ny);
    const int j = r / int(nx), i = r - j * int(nx);
    const int ci[3] = {i, j, k};
    int ta[3][2], nta[3];
    for (int a = 0; a < 3; ++a) {
      const int t = ci[a] >> p.tb[a];
      nta[a] = 0;
      if (t < p.nt[a]) ta[a][nta[a]++] = t;
      if ((ci[a] & ((1 << p.tb[a]) - 1)) == 0 && t >= 1) ta[a][nta[a]++] = t - 1;
    }
    unsigned long long H = 0ull, L = 0ull;
    for (int u = 0; u < nta[2]; ++u)
      for (int v = 0; v < nta[1]; ++v)
        for (int w = 0; w < nta[0]; ++w) {
          const int tx = ta[0][w], ty = ta[1][v], tz = ta[2][u];
          const int t = tx + p.nt[0] * (ty + p.nt[1] * tz);
          const int local = (i - (tx << p.tb[0])) +
                            p.tn[0] * ((j - (ty << p.tb[1])) + p.tn[1] * (k - (tz << p.tb[2])));
          // four chunk blocks at a time, their loads in flight together
          // (integer sums: the order does not matter)
          const uint32_t c1 = tp->wl[t + 1];
          uint32_t c = tp->wl[t];
          const int64_t cs = int64_t(p.nodes) * W;
          for (; c + 4 <= c1; c += 4) {
            const unsigned long long* blk = part + int64_t(c) * cs + local;
            const unsigned long long h0 = blk[0], h1 = blk[cs], h2 = blk[2 * cs], h3 = blk[3 * cs];
            H += (h0 + h1) + (h2 + h3);
            if (!F32) {
              const unsigned long long l0 = blk[p.nodes], l1 = blk[cs + p.nodes],
                                       l2 = blk[2 * cs + p.nodes], l3 = blk[3 * cs + p.nodes];
              L += (l0 + l1) + (l2 + l3);
            }
          }
          for (; c < c1; ++c) {
            const unsigned long long* blk = part + int64_t(c) * cs + local;
            H += blk[0];
            if (!F32) L += blk[p.nodes];
          }
        }
    double v = double((long long)H);
    if (!F32) v = v + double((long long)L) * 2.3283064365386963e-10;  // 2^-32
    rho[g] = bad ? __builtin_nan("") : v * unscale;
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void bbox(osc_ctx* c, int64_t n, const double* x, const double* y, const double* z, double lo[3],
          double hi[3]) {
  DevBuf part(c->pool, sizeof(double) * 6 * kRedBlocks + 64);
  double* pp = part.as<double>();
  int nb = int(std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + kThreads - 1) / kThreads)));
  ProfScope ps(c, ST_DEP_BBOX);
  hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(kThreads), 0, c->stream, x, y, z, n, pp);
  OSC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kThreads), 0, c->stream, pp, nb,
                     pp + 6 * kRedBlocks, nullptr, ResizeCells{});
  OSC_HIP(hipGetLastError());
  double h[6];
  OSC_HIP(hipMemcpyAsync(h, pp + 6 * kRedBlocks, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  OSC_HIP(hipStreamSynchronize(c->stream));
  for (int a = 0; a < 3; ++a) {
    lo[a] = h[a];
    hi[a] = h[3 + a];
  }
}

void bbox_resize(osc_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                 const int cells[3], MeshDev* md) {
  DevBuf part(c->pool, sizeof(double) * 6 * kRedBlocks + 64);
  double* pp = part.as<double>();
  int nb = int(std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + kThreads - 1) / kThreads)));
  ProfScope ps(c, ST_DEP_BBOX);
  hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(kThreads), 0, c->stream, x, y, z, n, pp);
  OSC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kThreads), 0, c->stream, pp, nb,
                     pp + 6 * kRedBlocks, md, ResizeCells{{cells[0], cells[1], cells[2]}});
  OSC_HIP(hipGetLastError());
}

void device_sum_to(osc_ctx* c, const double* v, int64_t n, double* dst) {
  DevBuf part(c->pool, sizeof(double) * kRedBlocks);
  double* pp = part.as<double>();
  hipLaunchKernelGGL(k_sum_partial, dim3(kRedBlocks), dim3(kThreads), 0, c->stream, v, n, pp);
  OSC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kThreads), 0, c->stream, pp, kRedBlocks, dst);
  OSC_HIP(hipGetLastError());
}

double device_sum(osc_ctx* c, const double* v, int64_t n) {
  DevBuf part(c->pool, sizeof(double) * (kRedBlocks + 1));
  double* pp = part.as<double>();
  const int nb = kRedBlocks;
  hipLaunchKernelGGL(k_sum_partial, dim3(nb), dim3(kThreads), 0, c->stream, v, n, pp);
  OSC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kThreads), 0, c->stream, pp, nb, pp + kRedBlocks);
  OSC_HIP(hipGetLastError());
  double r = 0.0;
  OSC_HIP(hipMemcpyAsync(&r, pp + kRedBlocks, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  OSC_HIP(hipStreamSynchronize(c->stream));
  return r;
}

void fill_u32(osc_ctx* c, uint32_t* p, int64_t n, uint32_t v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(nblk(n)), dim3(kThreads), 0, c->stream, p, n, v);
  OSC_HIP(hipGetLastError());
}

// Digits of the LSD passes over nbits key bits: at most 8 bits per pass,
// split evenly (a key range needing no bits still takes one zero-bit pass).
struct SortPlan {
  int npass = 1;
  int bits[4] = {0, 0, 0, 0};
  int nb = 0;  // tiles of RS_TILE
  uint32_t mask(int p) const { return (1u << bits[p]) - 1u; }
};
static SortPlan sort_plan(int nbits, int64_t n) {
  SortPlan sp;
  sp.npass = nbits <= 0 ? 1 : (nbits + 7) / 8;
  int shift = 0;
  for (int p = 0; p < sp.npass; ++p) {
    sp.bits[p] = nbits > 0 ? (nbits - shift + (sp.npass - p) - 1) / (sp.npass - p) : 0;
    shift += sp.bits[p];
  }
  sp.nb = int((n + RS_TILE - 1) / RS_TILE);
  return sp;
}

// Sorts (key, value) pairs; values start as the input index.  hist holds
// 256 * nb counters; on entry it already holds the first pass's per-tile digit
// counts (k_cic_keys).  With pl.srec the payload (records, inverse
// permutation) is written from the sorted values at the end.
static void radix_sort(osc_ctx* c, DevBuf& ka, DevBuf& va, DevBuf& kb, DevBuf& vb, int64_t n,
                       const SortPlan& sp, DevBuf& hist, const SortPayload& pl,
                       bool& result_in_b) {
  const int nb = sp.nb;
  result_in_b = false;
  if (n == 0) return;
  DevBuf total(c->pool, sizeof(uint32_t) * 256);
  uint32_t *kin = ka.as<uint32_t>(), *vin = va.as<uint32_t>();
  uint32_t *kout = kb.as<uint32_t>(), *vout = vb.as<uint32_t>();
  int shift = 0;
  for (int p = 0; p < sp.npass; ++p) {
    const int bits = sp.bits[p];
    const uint32_t mask = sp.mask(p);
    if (p > 0) {
      hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_T), 0, c->stream, kin, n, shift, mask,
                         hist.as<uint32_t>(), nb);
      OSC_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_digit_scan, dim3(mask + 1), dim3(RS_T), 0, c->stream, hist.as<uint32_t>(),
                       nb, total.as<uint32_t>());
    OSC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_T), 0, c->stream, kin,
                       p == 0 ? nullptr : vin, kout, vout, n, shift, bits,
                       hist.as<const uint32_t>(), total.as<const uint32_t>(), nb);
    OSC_HIP(hipGetLastError());
    std::swap(kin, kout);
    std::swap(vin, vout);
    result_in_b = !result_in_b;
    shift += bits;
  }
  if (pl.srec) {
    hipLaunchKernelGGL(k_sort_payload, dim3(nblk(n)), dim3(kThreads), 0, c->stream, vin, n, pl);
    OSC_HIP(hipGetLastError());
  }
}

void queue_deposit_hash_check(osc_ctx* c, int64_t n, const double* x, const double* y,
                              const double* z) {
  char* base = static_cast<char*>(c->counters);
  auto* q = reinterpret_cast<unsigned long long*>(base + kHashQuery);
  const int nb = int(std::min<int64_t>(nblk(n), 4096));
  DevBuf part(c->pool, sizeof(uint64_t) * std::max(nb, 1));
  fill_u32(c, reinterpret_cast<uint32_t*>(q), 2, 0u);
  if (n > 0) {
    hipLaunchKernelGGL(k_hash_xyz, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream, x, y, z, n,
                       part.as<uint64_t>());
    OSC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_hash_sum, dim3(1), dim3(kThreads), 0, c->stream,
                       part.as<const uint64_t>(), nb, q);
    OSC_HIP(hipGetLastError());
  }
  static_assert(kHashQuery == kHashDeposit + 8, "hashes not adjacent");
  OSC_HIP(hipMemcpyAsync(c->hash_host, base + kHashDeposit, 2 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c->stream));
  OSC_HIP(hipEventRecord(c->hash_ev, c->stream));
}

bool deposit_hash_matches(osc_ctx* c) {
  OSC_HIP(hipEventSynchronize(c->hash_ev));
  return c->hash_host[0] == c->hash_host[1];
}

int64_t read_counter(osc_ctx* c, int idx) {
  uint32_t v = 0;
  OSC_HIP(hipMemcpyAsync(&v, static_cast<uint32_t*>(c->counters) + idx, sizeof(v),
                         hipMemcpyDeviceToHost, c->stream));
  OSC_HIP(hipStreamSynchronize(c->stream));
  return int64_t(v);
}

// Tile shape for a mesh of nn nodes: from 8^3 cells, the axis with the most
// tiles doubled until the tiles and the off-mesh bin fit kTileMaxBins bins;
// false when the node block of a tile would not fit the deposit's LDS
// accumulators or the gather's LDS node values (6 components)
static bool tile_shape(const int nn[3], bool f32, TileParams& p) {
  int cells[3];
  for (int a = 0; a < 3; ++a) {
    cells[a] = nn[a] - 1;
    if (cells[a] < 1) return false;
    p.tb[a] = 3;
  }
  for (;;) {
    int64_t nt = 1;
    for (int a = 0; a < 3; ++a) {
      p.nt[a] = (cells[a] + (1 << p.tb[a]) - 1) >> p.tb[a];
      nt *= p.nt[a];
    }
    if (nt + 1 <= kTileMaxBins) {
      p.ntiles = int(nt);
      break;
    }
    int a = 0;
    for (int b = 1; b < 3; ++b)
      if (p.nt[b] > p.nt[a]) a = b;
    if (++p.tb[a] > 8) return false;
  }
  p.nodes = 1;
  for (int a = 0; a < 3; ++a) {
    p.tn[a] = (1 << p.tb[a]) + 1;
    p.nodes *= p.tn[a];
  }
  p.words = f32 ? 1 : 2;
  if (p.nodes * p.words > kTileAccWords) return false;
  if (p.nodes * 7 * (f32 ? 4 : 8) > (f32 ? kTileGatherBytes32 : kTileGatherBytes64)) return false;
  p.dbits = 0;
  while ((1 << p.dbits) <= p.ntiles) ++p.dbits;
  return true;
}

// The deposit of deposit_cic through tiles (charge mode, not the harness rule)
static void deposit_tiled(osc_ctx* c, const DepositArgs& a, double* rho, SortedParticles* keep,
                          TileParams tp) {
  const int64_t n = a.n;
  const int64_t ntot = int64_t(a.nn[0]) * a.nn[1] * a.nn[2];
  uint32_t* cnt = static_cast<uint32_t*>(c->counters);
  const bool f32 = a.f32 != 0;
  for (int d = 0; d < 3; ++d) {
    tp.mn[d] = a.mn[d];
    tp.dinv[d] = 1.0 / a.delta[d];  // dxi = 1.d0/dx (:225-227)
    tp.d[d] = a.delta[d];
    tp.nlo[d] = a.nlo[d];
    tp.nn[d] = a.nn[d];
  }
  tp.md = a.md;
  CellParams cp{};
  for (int d = 0; d < 3; ++d) {
    cp.mn[d] = a.mn[d];
    cp.d[d] = a.delta[d];
    cp.dinv[d] = tp.dinv[d];
    cp.wofs[d] = a.nlo[d];
    cp.nn[d] = a.nn[d];
  }
  cp.md = a.md;
  cp.hasq = a.q != nullptr;
  cp.q_each = a.q_each;
  const int nbins = tp.ntiles + 1;
  const int nb = int((n + TS_P - 1) / TS_P);
  // chunks: about 512 work items for the whole bunch, 1024 to 32768 particles
  int chunk = 1024;
  while (chunk < 32768 && int64_t(chunk) * 512 < n) chunk <<= 1;
  const int64_t nwork = (n + chunk - 1) / chunk + tp.ntiles;
  const size_t rb = f32 ? sizeof(float4) : sizeof(double4);
  DevBuf plan(c->pool, sizeof(TilePlan));
  DevBuf hist(c->pool, sizeof(uint32_t) * size_t(nbins) * size_t(nb > 0 ? nb : 1));
  DevBuf total(c->pool, sizeof(uint32_t) * kTileMaxBins);
  DevBuf srec(c->pool, rb * (n + 1)), sv(c->pool, sizeof(uint32_t) * (n + 1));
  DevBuf sk(c->pool, f32 ? sizeof(uint32_t) * (n + 1) : 0);
  DevBuf part(c->pool, sizeof(unsigned long long) * size_t(nwork) * size_t(tp.nodes * tp.words));
  TilePlan* pl = plan.as<TilePlan>();
  // the max |q| word (the two last 32-bit words of the plan) starts at 0
  fill_u32(c, reinterpret_cast<uint32_t*>(&pl->qmax), 2, 0u);
  {
    ProfScope ps(c, ST_DEP_KEYS);
    DevBuf hpart(c->pool, a.hash ? sizeof(uint64_t) * size_t(nb) : 0);
    hipLaunchKernelGGL(k_tile_count, dim3(nb), dim3(TS_T), 0, c->stream, a.x, a.y, a.z, a.q, n,
                       tp, hist.as<uint32_t>(), nb, cnt, pl, hpart.as<uint64_t>());
    OSC_HIP(hipGetLastError());
    if (a.hash) {
      hipLaunchKernelGGL(k_hash_sum, dim3(1), dim3(kThreads), 0, c->stream,
                         hpart.as<const uint64_t>(), nb, a.hash);
      OSC_HIP(hipGetLastError());
    }
  }
  {
    ProfScope ps(c, ST_DEP_SORT);
    hipLaunchKernelGGL(k_digit_scan, dim3(nbins), dim3(RS_T), 0, c->stream, hist.as<uint32_t>(), nb,
                       total.as<uint32_t>());
    OSC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tile_plan, dim3(1), dim3(1024), 0, c->stream, total.as<const uint32_t>(),
                       tp.ntiles, n, a.q_each, int(a.q != nullptr), chunk, pl);
    OSC_HIP(hipGetLastError());
    uint32_t* skp = f32 ? sk.as<uint32_t>() : nullptr;
    auto scatter = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(nb), dim3(TS3_T), 0, c->stream, a.x, a.y, a.z, a.q, n, tp,
                         hist.as<const uint32_t>(), nb, pl, srec.as<double4>(), sv.as<uint32_t>(),
                         skp);
    };
    if (f32 && a.q)
      scatter(k_tile_scatter<true, true>);
    else if (f32)
      scatter(k_tile_scatter<true, false>);
    else if (a.q)
      scatter(k_tile_scatter<false, true>);
    else
      scatter(k_tile_scatter<false, false>);
    OSC_HIP(hipGetLastError());
  }
  {
    ProfScope ps(c, ST_DEP_CELLS);
    if (f32)
      hipLaunchKernelGGL(k_tile_accum<true>, dim3(unsigned(nwork)), dim3(TA_T), 0, c->stream,
                         srec.as<const double4>(), sk.as<const uint32_t>(),
                         total.as<const uint32_t>(), pl, tp, cp, chunk,
                         part.as<unsigned long long>());
    else
      hipLaunchKernelGGL(k_tile_accum<false>, dim3(unsigned(nwork)), dim3(TA_T), 0, c->stream,
                         srec.as<const double4>(), static_cast<const uint32_t*>(nullptr),
                         total.as<const uint32_t>(), pl, tp, cp, chunk,
                         part.as<unsigned long long>());
    OSC_HIP(hipGetLastError());
  }
  {
    ProfScope ps(c, ST_DEP_NODES);
    if (f32)
      hipLaunchKernelGGL(k_tile_nodes<true>, dim3(nblk(ntot)), dim3(kThreads), 0, c->stream,
                         part.as<const unsigned long long>(), pl, tp, rho);
    else
      hipLaunchKernelGGL(k_tile_nodes<false>, dim3(nblk(ntot)), dim3(kThreads), 0, c->stream,
                         part.as<const unsigned long long>(), pl, tp, rho);
    OSC_HIP(hipGetLastError());
  }
  if (keep) {
    keep->n = n;
    keep->sv = std::move(sv);
    keep->srec = std::move(srec);
    keep->f32 = f32;
    if (f32) keep->sk = std::move(sk);
    keep->tiled = true;
    keep->tile = tp;
    keep->tile.md = nullptr;
    keep->plan = std::move(plan);
    keep->total = std::move(total);
    keep->chunk = chunk;
    keep->nwork = nwork;
  }
}

void deposit_cic(osc_ctx* c, const DepositArgs& a, double* rho, SortedParticles* keep) {
  const int nx = a.nn[0], ny = a.nn[1], nz = a.nn[2];
  const int64_t ncell = int64_t(nx) * ny * nz;
  OSC_REQUIRE(ncell < (int64_t(1) << 31) && a.n < (int64_t(1) << 32) - 1, OSC_ERR_ARG,
              "grid or particle count too large for 32-bit keys");
  // counters: [0] dropped particles, [1] long runs (heavy cells); [2] is the gather's clamp count
  uint32_t* cnt = static_cast<uint32_t*>(c->counters);
  fill_u32(c, cnt, 4, 0u);
  const int64_t n = a.n;
  {
    TileParams tp{};
    if (c->tiled_deposit && !a.harness && n > 0 && tile_shape(a.nn, a.f32 != 0, tp)) {
      deposit_tiled(c, a, rho, keep, tp);
      return;
    }
  }
  // FP32 deposition (the charge-mode deposit of an FP32 context): 16-byte
  // weight records, FP32 sums
  const bool f32 = a.f32 && !a.harness;
  const size_t rb = f32 ? sizeof(float4) : sizeof(double4);
  DevBuf ka(c->pool, sizeof(uint32_t) * (n + 1)), va(c->pool, sizeof(uint32_t) * (n + 1));
  DevBuf kb(c->pool, sizeof(uint32_t) * (n + 1)), vb(c->pool, sizeof(uint32_t) * (n + 1));
  DevBuf rec(c->pool, rb * (n + 1));
  DevBuf srec(c->pool, keep ? rb * (n + 1) : 0);
  KeyParams kp{};
  CellParams cp{};
  for (int d = 0; d < 3; ++d) {
    kp.mn[d] = a.mn[d];
    kp.dinv[d] = 1.0 / a.delta[d];  // dxi = 1.d0/dx (:225-227)
    kp.base[d] = a.harness ? double(a.nlo[d]) : 1.0;
    kp.nlo[d] = a.nlo[d];
    kp.nn[d] = a.nn[d];
    kp.d[d] = a.delta[d];
    kp.wofs[d] = a.harness ? 1 : a.nlo[d];
    cp.mn[d] = a.mn[d];
    cp.d[d] = a.delta[d];
    cp.dinv[d] = kp.dinv[d];
    cp.wofs[d] = a.harness ? 1 : a.nlo[d];  // (ip-ilogbl+1)*dx  or  ip*dx
    cp.nn[d] = a.nn[d];
  }
  kp.sentinel = uint32_t(ncell);
  kp.md = a.md;
  cp.md = a.md;
  cp.hasq = a.q != nullptr;
  cp.harness = a.harness;
  cp.q_each = a.q_each;
  int nbits = 0;
  while ((uint64_t(1) << nbits) <= uint64_t(ncell)) ++nbits;
  const SortPlan sp = sort_plan(nbits, n);
  DevBuf hist(c->pool, sizeof(uint32_t) * 256 * size_t(sp.nb + 1));
  // one flag per cell, set where a cell's sums are written (S itself is not
  // cleared: only flagged cells are read); zeroed by the key kernel when its
  // grid has a thread per 4 words, else by a fill of their own
  const int64_t occ_words = (ncell + 3) / 4;
  DevBuf occ(c->pool, sizeof(uint32_t) * occ_words);
  if (n > 0 && int64_t(sp.nb) * KEY_T * 4 >= occ_words) {
    kp.clear = occ.as<uint32_t>();
    kp.nclear = occ_words;
  } else {
    ProfScope pf(c, ST_DEP_FIX);
    fill_u32(c, occ.as<uint32_t>(), occ_words, 0u);
  }
  if (n > 0) {
    ProfScope ps(c, ST_DEP_KEYS);
    DevBuf hpart(c->pool, a.hash ? sizeof(uint64_t) * sp.nb : 0);
    if (f32)
      hipLaunchKernelGGL(k_cic_keys<true>, dim3(sp.nb), dim3(KEY_T), 0, c->stream, a.x, a.y, a.z,
                         a.q, n, kp, ka.as<uint32_t>(), rec.as<double4>(), cnt, sp.mask(0),
                         hist.as<uint32_t>(), sp.nb, hpart.as<uint64_t>());
    else
      hipLaunchKernelGGL(k_cic_keys<false>, dim3(sp.nb), dim3(KEY_T), 0, c->stream, a.x, a.y, a.z,
                         a.q, n, kp, ka.as<uint32_t>(), rec.as<double4>(), cnt, sp.mask(0),
                         hist.as<uint32_t>(), sp.nb, hpart.as<uint64_t>());
    OSC_HIP(hipGetLastError());
    if (a.hash) {
      hipLaunchKernelGGL(k_hash_sum, dim3(1), dim3(kThreads), 0, c->stream,
                         hpart.as<const uint64_t>(), sp.nb, a.hash);
      OSC_HIP(hipGetLastError());
    }
  }
  bool in_b = false;
  {
    ProfScope ps(c, ST_DEP_SORT);
    radix_sort(c, ka, va, kb, vb, n, sp, hist, SortPayload{}, in_b);
  }
  const uint32_t* sk = in_b ? kb.as<uint32_t>() : ka.as<uint32_t>();
  const uint32_t* sv = in_b ? vb.as<uint32_t>() : va.as<uint32_t>();
  // corner sums of every cell (zero where no particle), wave partials of runs
  // crossing wave boundaries, list of long runs
  const int64_t nw = (n + 63) / 64 + 1;
  const size_t ab = f32 ? sizeof(float) : sizeof(double);
  DevBuf S(c->pool, ab * 8 * ncell), head(c->pool, ab * 8 * nw), tail(c->pool, ab * 8 * nw),
      heavy(c->pool, sizeof(uint32_t) * nw);
  if (n > 0) {
    ProfScope ps_cells(c, ST_DEP_CELLS);
    if (f32 && keep)
      hipLaunchKernelGGL(k_seg_reduce32<true>, dim3(nblk(n)), dim3(kThreads), 0, c->stream, sk,
                         sv, rec.as<const float4>(), srec.as<float4>(), n, cnt, float(a.q_each),
                         int(a.q != nullptr), S.as<float>(), occ.as<uint8_t>(), head.as<float>(),
                         tail.as<float>(), ncell);
    else if (f32)
      hipLaunchKernelGGL(k_seg_reduce32<false>, dim3(nblk(n)), dim3(kThreads), 0, c->stream, sk,
                         sv, rec.as<const float4>(), nullptr, n, cnt, float(a.q_each),
                         int(a.q != nullptr), S.as<float>(), occ.as<uint8_t>(), head.as<float>(),
                         tail.as<float>(), ncell);
    else if (keep)
      hipLaunchKernelGGL(k_seg_reduce<true>, dim3(nblk(n)), dim3(kThreads), 0, c->stream, sk, sv,
                         rec.as<const double4>(), srec.as<double4>(), n, cnt,
                         cp, S.as<double>(), occ.as<uint8_t>(), head.as<double>(),
                         tail.as<double>());
    else
      hipLaunchKernelGGL(k_seg_reduce<false>, dim3(nblk(n)), dim3(kThreads), 0, c->stream, sk, sv,
                         rec.as<const double4>(), nullptr, n, cnt, cp, S.as<double>(),
                         occ.as<uint8_t>(), head.as<double>(), tail.as<double>());
    OSC_HIP(hipGetLastError());
    ps_cells.stop();
    ProfScope pf(c, ST_DEP_FIX);
    if (f32) {
      hipLaunchKernelGGL(k_seg_fix<float>, dim3(nblk(nw)), dim3(kThreads), 0, c->stream, sk, n,
                         cnt, ncell, head.as<const float>(), tail.as<const float>(),
                         S.as<float>(), occ.as<uint8_t>(), heavy.as<uint32_t>(), cnt + 1);
      OSC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_seg_heavy<float>, dim3(256), dim3(kThreads), 0, c->stream, sk, n, cnt,
                         ncell, head.as<const float>(), tail.as<const float>(), S.as<float>(),
                         occ.as<uint8_t>(), heavy.as<const uint32_t>(), cnt + 1);
    } else {
      hipLaunchKernelGGL(k_seg_fix<double>, dim3(nblk(nw)), dim3(kThreads), 0, c->stream, sk, n,
                         cnt, ncell, head.as<const double>(), tail.as<const double>(),
                         S.as<double>(), occ.as<uint8_t>(), heavy.as<uint32_t>(), cnt + 1);
      OSC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_seg_heavy<double>, dim3(256), dim3(kThreads), 0, c->stream, sk, n, cnt,
                         ncell, head.as<const double>(), tail.as<const double>(),
                         S.as<double>(), occ.as<uint8_t>(), heavy.as<const uint32_t>(), cnt + 1);
    }
    OSC_HIP(hipGetLastError());
  }
  {
    ProfScope ps(c, ST_DEP_NODES);
    // fewer particles than cells (most cells empty): load the sums of flagged
    // cells only
    const int sparse = n < ncell ? 1 : 0;
    if (f32)
      hipLaunchKernelGGL(k_node<float>, dim3(nblk(ncell)), dim3(kThreads), 0, c->stream,
                         S.as<const float>(), occ.as<const uint8_t>(), nx, ny, nz, sparse,
                         a.harness, a.scale_after, rho);
    else
      hipLaunchKernelGGL(k_node<double>, dim3(nblk(ncell)), dim3(kThreads), 0, c->stream,
                         S.as<const double>(), occ.as<const uint8_t>(), nx, ny, nz, sparse,
                         a.harness, a.scale_after, rho);
    OSC_HIP(hipGetLastError());
  }
  if (keep) {
    keep->n = n;
    keep->sv = std::move(in_b ? vb : va);
    keep->srec = std::move(srec);
    keep->f32 = f32;
    if (f32) keep->sk = std::move(in_b ? kb : ka);
  }
}

void sort_for_gather(osc_ctx* c, const InterpArgs& a, int64_t n, const double* x, const double* y,
                     const double* z, SortedParticles* out) {
  const int64_t ncell = int64_t(a.nn[0]) * a.nn[1] * a.nn[2];
  OSC_REQUIRE(ncell < (int64_t(1) << 31) && n < (int64_t(1) << 32) - 1, OSC_ERR_ARG,
              "grid or point count too large for 32-bit keys");
  DevBuf ka(c->pool, sizeof(uint32_t) * (n + 1)), va(c->pool, sizeof(uint32_t) * (n + 1));
  DevBuf kb(c->pool, sizeof(uint32_t) * (n + 1)), vb(c->pool, sizeof(uint32_t) * (n + 1));
  DevBuf rec(c->pool, sizeof(double4) * (n + 1)), srec(c->pool, sizeof(double4) * (n + 1));
  KeyParams kp{};
  for (int d = 0; d < 3; ++d) {
    kp.mn[d] = a.mn[d];
    kp.dinv[d] = 1.0 / a.delta[d];
    kp.nlo[d] = a.nlo[d];
    kp.nhi[d] = a.nhi[d];
    kp.nn[d] = a.nn[d];
  }
  kp.clamp = 1;
  kp.sentinel = uint32_t(ncell);
  uint32_t* cnt = static_cast<uint32_t*>(c->counters) + 3;  // unused drop count of the clamp rule
  fill_u32(c, cnt, 1, 0u);
  int nbits = 0;
  while ((uint64_t(1) << nbits) <= uint64_t(ncell)) ++nbits;
  const SortPlan sp = sort_plan(nbits, n);
  DevBuf hist(c->pool, sizeof(uint32_t) * 256 * size_t(sp.nb + 1));
  if (n > 0) {
    hipLaunchKernelGGL(k_cic_keys<false>, dim3(sp.nb), dim3(KEY_T), 0, c->stream, x, y, z,
                       static_cast<const double*>(nullptr), n, kp, ka.as<uint32_t>(),
                       rec.as<double4>(), cnt, sp.mask(0), hist.as<uint32_t>(), sp.nb,
                       static_cast<uint64_t*>(nullptr));
    OSC_HIP(hipGetLastError());
  }
  bool in_b = false;
  SortPayload pl{rec.as<const double4>(), srec.as<double4>()};
  radix_sort(c, ka, va, kb, vb, n, sp, hist, pl, in_b);
  out->n = n;
  out->sv = std::move(in_b ? vb : va);
  out->srec = std::move(srec);
}

}  // namespace osc
