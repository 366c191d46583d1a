// Trilinear field gather: interpolate_field (open_spacecharge_mod.f90:311-390)
// for a batch of points, bit-compatible with the upstream per-point routine.
#include "osc_internal.h"

namespace osc {

namespace {
constexpr int kThreads = 256;
// batches at least this large are sorted by cell before the gather


struct IP {
  double mn[3], delta[3], hinv[3];
  int nlo[3], nhi[3], nn[3];
};

__device__ __forceinline__ int clamp_idx(double x, double mn, double hinv, int nhi, bool& fl) {
  // ip = floor((x-min)*hxi+1); clamp to [1, nhi-1] (:320-355)
  const double f = floor((x - mn) * hinv + 1.0);
  int ip;
  if (!(f >= 1.0)) {
    ip = 1;
    fl = true;
  } else if (f > double(nhi - 1)) {
    ip = nhi - 1;
    fl = true;
  } else {
    ip = int(f);
  }
  return ip;
}

__device__ __forceinline__ double interp8(const double* __restrict__ F, int64_t sy, int64_t sz,
                                          int64_t b, double ab, double de, double gh) {
  // E = sum of 8 terms in the upstream order (:369-376)
  const int64_t i1 = 1, j1 = sy, k1 = sz;
  return F[b] * ab * de * gh + F[b + j1] * ab * (1. - de) * gh +
         F[b + j1 + k1] * ab * (1. - de) * (1. - gh) + F[b + k1] * ab * de * (1. - gh) +
         F[b + i1 + k1] * (1. - ab) * de * (1. - gh) +
         F[b + i1 + j1 + k1] * (1. - ab) * (1. - de) * (1. - gh) +
         F[b + i1 + j1] * (1. - ab) * (1. - de) * gh + F[b + i1] * (1. - ab) * de * gh;
}

__global__ void k_interp(const double* __restrict__ ef, const double* __restrict__ bf, int64_t n,
                         const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, IP p, double* __restrict__ E,
                         double* __restrict__ B, int8_t* __restrict__ clamped,
                         uint32_t* __restrict__ nclamp) {
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint32_t mine = 0;
  const int64_t sy = p.nn[0], sz = int64_t(p.nn[0]) * p.nn[1], nt = sz * p.nn[2];
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double xi = x[i], yi = y[i], zi = z[i];
    bool fl = false;
    const int ip = clamp_idx(xi, p.mn[0], p.hinv[0], p.nhi[0], fl);
    const int jp = clamp_idx(yi, p.mn[1], p.hinv[1], p.nhi[1], fl);
    const int kp = clamp_idx(zi, p.mn[2], p.hinv[2], p.nhi[2], fl);
    const double ab = ((p.mn[0] - xi) + ip * p.delta[0]) * p.hinv[0];
    const double de = ((p.mn[1] - yi) + jp * p.delta[1]) * p.hinv[1];
    const double gh = ((p.mn[2] - zi) + kp * p.delta[2]) * p.hinv[2];
    // array element (ip, jp, kp) of a field with lower bound nlo; kept inside
    // the array when nlo differs from 1 (upstream assumes nlo = 1)
    int i0 = ip - p.nlo[0], j0 = jp - p.nlo[1], k0 = kp - p.nlo[2];
    i0 = i0 < 0 ? 0 : (i0 > p.nn[0] - 2 ? p.nn[0] - 2 : i0);
    j0 = j0 < 0 ? 0 : (j0 > p.nn[1] - 2 ? p.nn[1] - 2 : j0);
    k0 = k0 < 0 ? 0 : (k0 > p.nn[2] - 2 ? p.nn[2] - 2 : k0);
    const int64_t b = i0 + j0 * sy + k0 * sz;
    if (E && ef)
#pragma unroll
      for (int c = 0; c < 3; ++c) E[3 * i + c] = interp8(ef + c * nt, sy, sz, b, ab, de, gh);
    if (B && bf)
#pragma unroll
      for (int c = 0; c < 3; ++c) B[3 * i + c] = interp8(bf + c * nt, sy, sz, b, ab, de, gh);
    if (clamped) clamped[i] = fl ? 1 : 0;
    mine += fl ? 1u : 0u;
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nclamp, cnt);
}
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// node records: rec[node*RS + 0..2] = E, [3..5] = B (RS = 8, 64-byte records)
// so that the 8 nodes of a cell are 4 pairs of adjacent records.  Each wave
// reads 64 nodes of the six planar components (512 contiguous bytes per load)
// and writes their 4 KB of records through LDS as 4 contiguous 1 KB stores.
__global__ void __launch_bounds__(kThreads)
    k_pack_nodes(const double* __restrict__ ef, const double* __restrict__ bf, int64_t nt,
                 double* __restrict__ rec) {
  __shared__ double stage[kThreads / 64][64 * 9];
  const int lane = threadIdx.x & 63;
  double* w = stage[threadIdx.x >> 6];
  // wave-uniform loop over chunks of 64 nodes
  for (int64_t i0 = blockIdx.x * int64_t(blockDim.x) + (threadIdx.x & ~63); i0 < nt;
       i0 += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = i0 + lane;
    const int nn = nt - i0 < 64 ? int(nt - i0) : 64;
    if (i < nt) {
      // record of node i at 9-double stride in LDS (spreads the banks)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        w[lane * 9 + c] = ef ? ef[c * nt + i] : 0.0;
        w[lane * 9 + 3 + c] = bf ? bf[c * nt + i] : 0.0;
      }
      w[lane * 9 + 6] = 0.0;
      w[lane * 9 + 7] = 0.0;
    }
    wave_lds_sync();
    double2* o = reinterpret_cast<double2*>(rec + 8 * i0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = q * 64 + lane;  // 16-byte element of the chunk: record e / 4, part e % 4
      if ((e >> 2) < nn) {
        const double* src = w + (e >> 2) * 9 + 2 * (e & 3);
        o[e] = make_double2(src[0], src[1]);
      }
    }
    wave_lds_sync();  // read before the next chunk overwrites the stage
  }
}

__device__ __forceinline__ double wsum8(const double* __restrict__ R, int64_t sy, int64_t sz,
                                        int64_t b, int c, double ab, double de, double gh) {
  // same 8 terms and order as interp8, reading records (stride 8 doubles)
  const int64_t i1 = 8, j1 = 8 * sy, k1 = 8 * sz, o = 8 * b + c;
  return R[o] * ab * de * gh + R[o + j1] * ab * (1. - de) * gh +
         R[o + j1 + k1] * ab * (1. - de) * (1. - gh) + R[o + k1] * ab * de * (1. - gh) +
         R[o + i1 + k1] * (1. - ab) * de * (1. - gh) +
         R[o + i1 + j1 + k1] * (1. - ab) * (1. - de) * (1. - gh) +
         R[o + i1 + j1] * (1. - ab) * (1. - de) * gh + R[o + i1] * (1. - ab) * de * gh;
}

// SORTED: slot s holds particle sv[s] with record srec[s] = (x, y, z, q);
// otherwise slot s is particle s with coordinates x[s], y[s], z[s]
template <bool SORTED>
__global__ void k_interp_rec(const double* __restrict__ R, int wantE, int wantB, int64_t n,
                             const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double4* __restrict__ srec,
                             const uint32_t* __restrict__ sv, IP p, double* __restrict__ E,
                             double* __restrict__ B, int8_t* __restrict__ clamped,
                             uint32_t* __restrict__ nclamp) {
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint32_t mine = 0;
  const int64_t sy = p.nn[0], sz = int64_t(p.nn[0]) * p.nn[1];
  for (int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; s < n;
       s += int64_t(gridDim.x) * blockDim.x) {
    int64_t i;
    double xi, yi, zi;
    if (SORTED) {
      i = int64_t(sv[s]);
      const double4 r = srec[s];
      xi = r.x;
      yi = r.y;
      zi = r.z;
    } else {
      i = s;
      xi = x[s];
      yi = y[s];
      zi = z[s];
    }
    bool fl = false;
    const int ip = clamp_idx(xi, p.mn[0], p.hinv[0], p.nhi[0], fl);
    const int jp = clamp_idx(yi, p.mn[1], p.hinv[1], p.nhi[1], fl);
    const int kp = clamp_idx(zi, p.mn[2], p.hinv[2], p.nhi[2], fl);
    const double ab = ((p.mn[0] - xi) + ip * p.delta[0]) * p.hinv[0];
    const double de = ((p.mn[1] - yi) + jp * p.delta[1]) * p.hinv[1];
    const double gh = ((p.mn[2] - zi) + kp * p.delta[2]) * p.hinv[2];
    int i0 = ip - p.nlo[0], j0 = jp - p.nlo[1], k0 = kp - p.nlo[2];
    i0 = i0 < 0 ? 0 : (i0 > p.nn[0] - 2 ? p.nn[0] - 2 : i0);
    j0 = j0 < 0 ? 0 : (j0 > p.nn[1] - 2 ? p.nn[1] - 2 : j0);
    k0 = k0 < 0 ? 0 : (k0 > p.nn[2] - 2 ? p.nn[2] - 2 : k0);
    const int64_t b = i0 + j0 * sy + k0 * sz;
    if (wantE)
#pragma unroll
      for (int c = 0; c < 3; ++c) E[3 * i + c] = wsum8(R, sy, sz, b, c, ab, de, gh);
    if (wantB)
#pragma unroll
      for (int c = 0; c < 3; ++c) B[3 * i + c] = wsum8(R, sy, sz, b, 3 + c, ab, de, gh);
    if (clamped) clamped[i] = fl ? 1 : 0;
    mine += fl ? 1u : 0u;
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nclamp, cnt);
}

// wsum8 over node records staged in LDS: row r = dj + 2 dk holds the records
// of nodes q0 + 0..kNodeWin-1 (6 doubles each) of that row offset, q = the
// cell's position in the rows; same 8 terms in the same order
constexpr int kNodeWin = 16;
__device__ __forceinline__ double wsum8_lds(const double* st, int q, int c, double ab, double de,
                                            double gh) {
  auto N = [&](int r, int d) { return st[((r * kNodeWin) + q + d) * 6 + c]; };
  return N(0, 0) * ab * de * gh + N(1, 0) * ab * (1. - de) * gh +
         N(3, 0) * ab * (1. - de) * (1. - gh) + N(2, 0) * ab * de * (1. - gh) +
         N(2, 1) * (1. - ab) * de * (1. - gh) + N(3, 1) * (1. - ab) * (1. - de) * (1. - gh) +
         N(1, 1) * (1. - ab) * (1. - de) * gh + N(0, 1) * (1. - ab) * de * gh;
}

// cell-ordered gather: slot s (record srec[s], particle sv[s]) leaves
// (E, B, clamped, 0) in the 64-byte record out8[sv[s]]; the values are those
// of k_interp_rec (same terms, same order).  When the cells of a wave's 64
// particles lie within kNodeWin - 1 consecutive cells of one row (96 % of
// the waves of a 10^7-particle Gaussian bunch on 128^3), the wave loads the 4
// node rows it needs once, one record per lane, into LDS and every lane sums
// from there: one lane per node record instead of 8 record reads per lane
// (the gather is bound by the vector-memory data path, not HBM).  Other waves
// read the records directly.  Each wave then stages its 64 results in LDS
// (9-double stride) and stores each record with four lanes (16 bytes each)
// at its particle's original position: a random but whole 64-byte line per
// record, about as costly as the random read it replaces in the return to
// particle order, which then streams (k_unpack).
constexpr int kRecPad = 9;
constexpr int kUnpackOccLds = 17 * 1024;
static_assert(4 * kNodeWin * 6 <= 64 * kRecPad, "node rows exceed the wave's stage");
__global__ void __launch_bounds__(kThreads)
    k_interp_cells(const double* __restrict__ R, int wantE, int wantB, int64_t n,
                   const double4* __restrict__ srec, const uint32_t* __restrict__ sv, IP p,
                   double* __restrict__ out8, uint32_t* __restrict__ nclamp) {
  __shared__ uint32_t cnt;
  __shared__ double stage[kThreads / 64][64 * kRecPad];
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint32_t mine = 0;
  const int lane = threadIdx.x & 63;
  double* st = stage[threadIdx.x >> 6];
  const int64_t sy = p.nn[0], sz = int64_t(p.nn[0]) * p.nn[1];
  // wave-uniform loop over chunks of 64 slots
  for (int64_t wb = blockIdx.x * int64_t(blockDim.x) + (threadIdx.x & ~63); wb < n;
       wb += int64_t(gridDim.x) * blockDim.x) {
    const int64_t s = wb + lane;
    double v[6] = {0, 0, 0, 0, 0, 0};
    bool fl = false;
    double ab = 0, de = 0, gh = 0;
    int64_t b = 0;
    if (s < n) {
      const double4 r = srec[s];
      const double xi = r.x, yi = r.y, zi = r.z;
      const int ip = clamp_idx(xi, p.mn[0], p.hinv[0], p.nhi[0], fl);
      const int jp = clamp_idx(yi, p.mn[1], p.hinv[1], p.nhi[1], fl);
      const int kp = clamp_idx(zi, p.mn[2], p.hinv[2], p.nhi[2], fl);
      ab = ((p.mn[0] - xi) + ip * p.delta[0]) * p.hinv[0];
      de = ((p.mn[1] - yi) + jp * p.delta[1]) * p.hinv[1];
      gh = ((p.mn[2] - zi) + kp * p.delta[2]) * p.hinv[2];
      int i0 = ip - p.nlo[0], j0 = jp - p.nlo[1], k0 = kp - p.nlo[2];
      i0 = i0 < 0 ? 0 : (i0 > p.nn[0] - 2 ? p.nn[0] - 2 : i0);
      j0 = j0 < 0 ? 0 : (j0 > p.nn[1] - 2 ? p.nn[1] - 2 : j0);
      k0 = k0 < 0 ? 0 : (k0 > p.nn[2] - 2 ? p.nn[2] - 2 : k0);
      b = i0 + j0 * sy + k0 * sz;
      mine += fl ? 1u : 0u;
    }
    // the wave's cell range (lanes past n take the first lane's cell)
    const long long b0 = __shfl(static_cast<long long>(b), 0);
    long long lo = s < n ? static_cast<long long>(b) : b0, hi = lo;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const long long ol = __shfl_xor(lo, off), oh = __shfl_xor(hi, off);
      lo = ol < lo ? ol : lo;
      hi = oh > hi ? oh : hi;
    }
    if (hi - lo <= kNodeWin - 2) {
      // lane l loads node lo + (l % kNodeWin) of row l / kNodeWin, up to the
      // last node the wave's cells use (hi + 1 in each row)
      const int r = lane / kNodeWin, q = lane % kNodeWin;
      if (q <= int(hi - lo) + 1) {
        const int64_t node = lo + q + ((r & 1) ? sy : 0) + ((r & 2) ? sz : 0);
        const double* src = R + 8 * node;
#pragma unroll
        for (int c = 0; c < 6; ++c) st[lane * 6 + c] = src[c];
      }
      wave_lds_sync();
      if (s < n) {
        const int q0 = int(b - lo);
        if (wantE)
#pragma unroll
          for (int c = 0; c < 3; ++c) v[c] = wsum8_lds(st, q0, c, ab, de, gh);
        if (wantB)
#pragma unroll
          for (int c = 0; c < 3; ++c) v[3 + c] = wsum8_lds(st, q0, 3 + c, ab, de, gh);
      }
      wave_lds_sync();  // the node rows are read before the results overwrite them
    } else if (s < n) {
      if (wantE)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = wsum8(R, sy, sz, b, c, ab, de, gh);
      if (wantB)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[3 + c] = wsum8(R, sy, sz, b, 3 + c, ab, de, gh);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) st[lane * kRecPad + c] = v[c];
    st[lane * kRecPad + 6] = fl ? 1.0 : 0.0;
    st[lane * kRecPad + 7] = 0.0;
    const uint32_t dst = s < n ? sv[s] : 0u;
    // LDS operations of one wave complete in order; the fences keep the
    // compiler from moving the accesses across (records written above are read
    // by other lanes below, and rewritten in the next chunk)
    wave_lds_sync();
    // round j: lanes 4k..4k+3 store the 4 parts of record 16 j + k
    const int part = lane & 3;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 16 * j + (lane >> 2);
      const uint32_t d = uint32_t(__shfl(int(dst), r));
      if (wb + r < n) {
        const double* sr = st + r * kRecPad + 2 * part;
        reinterpret_cast<double2*>(out8 + 8 * int64_t(d))[part] = make_double2(sr[0], sr[1]);
      }
    }
    wave_lds_sync();
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nclamp, cnt);
}

// ---------------------------------------------------------------------------
// FP32 particle path (the fused step of an FP32 context)
// ---------------------------------------------------------------------------
// node records in FP32: (E0, E1, E2, B0, B1, B2, 0, 0), 32 bytes per node
__global__ void __launch_bounds__(kThreads)
    k_pack_nodes32(const double* __restrict__ ef, const double* __restrict__ bf, int64_t nt,
                   float* __restrict__ rec) {
  __shared__ float stage[kThreads / 64][64 * 9];
  const int lane = threadIdx.x & 63;
  float* w = stage[threadIdx.x >> 6];
  for (int64_t i0 = blockIdx.x * int64_t(blockDim.x) + (threadIdx.x & ~63); i0 < nt;
       i0 += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = i0 + lane;
    const int nn = nt - i0 < 64 ? int(nt - i0) : 64;
    if (i < nt) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        w[lane * 9 + c] = ef ? float(ef[c * nt + i]) : 0.f;
        w[lane * 9 + 3 + c] = bf ? float(bf[c * nt + i]) : 0.f;
      }
      w[lane * 9 + 6] = 0.f;
      w[lane * 9 + 7] = 0.f;
    }
    wave_lds_sync();
    float4* o = reinterpret_cast<float4*>(rec + 8 * i0);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = q * 64 + lane;  // 16-byte element: record e / 2, half e % 2
      if ((e >> 1) < nn) {
        const float* src = w + (e >> 1) * 9 + 4 * (e & 1);
        o[e] = make_float4(src[0], src[1], src[2], src[3]);
      }
    }
    wave_lds_sync();
  }
}

template <typename A, typename F>
__device__ __forceinline__ A wsum8_any(F&& N, A ab, A de, A gh) {
  const A one = A(1);
  return N(0, 0) * ab * de * gh + N(1, 0) * ab * (one - de) * gh +
         N(3, 0) * ab * (one - de) * (one - gh) + N(2, 0) * ab * de * (one - gh) +
         N(2, 1) * (one - ab) * de * (one - gh) + N(3, 1) * (one - ab) * (one - de) * (one - gh) +
         N(1, 1) * (one - ab) * (one - de) * gh + N(0, 1) * (one - ab) * de * gh;
}

// The cell-ordered gather of k_interp_cells in FP32: the cell is the deposit's
// sorted key and the weights its FP32 records (the deposit and interpolate
// rules agree for every particle on the mesh); a dropped particle (sentinel
// key, off the mesh) takes the clamp rule from its FP64 coordinates.  Node
// rows staged in LDS as in k_interp_cells; the results (E, B, clamped, 0)
// leave as the 64-byte FP64 records of k_interp_cells at the particle's
// original position (measured: 32-byte FP32 records at random places,
// 0.34 ms for 10^7 points, are slower to store than 64-byte ones), for
// k_unpack.
__global__ void __launch_bounds__(kThreads)
    k_interp_cells32(const float* __restrict__ R, int wantE, int wantB, int64_t n,
                     const float4* __restrict__ srec, const uint32_t* __restrict__ sk,
                     const uint32_t* __restrict__ sv, const double* __restrict__ x,
                     const double* __restrict__ y, const double* __restrict__ z, IP p,
                     double* __restrict__ out8, uint32_t* __restrict__ nclamp) {
  __shared__ uint32_t cnt;
  __shared__ float stage[kThreads / 64][64 * 6];
  __shared__ double ostage[kThreads / 64][64 * kRecPad];
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint32_t mine = 0;
  const int lane = threadIdx.x & 63;
  float* st = stage[threadIdx.x >> 6];
  double* ot = ostage[threadIdx.x >> 6];
  const int64_t sy = p.nn[0], sz = int64_t(p.nn[0]) * p.nn[1];
  const uint32_t ncell = uint32_t(p.nn[0]) * uint32_t(p.nn[1]) * uint32_t(p.nn[2]);
  for (int64_t wb = blockIdx.x * int64_t(blockDim.x) + (threadIdx.x & ~63); wb < n;
       wb += int64_t(gridDim.x) * blockDim.x) {
    const int64_t s = wb + lane;
    float v[6] = {0, 0, 0, 0, 0, 0};
    bool fl = false;
    float ab = 0, de = 0, gh = 0;
    int64_t b = 0;
    if (s < n) {
      const uint32_t key = sk[s];
      if (key < ncell) {
        const float4 r = srec[s];
        ab = r.x;
        de = r.y;
        gh = r.z;
        b = int64_t(key);
      } else {
        const uint32_t i = sv[s];
        const double xi = x[i], yi = y[i], zi = z[i];
        const int ip = clamp_idx(xi, p.mn[0], p.hinv[0], p.nhi[0], fl);
        const int jp = clamp_idx(yi, p.mn[1], p.hinv[1], p.nhi[1], fl);
        const int kp = clamp_idx(zi, p.mn[2], p.hinv[2], p.nhi[2], fl);
        ab = float(((p.mn[0] - xi) + ip * p.delta[0]) * p.hinv[0]);
        de = float(((p.mn[1] - yi) + jp * p.delta[1]) * p.hinv[1]);
        gh = float(((p.mn[2] - zi) + kp * p.delta[2]) * p.hinv[2]);
        int i0 = ip - p.nlo[0], j0 = jp - p.nlo[1], k0 = kp - p.nlo[2];
        i0 = i0 < 0 ? 0 : (i0 > p.nn[0] - 2 ? p.nn[0] - 2 : i0);
        j0 = j0 < 0 ? 0 : (j0 > p.nn[1] - 2 ? p.nn[1] - 2 : j0);
        k0 = k0 < 0 ? 0 : (k0 > p.nn[2] - 2 ? p.nn[2] - 2 : k0);
        b = i0 + j0 * sy + k0 * sz;
      }
      mine += fl ? 1u : 0u;
    }
    const long long b0 = __shfl(static_cast<long long>(b), 0);
    long long lo = s < n ? static_cast<long long>(b) : b0, hi = lo;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const long long ol = __shfl_xor(lo, off), oh = __shfl_xor(hi, off);
      lo = ol < lo ? ol : lo;
      hi = oh > hi ? oh : hi;
    }
    if (hi - lo <= kNodeWin - 2) {
      const int r = lane / kNodeWin, q = lane % kNodeWin;
      if (q <= int(hi - lo) + 1) {
        const int64_t node = lo + q + ((r & 1) ? sy : 0) + ((r & 2) ? sz : 0);
        const float4* src = reinterpret_cast<const float4*>(R + 8 * node);
        const float4 e0 = src[0], e1 = src[1];
        st[lane * 6 + 0] = e0.x;
        st[lane * 6 + 1] = e0.y;
        st[lane * 6 + 2] = e0.z;
        st[lane * 6 + 3] = e0.w;
        st[lane * 6 + 4] = e1.x;
        st[lane * 6 + 5] = e1.y;
      }
      wave_lds_sync();
      if (s < n) {
        // compile-time component indices keep v in registers
        const int q0 = int(b - lo);
#pragma unroll
        for (int c = 0; c < 6; ++c)
          if (c < 3 ? wantE : wantB)
            v[c] = wsum8_any<float>(
                [&](int rr, int d) { return st[((rr * kNodeWin) + q0 + d) * 6 + c]; }, ab, de, gh);
      }
      wave_lds_sync();
    } else if (s < n) {
      const int64_t i1 = 8, j1 = 8 * sy, k1 = 8 * sz;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        if (!(c < 3 ? wantE : wantB)) continue;
        const int64_t o = 8 * b + c;
        auto N = [&](int rr, int d) {
          return R[o + ((rr & 1) ? j1 : 0) + ((rr & 2) ? k1 : 0) + (d ? i1 : 0)];
        };
        v[c] = wsum8_any<float>(N, ab, de, gh);
      }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) ot[lane * kRecPad + c] = double(v[c]);
    ot[lane * kRecPad + 6] = fl ? 1.0 : 0.0;
    ot[lane * kRecPad + 7] = 0.0;
    const uint32_t dst = s < n ? sv[s] : 0u;
    wave_lds_sync();
    // round j: lanes 4k..4k+3 store the 4 parts of record 16 j + k
    const int part = lane & 3;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 16 * j + (lane >> 2);
      const uint32_t d = uint32_t(__shfl(int(dst), r));
      if (wb + r < n) {
        const double* sr = ot + r * kRecPad + 2 * part;
        reinterpret_cast<double2*>(out8 + 8 * int64_t(d))[part] = make_double2(sr[0], sr[1]);
      }
    }
    wave_lds_sync();
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nclamp, cnt);
}

// The cell-ordered gather of a step with fewer particles than twice the
// nodes: each particle reads its 8 nodes from the planar fields where they
// lie (neighbours in cell order share their lines), so that no node records
// of the whole mesh are packed.
// Cells and weights as in k_interp_cells (FP64: from the record's
// coordinates) and k_interp_cells32 (F32: the deposit's key and FP32
// weights; off the mesh, the clamp rule from the coordinates), the same 8
// terms in the same order over the same node values (F32: rounded to float
// as the packed records hold them): the results are those of the packed
// kernels.  (E, B, clamped, 0) leave as the 64-byte record out8[sv[s]] for
// k_unpack.
template <bool F32>
__global__ void __launch_bounds__(kThreads)
    k_interp_planar(const double* __restrict__ ef, const double* __restrict__ bf, int wantE,
                    int wantB, int64_t n, const double4* __restrict__ srec,
                    const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sv,
                    const double* __restrict__ x, const double* __restrict__ y,
                    const double* __restrict__ z, IP p, double* __restrict__ out8,
                    uint32_t* __restrict__ nclamp) {
  using A = typename std::conditional<F32, float, double>::type;
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint32_t mine = 0;
  const int64_t sy = p.nn[0], sz = int64_t(p.nn[0]) * p.nn[1], nt = sz * p.nn[2];
  const uint32_t ncell = uint32_t(nt);
  for (int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; s < n;
       s += int64_t(gridDim.x) * blockDim.x) {
    bool fl = false;
    A ab, de, gh;
    int64_t b;
    const uint32_t key = F32 ? sk[s] : 0u;
    if (F32 && key < ncell) {
      const float4 r = reinterpret_cast<const float4*>(srec)[s];
      ab = A(r.x);
      de = A(r.y);
      gh = A(r.z);
      b = int64_t(key);
    } else {
      double xi, yi, zi;
      if (F32) {
        const uint32_t i = sv[s];
        xi = x[i];
        yi = y[i];
        zi = z[i];
      } else {
        const double4 r = srec[s];
        xi = r.x;
        yi = r.y;
        zi = r.z;
      }
      const int ip = clamp_idx(xi, p.mn[0], p.hinv[0], p.nhi[0], fl);
      const int jp = clamp_idx(yi, p.mn[1], p.hinv[1], p.nhi[1], fl);
      const int kp = clamp_idx(zi, p.mn[2], p.hinv[2], p.nhi[2], fl);
      ab = A(((p.mn[0] - xi) + ip * p.delta[0]) * p.hinv[0]);
      de = A(((p.mn[1] - yi) + jp * p.delta[1]) * p.hinv[1]);
      gh = A(((p.mn[2] - zi) + kp * p.delta[2]) * p.hinv[2]);
      int i0 = ip - p.nlo[0], j0 = jp - p.nlo[1], k0 = kp - p.nlo[2];
      i0 = i0 < 0 ? 0 : (i0 > p.nn[0] - 2 ? p.nn[0] - 2 : i0);
      j0 = j0 < 0 ? 0 : (j0 > p.nn[1] - 2 ? p.nn[1] - 2 : j0);
      k0 = k0 < 0 ? 0 : (k0 > p.nn[2] - 2 ? p.nn[2] - 2 : k0);
      b = i0 + j0 * sy + k0 * sz;
    }
    mine += fl ? 1u : 0u;
    A v[6] = {A(0), A(0), A(0), A(0), A(0), A(0)};
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (c < 3 ? wantE : wantB) {
        const double* F = (c < 3 ? ef : bf) + (c % 3) * nt;
        v[c] = wsum8_any<A>(
            [&](int r2, int d) {
              return A(F[b + d + ((r2 & 1) ? sy : 0) + ((r2 & 2) ? sz : 0)]);
            },
            ab, de, gh);
      }
    double2* o = reinterpret_cast<double2*>(out8 + 8 * int64_t(sv[s]));
    o[0] = make_double2(double(v[0]), double(v[1]));
    o[1] = make_double2(double(v[2]), double(v[3]));
    o[2] = make_double2(double(v[4]), double(v[5]));
    o[3] = make_double2(fl ? 1.0 : 0.0, 0.0);
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nclamp, cnt);
}

// original order: the 64-byte records out8[i] = (E0, E1, E2, B0, B1, B2,
// flag, pad) into E(3,n), B(3,n) and the clamp flags.  Each wave moves 64
// records through LDS: read as 4 contiguous KB, written as contiguous runs of
// E and B (512 bytes per store instruction) and 64 flag bytes.
__global__ void __launch_bounds__(kThreads)
    k_unpack(const double* __restrict__ out8, int64_t n, double* __restrict__ E,
             double* __restrict__ B, int8_t* __restrict__ clamped) {
  __shared__ double stage[kThreads / 64][64 * 8];
  const int lane = threadIdx.x & 63;
  double* w = stage[threadIdx.x >> 6];
  const double2* in2 = reinterpret_cast<const double2*>(out8);
  // wave-uniform loop over chunks of 64 records
  for (int64_t i0 = blockIdx.x * int64_t(blockDim.x) + (threadIdx.x & ~63); i0 < n;
       i0 += int64_t(gridDim.x) * blockDim.x) {
    const int nrec = n - i0 < 64 ? int(n - i0) : 64;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = lane + 64 * q;  // 16-byte element of the chunk, record e / 4
      if ((e >> 2) < nrec) {
        const double2 v = in2[4 * i0 + e];
        w[2 * e] = v.x;
        w[2 * e + 1] = v.y;
      }
    }
    wave_lds_sync();
    if (E)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int e = lane + 64 * k;  // element of the chunk's E: record e / 3, component e % 3
        if (e / 3 < nrec) E[3 * i0 + e] = w[(e / 3) * 8 + e % 3];
      }
    if (B)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int e = lane + 64 * k;
        if (e / 3 < nrec) B[3 * i0 + e] = w[(e / 3) * 8 + 3 + e % 3];
      }
    if (clamped && lane < nrec) clamped[i0 + lane] = w[lane * 8 + 6] != 0.0 ? 1 : 0;
    wave_lds_sync();  // read before the next chunk overwrites the stage
  }
}

// ---------------------------------------------------------------------------
// tile gather: the particles of the tiled deposit, in bin order
// ---------------------------------------------------------------------------
constexpr int TG_T = 1024;

// One workgroup per gather item (kTileGatherChunk particles of one tile's
// bin): the E and B values of the tile's nodes are staged in LDS as node
// records of 7 values (FP64; F32: the values of the FP32 node records; the
// odd stride in 4-byte words spreads random nodes over the LDS banks), then
// each particle sums its 8 nodes from there in the upstream
// order and leaves (E, B, clamped, 0) in the 64-byte record out8[sv[s]] for
// k_unpack.  FP64: cell and weights from the coordinates by the
// interpolate_field rule, the values of k_interp_cells; F32: the deposit's cell
// key and FP32 weights, the values of k_interp_cells32.  Tiles holding fewer
// than kTileGatherStage particles, and particles whose cell lies outside the
// tile (a gather mesh other than the deposit's), read the planar fields.
constexpr int TG_PER = kTileGatherChunk / TG_T;
constexpr int kTileGatherStage = 512;
static_assert(kTileGatherChunk % TG_T == 0, "whole particles per thread");
template <bool F32>
__global__ void __launch_bounds__(TG_T)
    k_tile_gather(const double* __restrict__ ef, const double* __restrict__ bf, int wantE,
                  int wantB, TileParams tg, const TilePlan* __restrict__ tp,
                  const uint32_t* __restrict__ total, const double4* __restrict__ srec,
                  const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sv, IP p,
                  double* __restrict__ out8, uint32_t* __restrict__ nclamp) {
  using A = typename std::conditional<F32, float, double>::type;
  constexpr int NS = 7;  // LDS values per node record
  __shared__ double lds_raw[(F32 ? kTileGatherBytes32 : kTileGatherBytes64) / 8];
  __shared__ int s_t;
  __shared__ uint32_t cnt;
  A* L = reinterpret_cast<A*>(lds_raw);
  if (threadIdx.x == 0) {
    int t = -1;
    const int b = blockIdx.x, ntiles = tg.ntiles;
    if (b < int(tp->gwl[ntiles])) {
      int lo = 0, hi = ntiles - 1;  // the last tile t with gwl[t] <= b
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (int(tp->gwl[mid]) <= b) lo = mid;
        else hi = mid - 1;
      }
      t = lo;
    }
    s_t = t;
    cnt = 0;
  }
  __syncthreads();
  const int t = s_t;
  if (t < 0) return;  // uniform
  const int tx = t % tg.nt[0], ty = (t / tg.nt[0]) % tg.nt[1], tz = t / (tg.nt[0] * tg.nt[1]);
  const int ox = tx << tg.tb[0], oy = ty << tg.tb[1], oz = tz << tg.tb[2];
  const int nodes = tg.nodes, tn0 = tg.tn[0], tn01 = tg.tn[0] * tg.tn[1];
  const int64_t sy = p.nn[0], sz = int64_t(p.nn[0]) * p.nn[1], nt = sz * p.nn[2];
  const uint32_t count = total[t];
  const bool staged = count >= uint32_t(kTileGatherStage);  // uniform
  if (staged) {
    for (int node = threadIdx.x; node < nodes; node += TG_T) {
      const int lk = node / tn01, rr = node - lk * tn01, lj = rr / tn0, li = rr - lj * tn0;
      const int gi = ox + li, gj = oy + lj, gk = oz + lk;
      const bool in = gi < p.nn[0] && gj < p.nn[1] && gk < p.nn[2];
      const int64_t g = in ? gi + sy * gj + sz * gk : 0;
      A v[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double* F = c < 3 ? ef : bf;
        v[c] = A((in && (c < 3 ? wantE : wantB)) ? F[(c % 3) * nt + g] : 0.0);
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) L[NS * node + c] = v[c];
    }
    __syncthreads();
  }
  const int64_t s0 = int64_t(tp->bstart[t]) +
                     int64_t(int(blockIdx.x) - int(tp->gwl[t])) * kTileGatherChunk;
  const int64_t send = int64_t(tp->bstart[t]) + count;
  const int wx = 1 << tg.tb[0], wy = 1 << tg.tb[1], wz = 1 << tg.tb[2];
  // the thread's particles: records first, so that their loads overlap
  double4 rec[TG_PER];
  uint32_t key[TG_PER], dst[TG_PER];
#pragma unroll
  for (int k = 0; k < TG_PER; ++k) {
    const int64_t s = s0 + k * TG_T + threadIdx.x;
    rec[k] = make_double4(0.0, 0.0, 0.0, 0.0);
    key[k] = 0;
    dst[k] = 0;
    if (s < send) {
      if (F32) {
        const float4 r = reinterpret_cast<const float4*>(srec)[s];
        rec[k] = make_double4(r.x, r.y, r.z, 0.0);
        key[k] = sk[s];
      } else {
        rec[k] = srec[s];
      }
      dst[k] = sv[s];
    }
  }
  uint32_t mine = 0;
#pragma unroll
  for (int k = 0; k < TG_PER; ++k) {
    const int64_t s = s0 + k * TG_T + threadIdx.x;
    if (s >= send) continue;
    A ab, de, gh;
    int i0, j0, k0;
    bool fl = false;
    if (F32) {
      const uint32_t nx = uint32_t(p.nn[0]), ny = uint32_t(p.nn[1]);
      i0 = int(key[k] % nx);
      j0 = int((key[k] / nx) % ny);
      k0 = int(key[k] / (nx * ny));
      ab = A(rec[k].x);
      de = A(rec[k].y);
      gh = A(rec[k].z);
    } else {
      const double xr = rec[k].x, yr = rec[k].y, zr = rec[k].z;
      const int ip = clamp_idx(xr, p.mn[0], p.hinv[0], p.nhi[0], fl);
      const int jp = clamp_idx(yr, p.mn[1], p.hinv[1], p.nhi[1], fl);
      const int kp = clamp_idx(zr, p.mn[2], p.hinv[2], p.nhi[2], fl);
      ab = A(((p.mn[0] - xr) + ip * p.delta[0]) * p.hinv[0]);
      de = A(((p.mn[1] - yr) + jp * p.delta[1]) * p.hinv[1]);
      gh = A(((p.mn[2] - zr) + kp * p.delta[2]) * p.hinv[2]);
      i0 = ip - p.nlo[0];
      j0 = jp - p.nlo[1];
      k0 = kp - p.nlo[2];
      i0 = i0 < 0 ? 0 : (i0 > p.nn[0] - 2 ? p.nn[0] - 2 : i0);
      j0 = j0 < 0 ? 0 : (j0 > p.nn[1] - 2 ? p.nn[1] - 2 : j0);
      k0 = k0 < 0 ? 0 : (k0 > p.nn[2] - 2 ? p.nn[2] - 2 : k0);
    }
    mine += fl ? 1u : 0u;
    const int li = i0 - ox, lj = j0 - oy, lk = k0 - oz;
    A v[6] = {A(0), A(0), A(0), A(0), A(0), A(0)};
    if (staged && li >= 0 && li < wx && lj >= 0 && lj < wy && lk >= 0 && lk < wz) {
      const int q0 = li + tn0 * lj + tn01 * lk;
#pragma unroll
      for (int c = 0; c < 6; ++c)
        if (c < 3 ? wantE : wantB)
          v[c] = wsum8_any<A>(
              [&](int r2, int d) {
                return L[NS * (q0 + d + ((r2 & 1) ? tn0 : 0) + ((r2 & 2) ? tn01 : 0)) + c];
              },
              ab, de, gh);
    } else {
      const int64_t b = i0 + sy * j0 + sz * k0;
#pragma unroll
      for (int c = 0; c < 6; ++c)
        if (c < 3 ? wantE : wantB) {
          const double* F = (c < 3 ? ef : bf) + (c % 3) * nt;
          v[c] = wsum8_any<A>(
              [&](int r2, int d) {
                return A(F[b + d + ((r2 & 1) ? sy : 0) + ((r2 & 2) ? sz : 0)]);
              },
              ab, de, gh);
        }
    }
    double2* o = reinterpret_cast<double2*>(out8 + 8 * int64_t(dst[k]));
    o[0] = make_double2(double(v[0]), double(v[1]));
    o[1] = make_double2(double(v[2]), double(v[3]));
    o[2] = make_double2(double(v[4]), double(v[5]));
    o[3] = make_double2(fl ? 1.0 : 0.0, 0.0);
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nclamp, cnt);
}

// The off-mesh bin of the tiled deposit (slots bstart[ntiles] on): the
// interpolate_field clamp rule from the coordinates (FP64: the record; F32:
// x, y, z of the particle, the weights rounded as in k_interp_cells32),
// the planar fields read directly.
template <bool F32>
__global__ void __launch_bounds__(kThreads)
    k_tile_gather_off(const double* __restrict__ ef, const double* __restrict__ bf, int wantE,
                      int wantB, int ntiles, const TilePlan* __restrict__ tp,
                      const uint32_t* __restrict__ total, const double4* __restrict__ srec,
                      const uint32_t* __restrict__ sv, const double* __restrict__ x,
                      const double* __restrict__ y, const double* __restrict__ z, IP p,
                      double* __restrict__ out8, uint32_t* __restrict__ nclamp) {
  using A = typename std::conditional<F32, float, double>::type;
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int64_t sy = p.nn[0], sz = int64_t(p.nn[0]) * p.nn[1], nt = sz * p.nn[2];
  const int64_t s0 = tp->bstart[ntiles], s1 = s0 + total[ntiles];
  uint32_t mine = 0;
  for (int64_t s = s0 + blockIdx.x * int64_t(blockDim.x) + threadIdx.x; s < s1;
       s += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t dst = sv[s];
    double xi, yi, zi;
    if (F32) {
      xi = x[dst];
      yi = y[dst];
      zi = z[dst];
    } else {
      const double4 r = srec[s];
      xi = r.x;
      yi = r.y;
      zi = r.z;
    }
    bool fl = false;
    const int ip = clamp_idx(xi, p.mn[0], p.hinv[0], p.nhi[0], fl);
    const int jp = clamp_idx(yi, p.mn[1], p.hinv[1], p.nhi[1], fl);
    const int kp = clamp_idx(zi, p.mn[2], p.hinv[2], p.nhi[2], fl);
    const A ab = A(((p.mn[0] - xi) + ip * p.delta[0]) * p.hinv[0]);
    const A de = A(((p.mn[1] - yi) + jp * p.delta[1]) * p.hinv[1]);
    const A gh = A(((p.mn[2] - zi) + kp * p.delta[2]) * p.hinv[2]);
    int i0 = ip - p.nlo[0], j0 = jp - p.nlo[1], k0 = kp - p.nlo[2];
    i0 = i0 < 0 ? 0 : (i0 > p.nn[0] - 2 ? p.nn[0] - 2 : i0);
    j0 = j0 < 0 ? 0 : (j0 > p.nn[1] - 2 ? p.nn[1] - 2 : j0);
    k0 = k0 < 0 ? 0 : (k0 > p.nn[2] - 2 ? p.nn[2] - 2 : k0);
    const int64_t b = i0 + sy * j0 + sz * k0;
    A v[6] = {A(0), A(0), A(0), A(0), A(0), A(0)};
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (c < 3 ? wantE : wantB) {
        const double* F = (c < 3 ? ef : bf) + (c % 3) * nt;
        v[c] = wsum8_any<A>(
            [&](int rr, int d) { return A(F[b + d + ((rr & 1) ? sy : 0) + ((rr & 2) ? sz : 0)]); },
            ab, de, gh);
      }
    mine += fl ? 1u : 0u;
    double2* o = reinterpret_cast<double2*>(out8 + 8 * int64_t(dst));
    o[0] = make_double2(double(v[0]), double(v[1]));
    o[1] = make_double2(double(v[2]), double(v[3]));
    o[2] = make_double2(double(v[4]), double(v[5]));
    o[3] = make_double2(fl ? 1.0 : 0.0, 0.0);
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nclamp, cnt);
}

}  // namespace

void pack_nodes(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf, double* rec) {
  const int64_t nt = int64_t(a.nn[0]) * a.nn[1] * a.nn[2];
  const int64_t nb = std::min<int64_t>((nt + kThreads - 1) / kThreads, 65536 * 8);
  ProfScope pp(c, ST_GATHER_PACK);
  hipLaunchKernelGGL(k_pack_nodes, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream, ef, bf, nt,
                     rec);
  OSC_HIP(hipGetLastError());
}

void interpolate_sorted(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf,
                        const SortedParticles& sp, double* E, double* B, int8_t* clamped,
                        const double* packed, const double* x, const double* y,
                        const double* z) {
  IP p{};
  for (int d = 0; d < 3; ++d) {
    p.mn[d] = a.mn[d];
    p.delta[d] = a.delta[d];
    p.hinv[d] = 1.0 / a.delta[d];
    p.nlo[d] = a.nlo[d];
    p.nhi[d] = a.nhi[d];
    p.nn[d] = a.nn[d];
    OSC_REQUIRE(a.nn[d] >= 2, OSC_ERR_ARG, "interpolation needs at least 2 nodes per axis");
  }
  const int64_t nt = int64_t(a.nn[0]) * a.nn[1] * a.nn[2];
  uint32_t* cnt = static_cast<uint32_t*>(c->counters) + 2;
  fill_u32(c, cnt, 1, 0u);
  if (sp.tiled) {
    // bin order of the tiled deposit: each tile's nodes staged in LDS
    OSC_REQUIRE(!sp.f32 || (x && y && z), OSC_ERR_ARG, "FP32 gather: coordinates needed");
    const int64_t n = sp.n;
    if (n <= 0) return;
    DevBuf out8(c->pool, sizeof(double) * 8 * n);
    const TilePlan* tpl = sp.plan.as<const TilePlan>();
    const uint32_t* tot = sp.total.as<const uint32_t>();
    const int wantE = int(E && ef), wantB = int(B && bf);
    // gather items: at most one partial chunk per tile beyond the full ones
    const int64_t gitems = (n + kTileGatherChunk - 1) / kTileGatherChunk + sp.tile.ntiles;
    {
      ProfScope ps(c, ST_GATHER);
      if (sp.f32) {
        hipLaunchKernelGGL(k_tile_gather<true>, dim3(unsigned(gitems)), dim3(TG_T), 0, c->stream,
                           ef, bf, wantE, wantB, sp.tile, tpl, tot,
                           sp.srec.as<const double4>(), sp.sk.as<const uint32_t>(),
                           sp.sv.as<const uint32_t>(), p, out8.as<double>(), cnt);
        OSC_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_tile_gather_off<true>, dim3(64), dim3(kThreads), 0, c->stream, ef, bf,
                           wantE, wantB, sp.tile.ntiles, tpl, tot, sp.srec.as<const double4>(),
                           sp.sv.as<const uint32_t>(), x, y, z, p, out8.as<double>(), cnt);
      } else {
        hipLaunchKernelGGL(k_tile_gather<false>, dim3(unsigned(gitems)), dim3(TG_T), 0,
                           c->stream, ef, bf, wantE, wantB, sp.tile, tpl, tot,
                           sp.srec.as<const double4>(), static_cast<const uint32_t*>(nullptr),
                           sp.sv.as<const uint32_t>(), p, out8.as<double>(), cnt);
        OSC_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_tile_gather_off<false>, dim3(64), dim3(kThreads), 0, c->stream, ef,
                           bf, wantE, wantB, sp.tile.ntiles, tpl, tot, sp.srec.as<const double4>(),
                           sp.sv.as<const uint32_t>(), static_cast<const double*>(nullptr),
                           static_cast<const double*>(nullptr), static_cast<const double*>(nullptr),
                           p, out8.as<double>(), cnt);
      }
      OSC_HIP(hipGetLastError());
    }
    ProfScope po(c, ST_GATHER_OUT);
    const int64_t nb = std::min<int64_t>((n + kThreads - 1) / kThreads, 65536 * 8);
    hipLaunchKernelGGL(k_unpack, dim3(unsigned(nb)), dim3(kThreads), kUnpackOccLds, c->stream,
                       out8.as<const double>(), n, (E && ef) ? E : nullptr,
                       (B && bf) ? B : nullptr, clamped);
    OSC_HIP(hipGetLastError());
    return;
  }
  if (!packed && sp.n < 2 * nt) {
    // fewer particles than twice the nodes: node values from the planar
    // fields, no node records packed (at 256^3 the pack alone costs about
    // what the planar reads of 10^7 particles do)
    OSC_REQUIRE(!sp.f32 || (x && y && z), OSC_ERR_ARG, "FP32 gather: coordinates needed");
    const int64_t n = sp.n;
    if (n <= 0) return;
    DevBuf out8(c->pool, sizeof(double) * 8 * n);
    const int64_t nb = std::min<int64_t>((n + kThreads - 1) / kThreads, 65536 * 8);
    const int wantE = int(E && ef), wantB = int(B && bf);
    {
      ProfScope ps(c, ST_GATHER);
      if (sp.f32)
        hipLaunchKernelGGL(k_interp_planar<true>, dim3(unsigned(nb)), dim3(kThreads), 0,
                           c->stream, ef, bf, wantE, wantB, n, sp.srec.as<const double4>(),
                           sp.sk.as<const uint32_t>(), sp.sv.as<const uint32_t>(), x, y, z, p,
                           out8.as<double>(), cnt);
      else
        hipLaunchKernelGGL(k_interp_planar<false>, dim3(unsigned(nb)), dim3(kThreads), 0,
                           c->stream, ef, bf, wantE, wantB, n, sp.srec.as<const double4>(),
                           static_cast<const uint32_t*>(nullptr), sp.sv.as<const uint32_t>(),
                           static_cast<const double*>(nullptr), static_cast<const double*>(nullptr),
                           static_cast<const double*>(nullptr), p, out8.as<double>(), cnt);
      OSC_HIP(hipGetLastError());
    }
    ProfScope po(c, ST_GATHER_OUT);
    hipLaunchKernelGGL(k_unpack, dim3(unsigned(nb)), dim3(kThreads), kUnpackOccLds, c->stream,
                       out8.as<const double>(), n, wantE ? E : nullptr, wantB ? B : nullptr,
                       clamped);
    OSC_HIP(hipGetLastError());
    return;
  }
  if (sp.f32) {
    // FP32: 32-byte node and result records, the deposit's cells and weights
    OSC_REQUIRE(x && y && z && !packed, OSC_ERR_ARG, "FP32 gather: coordinates needed");
    DevBuf rec32(c->pool, sizeof(float) * 8 * nt);
    {
      ProfScope pp(c, ST_GATHER_PACK);
      const int64_t nbp = std::min<int64_t>((nt + kThreads - 1) / kThreads, 65536 * 8);
      hipLaunchKernelGGL(k_pack_nodes32, dim3(unsigned(nbp)), dim3(kThreads), 0, c->stream, ef,
                         bf, nt, rec32.as<float>());
      OSC_HIP(hipGetLastError());
    }
    const int64_t n = sp.n;
    if (n <= 0) return;
    DevBuf out8(c->pool, sizeof(double) * 8 * n);
    const int64_t nb = std::min<int64_t>((n + kThreads - 1) / kThreads, 65536 * 8);
    {
      ProfScope ps(c, ST_GATHER);
      hipLaunchKernelGGL(k_interp_cells32, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream,
                         rec32.as<const float>(), int(E && ef), int(B && bf), n,
                         sp.srec.as<const float4>(), sp.sk.as<const uint32_t>(),
                         sp.sv.as<const uint32_t>(), x, y, z, p, out8.as<double>(), cnt);
      OSC_HIP(hipGetLastError());
    }
    ProfScope po(c, ST_GATHER_OUT);
    hipLaunchKernelGGL(k_unpack, dim3(unsigned(nb)), dim3(kThreads), kUnpackOccLds, c->stream,
                       out8.as<const double>(), n, (E && ef) ? E : nullptr,
                       (B && bf) ? B : nullptr, clamped);
    OSC_HIP(hipGetLastError());
    return;
  }
  DevBuf rec(c->pool, packed ? 0 : sizeof(double) * 8 * nt);
  if (!packed) {
    pack_nodes(c, a, ef, bf, rec.as<double>());
    packed = rec.as<const double>();
  }
  const int64_t n = sp.n;
  if (n > 0) {
    DevBuf out8(c->pool, sizeof(double) * 8 * n);
    const int64_t nb = std::min<int64_t>((n + kThreads - 1) / kThreads, 65536 * 8);
    {
      ProfScope ps(c, ST_GATHER);
      hipLaunchKernelGGL(k_interp_cells, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream,
                         packed, int(E && ef), int(B && bf), n,
                         sp.srec.as<const double4>(), sp.sv.as<const uint32_t>(), p,
                         out8.as<double>(), cnt);
      OSC_HIP(hipGetLastError());
    }
    ProfScope po(c, ST_GATHER_OUT);
    // kUnpackOccLds bytes of dynamic LDS, unused by the kernel, cap it at 4
    // workgroups (4 waves per SIMD) per CU: measured 223 us against 233 us at
    // the 8 waves per SIMD its registers allow (10^7 records; 3 waves: 239 us)
    hipLaunchKernelGGL(k_unpack, dim3(unsigned(nb)), dim3(kThreads), kUnpackOccLds, c->stream,
                       out8.as<const double>(), n, (E && ef) ? E : nullptr, (B && bf) ? B : nullptr,
                       clamped);
    OSC_HIP(hipGetLastError());
  }
}

int64_t interpolate(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf,
                    int64_t n, const double* x, const double* y, const double* z, double* E,
                    double* B, int8_t* clamped) {
  IP p{};
  for (int d = 0; d < 3; ++d) {
    p.mn[d] = a.mn[d];
    p.delta[d] = a.delta[d];
    p.hinv[d] = 1.0 / a.delta[d];  // hxi = 1.d0/delta (:319)
    p.nlo[d] = a.nlo[d];
    p.nhi[d] = a.nhi[d];
    p.nn[d] = a.nn[d];
    OSC_REQUIRE(a.nn[d] >= 2, OSC_ERR_ARG, "interpolation needs at least 2 nodes per axis");
  }
  uint32_t* cnt = static_cast<uint32_t*>(c->counters) + 2;
  const int64_t nt = int64_t(a.nn[0]) * a.nn[1] * a.nn[2];
  if (n >= kGatherSortMin && n >= nt / 8) {
    // many points in arbitrary order: take them in cell order, so that the
    // node records are read with locality (random order reads every node
    // record line many times over).  The points of the last large deposit
    // come with their order (checked by content hash); others are sorted.
    if (c->kept_ok && c->kept.n == n) {
      {
        ProfScope ps(c, ST_GATHER_SORT);
        queue_deposit_hash_check(c, n, x, y, z);
      }
      // the node records do not depend on the order: packed while the host
      // waits for the hashes
      DevBuf rec(c->pool, sizeof(double) * 8 * nt);
      if (!c->kept.tiled) pack_nodes(c, a, ef, bf, rec.as<double>());
      if (deposit_hash_matches(c)) {
        interpolate_sorted(c, a, ef, bf, c->kept, E, B, clamped, rec.as<const double>());
        return 0;
      }
      SortedParticles sp;
      {
        ProfScope ps(c, ST_GATHER_SORT);
        sort_for_gather(c, a, n, x, y, z, &sp);
      }
      interpolate_sorted(c, a, ef, bf, sp, E, B, clamped,
                         c->kept.tiled ? nullptr : rec.as<const double>());
      return 0;
    }
    SortedParticles sp;
    {
      ProfScope ps(c, ST_GATHER_SORT);
      sort_for_gather(c, a, n, x, y, z, &sp);
    }
    interpolate_sorted(c, a, ef, bf, sp, E, B, clamped);
    return 0;
  }
  ProfScope ps(c, ST_GATHER);
  fill_u32(c, cnt, 1, 0u);
  if (n <= 0) return 0;
  if (n < nt / 8) {
    // few points: read the planar fields directly instead of packing the mesh
    int64_t nb = std::min<int64_t>((n + kThreads - 1) / kThreads, 65536 * 8);
    hipLaunchKernelGGL(k_interp, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream, ef, bf, n, x,
                       y, z, p, E, B, clamped, cnt);
    OSC_HIP(hipGetLastError());
    return 0;
  }
  DevBuf rec(c->pool, sizeof(double) * 8 * nt);
  int64_t nb = std::min<int64_t>((nt + kThreads - 1) / kThreads, 65536 * 8);
  ps.stop();
  {
    ProfScope pp(c, ST_GATHER_PACK);
    hipLaunchKernelGGL(k_pack_nodes, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream, ef, bf, nt,
                       rec.as<double>());
    OSC_HIP(hipGetLastError());
  }
  ProfScope pg(c, ST_GATHER);
  nb = std::min<int64_t>((n + kThreads - 1) / kThreads, 65536 * 8);
  hipLaunchKernelGGL(k_interp_rec<false>, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream,
                     rec.as<const double>(), int(E && ef), int(B && bf), n, x, y, z,
                     static_cast<const double4*>(nullptr), static_cast<const uint32_t*>(nullptr), p,
                     E, B, clamped, cnt);
  OSC_HIP(hipGetLastError());
  return 0;
}

void interpolate_packed(osc_ctx* c, const InterpArgs& a, const double* ef, const double* bf,
                        int64_t n, const double* x, const double* y, const double* z, double* E,
                        double* B, int8_t* clamped, double* rec, uint32_t* cnt) {
  IP p{};
  for (int d = 0; d < 3; ++d) {
    p.mn[d] = a.mn[d];
    p.delta[d] = a.delta[d];
    p.hinv[d] = 1.0 / a.delta[d];  // hxi = 1.d0/delta (:319)
    p.nlo[d] = a.nlo[d];
    p.nhi[d] = a.nhi[d];
    p.nn[d] = a.nn[d];
    OSC_REQUIRE(a.nn[d] >= 2, OSC_ERR_ARG, "interpolation needs at least 2 nodes per axis");
  }
  fill_u32(c, cnt, 1, 0u);
  if (n <= 0) return;
  // no profiling scopes here: their events would be created on demand
  const int64_t nt = int64_t(a.nn[0]) * a.nn[1] * a.nn[2];
  const int64_t nbp = std::min<int64_t>((nt + kThreads - 1) / kThreads, 65536 * 8);
  hipLaunchKernelGGL(k_pack_nodes, dim3(unsigned(nbp)), dim3(kThreads), 0, c->stream, ef, bf, nt,
                     rec);
  OSC_HIP(hipGetLastError());
  const int64_t nb = std::min<int64_t>((n + kThreads - 1) / kThreads, 65536 * 8);
  hipLaunchKernelGGL(k_interp_rec<false>, dim3(unsigned(nb)), dim3(kThreads), 0, c->stream, rec,
                     int(E && ef), int(B && bf), n, x, y, z, static_cast<const double4*>(nullptr),
                     static_cast<const uint32_t*>(nullptr), p, E, B, clamped, cnt);
  OSC_HIP(hipGetLastError());
}

}  // namespace osc
