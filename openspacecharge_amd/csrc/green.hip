// Green-function builders (FP64, on the device), their R2C spectra, the
// spectrum cache, and conversions to the upstream full-spectrum convention.
#include "igf_math.h"
#include "osc_internal.h"

namespace osc {

namespace {

constexpr int kThreads = 256;

inline int64_t nblk(int64_t n, int t = kThreads) {
  int64_t b = (n + t - 1) / t;
  return b > 65536 * 8 ? 65536 * 8 : b;
}

// primitive of osc_get_cgrn_freespace on the offset grid
// (open_spacecharge_mod.f90:674-691): F = prim(u,v,w) * factor.
// Grid (ceil(L0/64), L1, L2) of 64-thread blocks: no index division.
// Mirrored axes (the axis has no offset and even L): its coordinate is taken
// as (i + 1/2 - L/2) * d -- a half-integer times d, rounded once, so that the
// mirror point L-1-i is exactly its negative -- and only the lower half along
// it is evaluated.  xlafun2 is even in its first argument (x atan(yz/(rx)) is
// even in x, r depends on x^2): along `mev` the value is stored at i and
// L-1-i.  It is odd in its third argument (every term changes sign with z):
// along `mod` the negated value is stored at L-1-i; there the mirror half
// differs from a direct evaluation only by the rounding of the quotient
// inside the last logarithm.
//
// With md set, the spacing comes from the mesh record on the device (a
// captured solve replayed for meshes the host has not seen yet) and dx, dy,
// dz, factor and the minima are formed here with the host's operations in
// the host's order (build_offset_primitive); img selects the z offset of the
// cathode image solve, 2 min(3) + (max(3) - min(3)) (open_spacecharge_mod.f90:472-476).
__global__ void __launch_bounds__(64)
    k_prim_offset(double* __restrict__ F, int L0, int L1, int L2, double dx, double dy, double dz,
                  double umin, double vmin, double wmin, double factor, int icomp, int mev,
                  int mod, int half, int zlo, int zhi, int zbase,
                  const MeshDev* __restrict__ md, double gamma, int img) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, k = zbase + int(blockIdx.z);
  if (md) {
    dx = md->d[0];
    dy = md->d[1];
    dz = md->d[2] * gamma;
    factor = (icomp == 1 || icomp == 2) ? gamma / (dx * dy * dz) : 1.0 / (dx * dy * dz);
    umin = (0.5 - double(L0 / 2)) * dx;
    vmin = (0.5 - double(L1 / 2)) * dy;
    wmin = (0.5 - double(L2 / 2)) * dz;
    umin = umin + 0.0;
    vmin = vmin + 0.0;
    double oz = 0.0;
    if (img) oz = oz + 2 * md->mn[2] + (md->mx[2] - md->mn[2]);
    wmin = wmin + oz * gamma;
  }
  const bool m0 = mev == 0 || mod == 0, m1 = mev == 1 || mod == 1, m2 = mev == 2 || mod == 2;
  if (i >= (m0 ? L0 / 2 : L0)) return;
  // only the planes [zlo, zhi] are read (the near cells of a far-field build):
  // a plane outside is evaluated when its mirror image is inside
  if (k < zlo || k > zhi) {
    if (!m2 || L2 - 1 - k < zlo || L2 - 1 - k > zhi) return;
  }
  const double w = m2 ? (double(k) + (0.5 - double(L2 / 2))) * dz : double(k) * dz + wmin;
  const double v = m1 ? (double(j) + (0.5 - double(L1 / 2))) * dy : double(j) * dy + vmin;
  const double u = m0 ? (double(i) + (0.5 - double(L0 / 2))) * dx : double(i) * dx + umin;
  double g;
  if (icomp == 0) g = lafun2(u, v, w) * factor;
  else if (icomp == 1) g = xlafun2(u, v, w) * factor;
  else if (icomp == 2) g = xlafun2(v, w, u) * factor;
  else g = xlafun2(w, u, v) * factor;
  const int L[3] = {L0, L1, L2};
  const int64_t sy = L0, sz = int64_t(L0) * L1;
  const int64_t at = i + sy * j + sz * k;
  F[at] = g;
  // mirror images along an axis of `half` (bit a) are stored only for the
  // first plane past the lower half, the one plane of them the corner
  // difference of the lower half reads
  int64_t ev = -1;  // the mirror index along mev
  bool wev = false;
  if (mev >= 0) {
    const int c = mev == 0 ? i : (mev == 1 ? j : k);
    ev = at + (L[mev] - 1 - 2 * c) * (mev == 0 ? 1 : (mev == 1 ? sy : sz));
    wev = !((half >> mev) & 1) || c == L[mev] / 2 - 1;
    if (wev) F[ev] = g;
  }
  if (mod >= 0) {
    const int c = mod == 0 ? i : (mod == 1 ? j : k);
    const int64_t step = (L[mod] - 1 - 2 * c) * (mod == 0 ? 1 : (mod == 1 ? sy : sz));
    if (!((half >> mod) & 1) || c == L[mod] / 2 - 1) {
      F[at + step] = -g;
      if (wev) F[ev + step] = -g;
    }
  }
}

// 8-corner difference over 1:size-1 with the last planes left raw
// (open_spacecharge_mod.f90:697-705); output z index circularly shifted by sh
__global__ void __launch_bounds__(64)
    k_diff_offset(const double* __restrict__ F, double* __restrict__ G, int L0, int L1, int L2,
                  int sh, int zero_raw) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, kq = blockIdx.z;
  if (i >= L0) return;
  const int64_t P = int64_t(L0) * L1;
  int k = (kq - sh) % L2;
  if (k < 0) k += L2;
  const int64_t b = i + int64_t(L0) * j + P * k;
  double g;
  if (i < L0 - 1 && j < L1 - 1 && k < L2 - 1) {
    g = F[b + 1 + L0 + P] - F[b + L0 + P] - F[b + 1 + P] - F[b + 1 + L0] - F[b] + F[b + P] +
        F[b + L0] + F[b + 1];
  } else {
    // upstream keeps the raw primitive here; it never reaches the
    // extracted octant, and zeroing it removes its round-off from the FFT
    g = zero_raw ? 0.0 : F[b];
  }
  G[i + int64_t(L0) * j + P * kq] = g;
}

// legacy loops: osc_getgrnfree (core:476-493), osc_getgrnimageshift
// (core:1157-1174), osc_getgrnimageconvcorr (core:1207-1226)
__global__ void k_green_legacy(double* __restrict__ G, GreenSpec s) {
  const int L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
  const int64_t N = int64_t(L0) * L1 * L2;
  const double dx = s.delta[0], dy = s.delta[1], dz = s.delta[2];
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < N;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int ii = int(idx % L0);
    const int jj = int((idx / L0) % L1);
    const int kq = int(idx / (int64_t(L0) * L1));
    int kk = (kq - s.out_shift_z) % L2;
    if (kk < 0) kk += L2;
    double w;
    if (s.kind == GK_IMG_CONVCORR) {
      const int kp = s.lo[2] + kk;
      w = kp * dz + 2.0 * s.umin[2] - 2.0 * s.nlo3 * s.delta[2];
    } else {
      const int kp = s.lo[2] + (kk + s.shift[2]) % L2;
      w = kp * dz;
      if (s.kind == GK_IMG_SHIFT) w = w + s.zshft;
    }
    const int jp = s.lo[1] + (jj + s.shift[1]) % L1;
    const double v = jp * dy;
    const int ip = s.lo[0] + (ii + s.shift[0]) % L0;
    const double u = ip * dx;
    G[idx] = igf_legacy(s.igf, s.icomp, u, v, w, s.gamma, dx, dy, dz);
  }
}

// osc_getgrnpipe loops (core:652-713), term t = 0..3
// rfun (core:723-755) factorised over the grid: the cosines depend only on
// the x (resp. y) index and the mode, the z profile only on the z index and
// the mode pair.  k_pipe_tables evaluates them once with the expressions of
// rfun; k_pipe_rows sums, per (y, z) row and x mode m, the five y modes
// (j^n cos(n v pi/b) zfun_mn / kap_mn, each term as rfun forms it);
// k_green_pipe then adds the five x modes i^m cos(m u pi/a) times the row
// sums.  The 25 terms are grouped by m where rfun adds them in one sequence,
// so the values agree with rfun to a few ulps rather than bitwise: 5 terms
// per point instead of 25 quotients.
constexpr double kPi = 3.141592653589793115997963468544185161590576171875;  // acos(-1d0)

__device__ __forceinline__ double pipe_u(const GreenSpec& s, int ii, int si) {
  if (si == 1) return double(s.lo[0] + (ii + s.shift[0]) % s.L[0]) * s.delta[0];
  return 2.0 * s.umin[0] + ii * s.delta[0];
}
__device__ __forceinline__ double pipe_v(const GreenSpec& s, int jj, int sj) {
  if (sj == 1) return double(s.lo[1] + (jj + s.shift[1]) % s.L[1]) * s.delta[1];
  return 2.0 * s.umin[1] + jj * s.delta[1];
}

// cu[ii*5 + m-1] = cos(m u piainv), cv[jj*5 + n-1] = cos(n v pibinv),
// zf[kk*25 + (m-1)*5 + n-1] = zfun(w), kap[(m-1)*5 + n-1] = kapmn
__global__ void k_pipe_tables(GreenSpec s, int si, int sj, double* __restrict__ cu,
                              double* __restrict__ cv, double* __restrict__ zf,
                              double* __restrict__ kap) {
  const double ainv = 1.0 / s.a, binv = 1.0 / s.b;
  const double piainv = kPi * ainv, pibinv = kPi * binv;
  const double hz = s.delta[2], gam = s.gamma;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < s.L[0]) {
    const double u = pipe_u(s, t, si);
    for (int m = 1; m <= 5; ++m) cu[t * 5 + m - 1] = cos(m * u * piainv);
  }
  if (t < s.L[1]) {
    const double v = pipe_v(s, t, sj);
    for (int n = 1; n <= 5; ++n) cv[t * 5 + n - 1] = cos(n * v * pibinv);
  }
  if (t < s.L[2]) {
    const int kp = s.lo[2] + (t + s.shift[2]) % s.L[2];
    const double w = kp * hz;
    for (int m = 1; m <= 5; ++m)
      for (int n = 1; n <= 5; ++n) {
        const double km = m * kPi * ainv, kn = n * kPi * binv;
        const double kapmn = sqrt(km * km + kn * kn);
        double zfun = (exp(-(kapmn * fabs(gam * w - hz))) - 2.0 * exp(-(kapmn * fabs(gam * w))) +
                       exp(-(kapmn * fabs(gam * w + hz)))) /
                      (hz * hz * (kapmn * kapmn));
        if (w == 0) zfun = zfun + 2.0 / (hz * kapmn);
        zf[t * 25 + (m - 1) * 5 + n - 1] = zfun;
        if (t == 0) kap[(m - 1) * 5 + n - 1] = kapmn;
      }
  }
}

// the y-mode sums of every (y, z) row: A[(jj + L1 kk) 5 + m-1]
__global__ void __launch_bounds__(64)
    k_pipe_rows(GreenSpec s, int sj, const double* __restrict__ cv, const double* __restrict__ zf,
                const double* __restrict__ kap, double* __restrict__ A) {
  const int jj = blockIdx.x * 64 + threadIdx.x, kk = blockIdx.y;
  if (jj >= s.L[1]) return;
  double* a = A + (int64_t(jj) + int64_t(s.L[1]) * kk) * 5;
  for (int m = 1; m <= 5; ++m) {
    double acc = 0.0;
    int jn = 1;
    for (int n = 1; n <= 5; ++n) {
      jn *= sj;  // j**n
      acc = acc + double(jn) * cv[jj * 5 + n - 1] * zf[kk * 25 + (m - 1) * 5 + n - 1] /
                      kap[(m - 1) * 5 + n - 1];
    }
    a[m - 1] = acc;
  }
}

// osc_getgrnpipe loops (core:652-713) for one of the four terms
// (s.term: bit 1 -> image in x, bit 0 -> image in y); 64 x 4 threads per
// workgroup (one wave per row), grid (ceil(L0/64), ceil(L1/4), L2)
__global__ void __launch_bounds__(256)
    k_green_pipe(double* __restrict__ G, GreenSpec s, const double* __restrict__ cu,
                 const double* __restrict__ A) {
  const int ii = blockIdx.x * 64 + (threadIdx.x & 63);
  const int jj = blockIdx.y * 4 + (threadIdx.x >> 6), kk = blockIdx.z;
  if (ii >= s.L[0] || jj >= s.L[1]) return;
  const int si = (s.term & 2) ? -1 : 1;
  const double ainv = 1.0 / s.a, binv = 1.0 / s.b;
  const double* a = A + (int64_t(jj) + int64_t(s.L[1]) * kk) * 5;  // uniform
  double res = 0.0;
  int im = 1;
  for (int m = 1; m <= 5; ++m) {
    im *= si;  // i**m
    res = res + double(im) * cu[ii * 5 + m - 1] * a[m - 1];
  }
  res = res * 2.0 * kPi * ainv * binv;
  G[ii + int64_t(s.L[0]) * (jj + int64_t(s.L[1]) * kk)] = res;
}

// The pipe Green spectrum from its separable form: G_t(i, j, k) =
// C sum_{m,n} A[i,m] B[j,n] Z[k,m,n] with A = i^m cos(m u pi/a) (cu),
// B = j^n cos(n v pi/b) (cv), Z = zfun_mn(w) / kap_mn (zf, kap) and
// C = 2 pi / (a b), so that its DFT is C sum_{m,n} A^[kx,m] B^[ky,n]
// Z^[kz,m,n] with 1-D DFTs of the profiles.
// k_pipe_dft: out[k P + q] = sum_j w(j, q) exp(-2 pi i k j / L) for k < K,
// q < P, with the length's twiddle table; w = s^(q+1) prof (s = sign, the
// mode's power of i or j), or prof / kap[q] when kap is given.  One wave per
// output: lane l sums j = l, l + 64, ..., then the wave adds the lanes.
__global__ void __launch_bounds__(64)
    k_pipe_dft(const double* __restrict__ prof, int P, int L, int sign,
               const double* __restrict__ kap, const cpx<double>* __restrict__ tw,
               cpx<double>* __restrict__ out) {
  const int k = blockIdx.x, q = blockIdx.y, lane = threadIdx.x;
  const double sq = (sign < 0 && (q % 2) == 0) ? -1.0 : 1.0;  // (-1)^(q+1)
  double re = 0.0, im = 0.0;
  for (int j = lane; j < L; j += 64) {
    const double v = kap ? prof[j * P + q] / kap[q] : sq * prof[j * P + q];
    const cpx<double> w = tw[(k * j) % L];
    re = re + v * w.x;
    im = im + v * w.y;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    re = re + __shfl_xor(re, o, 64);
    im = im + __shfl_xor(im, o, 64);
  }
  if (lane == 0) out[k * P + q] = mk(re, im);
}

// k_pipe_spec: one (ky, kz) row of the half spectrum (KXP entries, zero past
// KX): T[m] = C sum_n B^[ky,n] Z^[kz,m,n], then sum_m A^[kx,m] T[m]
__global__ void __launch_bounds__(256)
    k_pipe_spec(const cpx<double>* __restrict__ ax, const cpx<double>* __restrict__ by,
                const cpx<double>* __restrict__ cz, int KX, int KXP, int L1, double C,
                cpx<double>* __restrict__ spec) {
  // one wave per row, four rows per workgroup
  const int ky = blockIdx.x * 4 + (threadIdx.x >> 6), kz = blockIdx.y;
  if (ky >= L1) return;
  cpx<double> T[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    cpx<double> acc = mk(0.0, 0.0);
#pragma unroll
    for (int n = 0; n < 5; ++n) acc = add(acc, mul(by[ky * 5 + n], cz[kz * 25 + m * 5 + n]));
    T[m] = mk(acc.x * C, acc.y * C);
  }
  cpx<double>* row = spec + (int64_t(kz) * L1 + ky) * KXP;
  for (int kx = threadIdx.x & 63; kx < KXP; kx += 64) {
    cpx<double> v = mk(0.0, 0.0);
    if (kx < KX)
#pragma unroll
      for (int m = 0; m < 5; ++m) v = add(v, mul(ax[kx * 5 + m], T[m]));
    row[kx] = v;
  }
}

__global__ void k_convert(const cpx<double>* __restrict__ in, cpx<float>* __restrict__ out,
                          int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = mk(float(in[i].x), float(in[i].y));
}

}  // namespace

// Axes along which the solve-path primitive of an E component is mirrored
// (k_prim_offset): mev, where xlafun2 is even (its first argument), and mod,
// where it is odd (its third argument); -1 where the axis has an offset or odd
// length, for phi or a Green function returned to the caller, and unless
// `fast` (the FP32 convolution).  The mirrored coordinate (i + 1/2 - L/2) d
// is rounded once where upstream rounds (i d) and then its sum with
// (1/2 - L/2) d; xlafun2's terms log(r + y) and log((r - z)/(r + z)) cancel
// catastrophically where |y| or |z| dominates r, so that an ulp of coordinate
// moves the primitive far more than an ulp of its value.  Measured on an
// 8x8x256 mesh (512-point z lines): the mirrored build puts E 3e-11 of
// max|E| from a build in upstream's arithmetic, which a restatement with
// upstream's coordinates matches to 4e-14.  The FP64 path therefore
// evaluates every point with upstream's arithmetic and mirrors nothing; the
// FP32 convolution (rounding near 1e-7) keeps the quarter-grid evaluation.
static void prim_mirrors(const GreenSpec& s, bool fast, int& mev, int& mod) {
  auto mirrorable = [&](int ax) { return s.offset[ax] == 0.0 && s.L[ax] % 2 == 0; };
  mev = mod = -1;
  if (fast && s.zero_raw && s.icomp >= 1) {
    const int ea = s.icomp - 1, oa = (s.icomp + 1) % 3;
    if (mirrorable(ea)) mev = ea;
    if (mirrorable(oa)) mod = oa;
  }
}

// primitive F of an offset-kind Green function (open_spacecharge_mod.f90:662-691).
// lower_halves: the reader (the corner-difference x pass) loads only the
// lower half, plus one plane, along mirrored y and z axes (unshifted z), so
// the mirror images beyond that are not stored.
static void build_offset_primitive(osc_ctx* c, const GreenSpec& s, double* F, bool fast,
                                   bool lower_halves = false, int zlo = 0, int zhi = -1) {
  if (zhi < 0 && zlo == 0) zhi = s.L[2] - 1;  // every plane
  const double dx = s.delta[0], dy = s.delta[1], dz = s.delta[2] * s.gamma;
  const double factor =
      (s.icomp == 1 || s.icomp == 2) ? s.gamma / (dx * dy * dz) : 1.0 / (dx * dy * dz);
  double umin = (0.5 - double(s.L[0] / 2)) * dx;
  double vmin = (0.5 - double(s.L[1] / 2)) * dy;
  double wmin = (0.5 - double(s.L[2] / 2)) * dz;
  umin = umin + s.offset[0];
  vmin = vmin + s.offset[1];
  wmin = wmin + s.offset[2] * s.gamma;
  const dim3 g3(unsigned((s.L[0] + 63) / 64), unsigned(s.L[1]), unsigned(s.L[2]));
  // the FP32 solve path evaluates E components on a quarter of the grid (even
  // mirror along their own axis, odd mirror along the axis of xlafun2's third
  // argument, see k_prim_offset); every other build keeps upstream's
  // coordinate arithmetic on the whole grid
  int mev, mod;
  prim_mirrors(s, fast, mev, mod);
  dim3 gp = g3;
  if (mev == 0 || mod == 0) gp.x = unsigned((s.L[0] / 2 + 63) / 64);
  if (mev == 1 || mod == 1) gp.y = unsigned(s.L[1] / 2);
  // the planes launched: those of [zlo, zhi], or of the lower half whose
  // mirror images are in it
  int zbase = zlo, ztop = zhi;
  if (mev == 2 || mod == 2) {
    zbase = std::max(0, std::min(zlo, s.L[2] - 1 - zhi));
    ztop = std::min(s.L[2] / 2 - 1, std::max(zhi, s.L[2] - 1 - zlo));
  }
  if (ztop < zbase) return;
  gp.z = unsigned(ztop - zbase + 1);
  // a solve being captured for replay takes the spacing from the device mesh
  // record; only the zero offset of the free solve and the image offset of
  // the cathode solve occur there
  const MeshDev* md = c->green_md;
  OSC_REQUIRE(!md || (s.offset[0] == 0.0 && s.offset[1] == 0.0), OSC_ERR_ARG,
              "device-side Green parameters: unsupported offset");
  int half = 0;
  if (lower_halves) {
    if (mev == 1 || mod == 1) half |= 2;
    if ((mev == 2 || mod == 2) && s.out_shift_z == 0) half |= 4;
  }
  launch(k_prim_offset, gp, dim3(64), 0, c->stream, F, s.L[0], s.L[1], s.L[2], dx, dy,
                     dz, umin, vmin, wmin, factor, s.icomp, mev, mod, half, zlo, zhi, zbase,
                     md, s.gamma, int(md && s.out_shift_z != 0));
  OSC_HIP(hipGetLastError());
}

// The half spectrum of one pipe Green term (s.kind GK_PIPE) built from the
// separable form (k_pipe_dft, k_pipe_spec): no real-space array and no 3-D
// transform.  Agrees with the transformed real array (build_green_real, then
// spectrum_forward_full) to rounding.
void pipe_spectrum(osc_ctx* c, const GreenSpec& s, const SpecDims& d, cpx<double>* spec) {
  const int si = (s.term & 2) ? -1 : 1, sj = (s.term & 1) ? -1 : 1;
  const int L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
  const int lmax = std::max(L0, std::max(L1, L2));
  DevBuf tab(c->pool, sizeof(double) * (5 * (L0 + L1) + 25 * L2 + 25) +
                          sizeof(cpx<double>) * (5 * (d.KX + L1) + 25 * L2));
  // the complex profiles first (16-byte aligned), then the real tables
  cpx<double>* ax = tab.as<cpx<double>>();
  cpx<double>* by = ax + 5 * d.KX;
  cpx<double>* cz = by + 5 * L1;
  double* cu = reinterpret_cast<double*>(cz + 25 * L2);
  double* cv = cu + 5 * L0;
  double* zf = cv + 5 * L1;
  double* kap = zf + 25 * L2;
  ProfScope ps(c, ST_GREEN_EVAL);
  launch(k_pipe_tables, dim3((lmax + 63) / 64), dim3(64), 0, c->stream, s, si, sj, cu, cv, zf,
         kap);
  OSC_HIP(hipGetLastError());
  launch(k_pipe_dft, dim3(unsigned(d.KX), 5), dim3(64), 0, c->stream,
         static_cast<const double*>(cu), 5, L0, si, static_cast<const double*>(nullptr),
         twiddles<double>(c, L0), ax);
  launch(k_pipe_dft, dim3(unsigned(L1), 5), dim3(64), 0, c->stream,
         static_cast<const double*>(cv), 5, L1, sj, static_cast<const double*>(nullptr),
         twiddles<double>(c, L1), by);
  launch(k_pipe_dft, dim3(unsigned(L2), 25), dim3(64), 0, c->stream,
         static_cast<const double*>(zf), 25, L2, 1, static_cast<const double*>(kap),
         twiddles<double>(c, L2), cz);
  OSC_HIP(hipGetLastError());
  const double C = 2.0 * kPi * (1.0 / s.a) * (1.0 / s.b);
  launch(k_pipe_spec, dim3(unsigned((L1 + 3) / 4), unsigned(L2)), dim3(256), 0, c->stream,
         static_cast<const cpx<double>*>(ax), static_cast<const cpx<double>*>(by),
         static_cast<const cpx<double>*>(cz), d.KX, d.KXP, L1, C, spec);
  OSC_HIP(hipGetLastError());
}

void build_green_real(osc_ctx* c, const GreenSpec& s, double* g, bool fast) {
  const int64_t N = int64_t(s.L[0]) * s.L[1] * s.L[2];
  ProfScope ps(c, ST_GREEN_EVAL);
  if (s.kind == GK_OFFSET) {
    // osc_get_cgrn_freespace (open_spacecharge_mod.f90:633-707)
    DevBuf F(c->pool, sizeof(double) * N);
    build_offset_primitive(c, s, F.as<double>(), fast);
    const dim3 g3(unsigned((s.L[0] + 63) / 64), unsigned(s.L[1]), unsigned(s.L[2]));
    launch(k_diff_offset, g3, dim3(64), 0, c->stream, F.as<const double>(), g, s.L[0],
                       s.L[1], s.L[2], s.out_shift_z, s.zero_raw);
    OSC_HIP(hipGetLastError());
  } else if (s.kind == GK_PIPE) {
    const int si = (s.term & 2) ? -1 : 1, sj = (s.term & 1) ? -1 : 1;
    const int lmax = std::max(s.L[0], std::max(s.L[1], s.L[2]));
    DevBuf tab(c->pool, sizeof(double) * (5 * (s.L[0] + s.L[1]) + 25 * s.L[2] + 25 +
                                          5 * int64_t(s.L[1]) * s.L[2]));
    double* cu = tab.as<double>();
    double* cv = cu + 5 * s.L[0];
    double* zf = cv + 5 * s.L[1];
    double* kap = zf + 25 * s.L[2];
    double* rows = kap + 25;
    launch(k_pipe_tables, dim3((lmax + 63) / 64), dim3(64), 0, c->stream, s, si, sj, cu,
                       cv, zf, kap);
    OSC_HIP(hipGetLastError());
    launch(k_pipe_rows, dim3(unsigned((s.L[1] + 63) / 64), unsigned(s.L[2])), dim3(64), 0,
           c->stream, s, sj, cv, zf, kap, rows);
    OSC_HIP(hipGetLastError());
    const dim3 g3(unsigned((s.L[0] + 63) / 64), unsigned((s.L[1] + 3) / 4), unsigned(s.L[2]));
    launch(k_green_pipe, g3, dim3(256), 0, c->stream, g, s, cu, rows);
    OSC_HIP(hipGetLastError());
  } else {
    launch(k_green_legacy, dim3(nblk(N)), dim3(kThreads), 0, c->stream, g, s);
    OSC_HIP(hipGetLastError());
  }
}

// With a corner difference and sym (per axis: 0, +1 even, -1 odd about
// L/2 - 1, last plane zero), lines along y and z are computed only for the
// lower half of a symmetric axis until that axis is transformed, whose pass
// rebuilds the upper half as it loads (AxisIO::mh).
template <typename T>
void spectrum_forward_full(osc_ctx* c, const double* g, cpx<T>* spec, const SpecDims& d,
                           const CornerDiff* cd, const int* sym, GreenRef* zr) {
  const int L0 = d.L[0], L1 = d.L[1], L2 = d.L[2];
  ProfScope ps(c, ST_GREEN_FFT);
  const bool hy = cd && sym && sym[1] != 0 && L1 % 2 == 0;
  const bool hz = cd && sym && sym[2] != 0 && L2 % 2 == 0 && cd->sh == 0;
  const int Yh = hy ? L1 / 2 : L1, Zh = hz ? L2 / 2 : L2;
  R2CIO r{};
  r.in_sy = L0;
  r.in_sz = int64_t(L0) * L1;
  r.out_sy = d.KXP;
  r.out_sz = int64_t(d.KXP) * L1;
  r.nrows = int64_t(Yh) * Zh;
  r.ny_rows = Yh;
  r.n_in = L0;
  if (cd)
    launch_xr2c_diff<T>(L0, g, spec, r, *cd, twiddles<T>(c, L0), c->stream);
  else
    launch_xr2c<T>(L0, g, spec, r, twiddles<T>(c, L0), c->stream);
  AxisIO ay{};
  ay.s_b = ay.d_b = 1;
  ay.s_p = ay.d_p = d.KXP;
  ay.s_o1 = ay.d_o1 = int64_t(d.KXP) * L1;
  ay.nb = d.KX;
  ay.n_in = L1;
  ay.w0 = 0;
  ay.n_out = L1;
  ay.mh = hy ? L1 / 2 : 0;
  ay.ms = hy ? sym[1] : 0;
  launch_axis<T>(L1, -1, spec, spec, ay, Zh, 1, twiddles<T>(c, L1), c->stream);
  if (zr) {
    zr->g = spec;
    zr->zr = 1;
    zr->zmh = hz ? L2 / 2 : 0;
    zr->zms = hz ? sym[2] : 0;
    return;
  }
  AxisIO az{};
  az.s_b = az.d_b = 1;
  az.s_p = az.d_p = int64_t(d.KXP) * L1;
  az.s_o1 = az.d_o1 = d.KXP;
  az.nb = d.KX;
  az.n_in = L2;
  az.w0 = 0;
  az.n_out = L2;
  az.mh = hz ? L2 / 2 : 0;
  az.ms = hz ? sym[2] : 0;
  launch_axis<T>(L2, -1, spec, spec, az, L1, 1, twiddles<T>(c, L2), c->stream);
}
template void spectrum_forward_full<double>(osc_ctx*, const double*, cpx<double>*, const SpecDims&,
                                            const CornerDiff*, const int*, GreenRef*);
template void spectrum_forward_full<float>(osc_ctx*, const double*, cpx<float>*, const SpecDims&,
                                           const CornerDiff*, const int*, GreenRef*);

void convert_spectrum(osc_ctx* c, const cpx<double>* in, cpx<float>* out, int64_t n) {
  ProfScope ps(c, ST_GREEN_FFT);
  launch(k_convert, dim3(nblk(n)), dim3(kThreads), 0, c->stream, in, out, n);
  OSC_HIP(hipGetLastError());
}

// Far-field cells of the FP32 build (CornerDiff::far): interior cells whose
// centres lie at |w| >= kFarEdges times the longest cell edge from the origin
// take the second-order expansion of the field kernel's cell integral in the
// x pass, and only the primitive planes the remaining near cells difference
// are evaluated.  Against 8-point Gauss quadrature of the kernel the
// expansion's worst error is about 1e-6 of the local field scale V / r^2 at
// 12 edges for cells 40 times longer in w than across (a relativistic
// bunch's), less for rounder cells; the FP64 primitive differences carry
// cancellation errors of 1e-5 .. 1e-4 relative there.  Not in a context
// that captures solves (osc_ctx_set_graphs: a captured solve knows the
// spacing only on the device, and its eager and replayed steps must agree).
// Returns the primitive planes needed, [zlo, zhi] (empty when zlo > zhi).
constexpr double kFarEdges = 12.0;
static bool far_field_applies(const osc_ctx* c, const GreenSpec& s, bool fast) {
  return fast && c->far_green && !c->graphs && s.zero_raw && s.icomp >= 1 && !c->green_md;
}
static void far_field_setup(osc_ctx* c, const GreenSpec& s, bool fast, CornerDiff& cd, int& zlo,
                            int& zhi) {
  zlo = 0;
  zhi = s.L[2] - 1;
  if (!far_field_applies(c, s, fast)) return;
  const double dx = s.delta[0], dy = s.delta[1], dw = s.delta[2] * s.gamma;
  cd.far = 1;
  cd.icomp = s.icomp;
  cd.factor = (s.icomp == 1 || s.icomp == 2) ? s.gamma / (dx * dy * dw) : 1.0 / (dx * dy * dw);
  cd.h[0] = dx;
  cd.h[1] = dy;
  cd.h[2] = dw;
  cd.c0[0] = (1.0 - double(s.L[0] / 2)) * dx + s.offset[0];
  cd.c0[1] = (1.0 - double(s.L[1] / 2)) * dy + s.offset[1];
  cd.c0[2] = (1.0 - double(s.L[2] / 2)) * dw + s.offset[2] * s.gamma;
  cd.wnear = kFarEdges * std::max(dx, std::max(dy, dw));
  int lo = s.L[2], hi = -1;
  for (int k = 0; k < s.L[2] - 1; ++k)
    if (!(std::fabs(cd.c0[2] + k * cd.h[2]) >= cd.wnear)) {
      lo = std::min(lo, k);
      hi = std::max(hi, k + 1);
    }
  zlo = lo;
  zhi = hi;
}

// Spectrum of the Green function of s in the context's precision: the real
// Green function is always built in FP64 (IGF primitives and 8-corner
// differences); an FP32 context transforms it in FP32.  When the z length has
// a compile-time plan the z transform is left to conv_solve, which takes it
// line by line inside the multiply (the array is z-real, GreenRef::zr).
// Cached per precision under the full specification.
template <typename T>
GreenRef green_spectrum(osc_ctx* c, const GreenSpec& s, DevBuf& keepalive) {
  const SpecDims d = SpecDims::of(s.L);
  const int64_t N = int64_t(s.L[0]) * s.L[1] * s.L[2];
  std::string key(reinterpret_cast<const char*>(&s), sizeof(s));
  // spectra built with far-field cells are kept apart from exact ones
  if (s.kind == GK_OFFSET && far_field_applies(c, s, sizeof(T) == 4)) key.push_back('F');
  GreenCache& gc = c->gcache;
  GreenCache::Entry* hit = nullptr;
  // a captured solve builds its spectra on every replay (the spacing is only
  // known on the device then): no cache
  const bool use_cache = gc.enable && !c->green_md;
  if (use_cache) {
    auto it = gc.map.find(key);
    if (it != gc.map.end()) {
      hit = it->second.get();
      hit->last = ++gc.tick;
      DevBuf& have = sizeof(T) == 4 ? hit->f32 : hit->f64;
      if (have.p) {
        ++gc.hits;
        return GreenRef{have.p, hit->zr, hit->zmh, hit->zms};
      }
    }
    ++gc.misses;
  }
  DevBuf g(c->pool, sizeof(double) * N);
  DevBuf sp(c->pool, sizeof(cpx<T>) * d.size());
  GreenRef ref{};
  GreenRef* zr = fft_has_plan(s.L[2]) ? &ref : nullptr;
  const bool fast = sizeof(T) == 4;  // mirrored quarter-grid primitive (prim_mirrors)
  if (s.kind == GK_OFFSET && fft_has_plan(s.L[0])) {
    // the x pass takes the 8-corner difference of the primitive as it loads
    // its rows, so the difference array is never stored
    CornerDiff cd{s.L[1], s.L[2], s.out_shift_z, s.zero_raw};
    int zlo, zhi;
    far_field_setup(c, s, fast, cd, zlo, zhi);
    if (zlo <= zhi) {
      ProfScope ps(c, ST_GREEN_EVAL);
      build_offset_primitive(c, s, g.as<double>(), fast, true, zlo, zhi);
    }
    // the corner difference of a primitive that is even along an axis is odd
    // about L/2 - 1 there, and even where the primitive is odd
    int mev, mod;
    prim_mirrors(s, fast, mev, mod);
    int sym[3] = {0, 0, 0};
    if (mev >= 0) sym[mev] = -1;
    if (mod >= 0) sym[mod] = 1;
    spectrum_forward_full<T>(c, g.as<const double>(), sp.as<cpx<T>>(), d, &cd, sym, zr);
  } else {
    build_green_real(c, s, g.as<double>(), fast);
    spectrum_forward_full<T>(c, g.as<const double>(), sp.as<cpx<T>>(), d, nullptr, nullptr, zr);
  }
  g.reset();
  ref.g = sp.p;
  if (!use_cache || (!hit && !gc.admit(key))) {
    keepalive = std::move(sp);
    return ref;
  }
  GreenCache::Entry* ep = hit;
  if (!ep) {
    auto e = std::make_unique<GreenCache::Entry>();
    ep = e.get();
    gc.map[key] = std::move(e);
  }
  ep->zr = ref.zr;
  ep->zmh = ref.zmh;
  ep->zms = ref.zms;
  const size_t bytes = sizeof(cpx<T>) * d.size();
  (sizeof(T) == 4 ? ep->f32 : ep->f64) = std::move(sp);
  ep->bytes += bytes;
  ep->last = ++gc.tick;
  gc.bytes += bytes;
  // keep the entry being returned even if the budget is exceeded
  if (gc.bytes > gc.max_bytes) {
    ep->last = ~uint64_t(0);
    gc.evict_to(gc.max_bytes);
    ep->last = gc.tick;
  }
  return ref;
}
template GreenRef green_spectrum<double>(osc_ctx*, const GreenSpec&, DevBuf&);
template GreenRef green_spectrum<float>(osc_ctx*, const GreenSpec&, DevBuf&);

// full upstream array F_s[g](k) = F_-[g](sigma_s k), sigma flipping axes with s=+1
void half_to_upstream(osc_ctx* c, const cpx<double>* spec_dev, const SpecDims& d, const int s[3],
                      double* full_host) {
  std::vector<cpx<double>> h(d.size());
  OSC_HIP(hipMemcpyAsync(h.data(), spec_dev, sizeof(cpx<double>) * d.size(),
                         hipMemcpyDeviceToHost, c->stream));
  OSC_HIP(hipStreamSynchronize(c->stream));
  const int L0 = d.L[0], L1 = d.L[1], L2 = d.L[2];
  cpx<double>* out = reinterpret_cast<cpx<double>*>(full_host);
  for (int k = 0; k < L2; ++k)
    for (int j = 0; j < L1; ++j)
      for (int i = 0; i < L0; ++i) {
        int a = s[0] > 0 ? (L0 - i) % L0 : i;
        int b = s[1] > 0 ? (L1 - j) % L1 : j;
        int e = s[2] > 0 ? (L2 - k) % L2 : k;
        cpx<double> v;
        if (a <= L0 / 2) {
          v = h[a + int64_t(d.KXP) * (b + int64_t(L1) * e)];
        } else {
          v = h[(L0 - a) + int64_t(d.KXP) * ((L1 - b) % L1 + int64_t(L1) * ((L2 - e) % L2))];
          v.y = -v.y;
        }
        out[i + int64_t(L0) * (j + int64_t(L1) * k)] = v;
      }
}

void upstream_to_half(osc_ctx* c, const double* full_host, const SpecDims& d, const int s[3],
                      cpx<double>* spec_dev) {
  const int L0 = d.L[0], L1 = d.L[1], L2 = d.L[2];
  std::vector<cpx<double>> h(d.size(), mk(0.0, 0.0));
  const cpx<double>* in = reinterpret_cast<const cpx<double>*>(full_host);
  for (int k = 0; k < L2; ++k)
    for (int j = 0; j < L1; ++j)
      for (int i = 0; i < d.KX; ++i) {
        int a = s[0] > 0 ? (L0 - i) % L0 : i;
        int b = s[1] > 0 ? (L1 - j) % L1 : j;
        int e = s[2] > 0 ? (L2 - k) % L2 : k;
        h[i + int64_t(d.KXP) * (j + int64_t(L1) * k)] = in[a + int64_t(L0) * (b + int64_t(L1) * e)];
      }
  OSC_HIP(hipMemcpyAsync(spec_dev, h.data(), sizeof(cpx<double>) * d.size(),
                         hipMemcpyHostToDevice, c->stream));
  OSC_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace osc
