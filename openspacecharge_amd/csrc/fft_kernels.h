// Pass kernels of the doubled-grid convolution pipeline and the generic
// per-axis complex FFT, built on fft_core.h.  Instantiated per precision in
// fft_f64.hip / fft_f32.hip.
//
// Spectrum layout: half spectrum S[kz][ky][kx], kx fastest with a padded row
// of KXP >= L/2+1 entries, ky rows of Ly, kz planes of Lz.
#pragma once
#include "fft_core.h"

namespace osc {

// ---------------------------------------------------------------------------
// argument blocks
// ---------------------------------------------------------------------------
// Batched strided line FFT.  Element (line b, position p) of grid block
// (o1, o2) is at base + b*s_b + p*s_p + o1*s_o1 + o2*s_o2 (input) and
// b*d_b + (p-w0)*d_p + o1*d_o1 + o2*d_o2 (output).  Positions >= n_in are
// zero and not read; only positions in [w0, w0+n_out) are written.
struct AxisIO {
  int64_t s_b, s_p, s_o1, s_o2;
  int64_t d_b, d_p, d_o1, d_o2;
  int nb, n_in, w0, n_out;
  // mh > 0: only elements [0, mh) of the input line are stored (mh = L/2), the
  // line being mirror-symmetric about L/2 - 1: element p in [mh, L-2] is
  // ms * element L-2-p and element L-1 is zero
  int mh, ms;
};

// element p of an axis input line (see AxisIO::mh)
template <typename T>
__device__ __forceinline__ cpx<T> axis_in(const cpx<T>* ib, const AxisIO& a, int L, int p) {
  if (p >= a.n_in) return mk<T>(0, 0);
  if (a.mh > 0 && p >= a.mh) {
    if (p == L - 1) return mk<T>(0, 0);
    const cpx<T> v = ib[(L - 2 - p) * a.s_p];
    return a.ms < 0 ? mk<T>(-v.x, -v.y) : v;
  }
  return ib[p * a.s_p];
}

// Real rows -> half spectra.  Row r (0 <= r < nrows) is (y, z) = (r % ny_rows,
// r / ny_rows); input row at in + y*in_sy + z*in_sz (n_in reals, rest zero),
// output row at out + y*out_sy + z*out_sz (L/2+1 complex).  Two rows are
// transformed together as the real and imaginary parts of one complex line.
struct R2CIO {
  int64_t in_sy, in_sz, out_sy, out_sz;
  int64_t nrows;
  int ny_rows, n_in;
};

// Epilogue of the inverse x pass for one output field: e = s1*val is stored
// to dst1 (if any); dst2 (if any) receives bmode 0: s2*e, or
// bmode 1: ((sgn*e)*beta)/c   (the two B-field formulas of the upstream code).
struct Epilogue {
  double* dst1;
  double* dst2;
  double s1, s2, beta, c;
  int bmode, pad_;
};

constexpr int kMaxOut = 6;
constexpr int kMaxTerm = 4;

// Half spectra -> real rows.  Rows as in R2CIO (spectrum side is "in"),
// output position x in [w0, w0+n_out) stored at dst + x-w0 + row offsets.
template <typename T>
struct C2RIO {
  int64_t in_sy, in_sz, in_so;
  int64_t out_sy, out_sz;
  int64_t nrows;
  int ny_rows, w0, n_out, pad_;
  Epilogue ep[kMaxOut];
};

// One spectral term: sgn * rho_hat(mirror(k)) * G(k)
template <typename T>
struct Term {
  const cpx<T>* g;
  T sgn;
  int mir;  // bit 0: x, bit 1: y, bit 2: z
  // z-real Green lines (ZinvIO::zreal): zmh > 0 when only z < zmh is stored,
  // the line being mirror-symmetric about Lz/2 - 1 with sign zms
  int zmh, zms;
};
template <typename T>
struct OutTerms {
  Term<T> t[kMaxTerm];
  int nt;
};
template <typename T>
struct ZinvIO {
  int64_t sy, sz;         // spectrum strides (KXP, KXP*Ly)
  int64_t o_sy, o_sz, o_so;  // output strides
  int Ly, Lz, nb, w0, n_out, nout;
  // 1: every output has one term, whose Green array is transformed along x
  // and y only; its z transform is taken per line inside the multiply
  // (k_zinv_g)
  int zreal;
  // > 0 (with zreal, every output one term with mir == 0): rho_hat has had
  // its x and y passes only, with zin stored z planes (the rest zero); its z
  // pass is taken inside the multiply, once per column for all outputs
  // (k_zconv)
  int zin;
  OutTerms<T> out[kMaxOut];
};

// ---------------------------------------------------------------------------
// launch geometry of the compile-time engine
// ---------------------------------------------------------------------------
template <typename T, int L>
struct Geo {
  static constexpr int TPL = PlanInfo<L>::TPL;
  // lines per workgroup: at least one 128-B cache line of columns (strided
  // passes read BX adjacent columns per row), within 1024 threads and 64 KB LDS
  static constexpr int bx_of(int tpl) {
    int bx = 256 / tpl;
    const int line = 128 / (int)sizeof(cpx<T>);
    if (bx < line) bx = line;
    if (bx < 1) bx = 1;
    while (bx > 1 && (bx * tpl > 1024 || LdsPad<T>::size(bx * L) * (int)sizeof(cpx<T>) > 65536))
      bx /= 2;
    return bx;
  }
  static constexpr int BX = PlanInfo<L>::ns ? bx_of(TPL) : 1;
  static constexpr int NT = BX * TPL;
};

// Tile widths (lines per workgroup) a strided pass can run with: half, the
// default and twice Geo's, where the threads (<= 1024) and LDS (<= 64 KB)
// allow, else the default.  The width changes only how many lines share a
// workgroup, not the arithmetic of a line: results are bit-identical.
template <typename T, int L>
struct AxisTiles {
  static constexpr int B0 = Geo<T, L>::BX, TPL = Geo<T, L>::TPL;
  static constexpr bool fits(int bx) {
    return bx >= 1 && bx * TPL <= 1024 &&
           LdsPad<T>::size(bx * L) * (int)sizeof(cpx<T>) <= 65536;
  }
  static constexpr int bx(int v) {
    return v == 0 ? (fits(B0 / 2) ? B0 / 2 : B0) : (v == 2 ? (fits(2 * B0) ? 2 * B0 : B0) : B0);
  }
};

// x-pass geometry (contiguous rows, so no cache-line constraint on the
// number of lines): half the lines of Geo per workgroup, half its LDS
template <typename T, int L>
struct XGeo {
  static constexpr int BX = Geo<T, L>::BX >= 2 ? Geo<T, L>::BX / 2 : 1;
  static constexpr int NT = BX * Geo<T, L>::TPL;
};

// generic engine geometry: one line per 64-thread workgroup
constexpr int kGenTPL = 64;
constexpr int kGenBX = 1;

template <typename T>
__device__ __forceinline__ cpx<T> ldc(const cpx<T>* p) {
  return *p;
}

// ---------------------------------------------------------------------------
// strided line FFT (y or z axis, or x with s_p = 1)
// ---------------------------------------------------------------------------
// MIR: the input lines are stored as halves (AxisIO::mh)
template <typename T, int L, int DIR, bool MIR, int BX = Geo<T, L>::BX>
__global__ void __launch_bounds__((BX * Geo<T, L>::TPL))
    k_axis(const cpx<T>* __restrict__ in, cpx<T>* out, AxisIO a, const cpx<T>* __restrict__ tw) {
  using F = FftTile<T, L, DIR, BX, false>;
  __shared__ cpx<T> lds[F::LDS];
  F f;
  // (bx, by, bz): the logical workgroup, XCD-grouped (xcd_block)
  const uint32_t gx = gridDim.x, gy = gridDim.y;
  const uint32_t lw = xcd_block(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
  const int bx = int(lw % gx), by = int((lw / gx) % gy), bz = int(lw / (gx * gy));
  const int bg = bx * BX + f.b;
  const bool ok = bg < a.nb;
  const cpx<T>* ib = in + (ok ? bg * a.s_b + by * a.s_o1 + bz * a.s_o2 : 0);
  if constexpr (MIR)
    f.load([&](int, int p) { return ok ? axis_in(ib, a, L, p) : mk<T>(0, 0); });
  else
    f.load([&](int, int p) { return (ok && p < a.n_in) ? ib[p * a.s_p] : mk<T>(0, 0); });
  f.run(lds, tw);
  cpx<T>* ob = out + (ok ? bg * a.d_b + by * a.d_o1 + bz * a.d_o2 : 0);
  f.store([&](int, int p, cpx<T> v) {
    const int q = p - a.w0;
    if (ok && q >= 0 && q < a.n_out) ob[q * a.d_p] = v;
  });
}

template <typename T, int DIR>
__global__ void __launch_bounds__(kGenBX* kGenTPL)
    k_axis_gen(const cpx<T>* __restrict__ in, cpx<T>* out, AxisIO a, const cpx<T>* __restrict__ tw,
               GenericPlan pl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  FftGeneric<T, DIR, kGenBX, kGenTPL> f(pl.L);
  const int bg = blockIdx.x * kGenBX + f.b;
  const bool ok = bg < a.nb;
  const cpx<T>* ib = in + (ok ? bg * a.s_b + blockIdx.y * a.s_o1 + blockIdx.z * a.s_o2 : 0);
  f.load(lds, [&](int, int p) { return ok ? axis_in(ib, a, pl.L, p) : mk<T>(0, 0); });
  f.run(pl, tw);
  cpx<T>* ob = out + (ok ? bg * a.d_b + blockIdx.y * a.d_o1 + blockIdx.z * a.d_o2 : 0);
  f.store([&](int, int p, cpx<T> v) {
    const int q = p - a.w0;
    if (ok && q >= 0 && q < a.n_out) ob[q * a.d_p] = v;
  });
}

// ---------------------------------------------------------------------------
// forward x pass: pairs of real rows -> half spectra
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void r2c_unpack(cpx<T> zk, cpx<T> zm, cpx<T>& A, cpx<T>& B) {
  // A = (Z[k] + conj Z[L-k]) / 2 ;  B = (Z[k] - conj Z[L-k]) / (2i)
  cpx<T> s = add(zk, conjg(zm)), d = sub(zk, conjg(zm));
  A = scale(s, T(0.5));
  B = mk(T(0.5) * d.y, T(-0.5) * d.x);
}

// rows ra (even) and rb = ra + 1 of an x pass as (y, z) = (r % ny_rows,
// r / ny_rows): one 32-bit division per thread (row counts are below 2^31,
// checked by the launchers)
struct RowPair {
  int ya, za, yb, zb;
};
__device__ __forceinline__ RowPair row_pair(int64_t ra, int ny_rows) {
  RowPair p;
  const uint32_t r = uint32_t(ra), ny = uint32_t(ny_rows);
  p.za = int(r / ny);
  p.ya = int(r - uint32_t(p.za) * ny);
  if (p.ya + 1 == ny_rows) {
    p.yb = 0;
    p.zb = p.za + 1;
  } else {
    p.yb = p.ya + 1;
    p.zb = p.za;
  }
  return p;
}

// The real rows of the x pass taken as the 8-corner difference of a primitive
// F (open_spacecharge_mod.f90:697-705): row (j, kq) of the result is built
// from rows (j, k), (j+1, k), (j, k+1), (j+1, k+1) of F with
// k = (kq - sh) mod L2; elements on the last plane of any axis keep the raw
// primitive, or 0 with zero_raw.  F has the strides of R2CIO's input and
// extents (n_in, L1, L2); R2CIO's rows may cover only the lower part in y
// (ny_rows < L1) and z.
struct CornerDiff {
  int L1, L2, sh, zero_raw;
  // far field (far != 0): an interior row whose cells lie at |w| >= wnear
  // from the origin is not differenced from the primitive but takes each
  // cell's integral of the field kernel x_a / r^3 (a = icomp - 1) from its
  // second-order expansion about the cell centre, times factor; h: the cell
  // edges, c0: the centre coordinates of cell 0 along each axis
  int far, icomp;
  double wnear, factor;
  double h[3], c0[3];
};

// row (j, kq) of the corner-difference source: its F row, whether it is
// interior and, for an interior row in the far field, far = true
__device__ __forceinline__ const double* diff_row(const double* F, const R2CIO& a,
                                                  const CornerDiff& cd, int j, int kq, bool ok,
                                                  bool& inner, bool& far) {
  inner = false;
  far = false;
  if (!ok) return F;
  int k = kq - cd.sh;  // kq < L2 and 0 <= sh < L2: one wrap at most
  if (k < 0) k += cd.L2;
  inner = j < cd.L1 - 1 && k < cd.L2 - 1;
  far = cd.far && inner && fabs(cd.c0[2] + k * cd.h[2]) >= cd.wnear;
  return F + j * a.in_sy + k * a.in_sz;
}

// the cell integral of x_a / r^3 over the cell centred at (u, v, w) with
// edges h, to second order: V (f + (hu^2 f_uu + hv^2 f_vv + hw^2 f_ww) / 24),
// the second derivatives 15 x_a x_b^2 / r^7 - 3 x_a / r^5 (b != a) and
// 15 x_a^3 / r^7 - 9 x_a / r^5 (b = a); times factor
__device__ __forceinline__ double far_cell(const CornerDiff& cd, double u, double v, double w) {
  const double r2 = u * u + v * v + w * w;
  const double ir2 = 1.0 / r2, ir3 = ir2 / sqrt(r2);
  const double xa = cd.icomp == 1 ? u : (cd.icomp == 2 ? v : w);
  const double f = xa * ir3;
  const double t15 = 15.0 * f * ir2 * ir2, t3 = 3.0 * f * ir2;
  double fuu = t15 * u * u - t3, fvv = t15 * v * v - t3, fww = t15 * w * w - t3;
  if (cd.icomp == 1) fuu -= 2.0 * t3;
  else if (cd.icomp == 2) fvv -= 2.0 * t3;
  else fww -= 2.0 * t3;
  const double hu = cd.h[0], hv = cd.h[1], hw = cd.h[2];
  return cd.factor * (hu * hv * hw) *
         (f + (hu * hu * fuu + hv * hv * fvv + hw * hw * fww) * (1.0 / 24.0));
}

// far_cell for an FP32 x pass: the cell centre formed in FP64 (its position
// is a large multiple of the edge), the expansion in FP32; K = factor times
// the cell volume, e2 = the squared edges over 24
__device__ __forceinline__ float far_cell_f(int icomp, float u, float v, float w, float K,
                                            float e2u, float e2v, float e2w) {
  const float r2 = u * u + v * v + w * w;
  const float ir = rsqrtf(r2), ir2 = ir * ir;
  const float xa = icomp == 1 ? u : (icomp == 2 ? v : w);
  const float f = xa * ir2 * ir;
  const float t15 = 15.0f * f * ir2 * ir2, t3 = 3.0f * f * ir2;
  float fuu = t15 * u * u - t3, fvv = t15 * v * v - t3, fww = t15 * w * w - t3;
  if (icomp == 1) fuu -= 2.0f * t3;
  else if (icomp == 2) fvv -= 2.0f * t3;
  else fww -= 2.0f * t3;
  return K * (f + e2u * fuu + e2v * fvv + e2w * fww);
}
// element p of a corner-difference row, with the terms in the order of the
// upstream loop (a left-to-right chain of adds and subtracts)
__device__ __forceinline__ double diff_elem(const double* q, const R2CIO& a, const CornerDiff& cd,
                                           bool inner, int p) {
  if (inner && p < a.n_in - 1) {
    const double* q10 = q + a.in_sy;
    const double* q01 = q + a.in_sz;
    const double* q11 = q01 + a.in_sy;
    return q11[p + 1] - q11[p] - q01[p + 1] - q10[p + 1] - q[p] + q01[p] + q10[p] + q[p + 1];
  }
  return cd.zero_raw ? 0.0 : q[p];
}

template <typename T, int L, bool DIFF>
__global__ void __launch_bounds__((XGeo<T, L>::NT))
    k_xr2c(const double* __restrict__ in, cpx<T>* __restrict__ out, R2CIO a,
           const cpx<T>* __restrict__ tw, CornerDiff cd) {
  constexpr int BX = XGeo<T, L>::BX;
  using F = FftTile<T, L, -1, BX, true>;
  __shared__ cpx<T> lds[F::LDS];
  F f;
  const int64_t ra = 2 * (int64_t(blockIdx.x) * BX + f.b), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const RowPair rp = row_pair(ra, a.ny_rows);
  if constexpr (DIFF) {
    bool ina, inb, fara, farb;
    const double* pa = diff_row(in, a, cd, rp.ya, rp.za, oka, ina, fara);
    const double* pb = diff_row(in, a, cd, rp.yb, rp.zb, okb, inb, farb);
    if (cd.far && (fara || farb)) {
      // far-field cell integrals (interior cells; the last cell of a row is
      // raw, i.e. zero, as zero_raw makes it)
      const double va = cd.c0[1] + rp.ya * cd.h[1], vb = cd.c0[1] + rp.yb * cd.h[1];
      int ka = rp.za - cd.sh, kb = rp.zb - cd.sh;
      if (ka < 0) ka += cd.L2;
      if (kb < 0) kb += cd.L2;
      const double wa = cd.c0[2] + ka * cd.h[2], wb = cd.c0[2] + kb * cd.h[2];
      if constexpr (sizeof(T) == 4) {
        const float K = float(cd.factor * (cd.h[0] * cd.h[1] * cd.h[2]));
        const float e2u = float(cd.h[0] * cd.h[0] / 24.0), e2v = float(cd.h[1] * cd.h[1] / 24.0),
                    e2w = float(cd.h[2] * cd.h[2] / 24.0);
        const float vaf = float(va), vbf = float(vb), waf = float(wa), wbf = float(wb);
        f.load([&](int, int p) {
          const float u = float(cd.c0[0] + p * cd.h[0]);
          T re = T(0), im = T(0);
          if (oka && p < a.n_in)
            re = fara ? (p < a.n_in - 1 ? far_cell_f(cd.icomp, u, vaf, waf, K, e2u, e2v, e2w)
                                        : T(0))
                      : T(diff_elem(pa, a, cd, ina, p));
          if (okb && p < a.n_in)
            im = farb ? (p < a.n_in - 1 ? far_cell_f(cd.icomp, u, vbf, wbf, K, e2u, e2v, e2w)
                                        : T(0))
                      : T(diff_elem(pb, a, cd, inb, p));
          return mk(re, im);
        });
      } else {
        f.load([&](int, int p) {
          const double u = cd.c0[0] + p * cd.h[0];
          T re = T(0), im = T(0);
          if (oka && p < a.n_in)
            re = fara ? (p < a.n_in - 1 ? T(far_cell(cd, u, va, wa)) : T(0))
                      : T(diff_elem(pa, a, cd, ina, p));
          if (okb && p < a.n_in)
            im = farb ? (p < a.n_in - 1 ? T(far_cell(cd, u, vb, wb)) : T(0))
                      : T(diff_elem(pb, a, cd, inb, p));
          return mk(re, im);
        });
      }
    } else {
      f.load([&](int, int p) {
        T re = (oka && p < a.n_in) ? T(diff_elem(pa, a, cd, ina, p)) : T(0);
        T im = (okb && p < a.n_in) ? T(diff_elem(pb, a, cd, inb, p)) : T(0);
        return mk(re, im);
      });
    }
  } else {
    const double* pa = in + (oka ? int64_t(rp.ya) * a.in_sy + int64_t(rp.za) * a.in_sz : 0);
    const double* pb = in + (okb ? int64_t(rp.yb) * a.in_sy + int64_t(rp.zb) * a.in_sz : 0);
    f.load([&](int, int p) {
      T re = (oka && p < a.n_in) ? T(pa[p]) : T(0);
      T im = (okb && p < a.n_in) ? T(pb[p]) : T(0);
      return mk(re, im);
    });
  }
  f.run(lds, tw);
  __syncthreads();
  f.dump(lds);
  __syncthreads();
  cpx<T>* oa = out + (oka ? int64_t(rp.ya) * a.out_sy + int64_t(rp.za) * a.out_sz : 0);
  cpx<T>* ob = out + (okb ? int64_t(rp.yb) * a.out_sy + int64_t(rp.zb) * a.out_sz : 0);
  for (int k = f.t; k <= L / 2; k += F::TPL) {
    cpx<T> A, B;
    r2c_unpack(lds[F::lidx(f.b, k)], lds[F::lidx(f.b, (L - k) % L)], A, B);
    if (oka) oa[k] = A;
    if (okb) ob[k] = B;
  }
}

template <typename T>
__global__ void __launch_bounds__(kGenBX* kGenTPL)
    k_xr2c_gen(const double* __restrict__ in, cpx<T>* __restrict__ out, R2CIO a,
               const cpx<T>* __restrict__ tw, GenericPlan pl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  using F = FftGeneric<T, -1, kGenBX, kGenTPL>;
  F f(pl.L);
  const int L = pl.L;
  const int64_t ra = 2 * (int64_t(blockIdx.x) * kGenBX + f.b), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const RowPair rp = row_pair(ra, a.ny_rows);
  const double* pa = in + (oka ? int64_t(rp.ya) * a.in_sy + int64_t(rp.za) * a.in_sz : 0);
  const double* pb = in + (okb ? int64_t(rp.yb) * a.in_sy + int64_t(rp.zb) * a.in_sz : 0);
  f.load(lds, [&](int, int p) {
    T re = (oka && p < a.n_in) ? T(pa[p]) : T(0);
    T im = (okb && p < a.n_in) ? T(pb[p]) : T(0);
    return mk(re, im);
  });
  f.run(pl, tw);
  const int buf = f.result_buf();
  cpx<T>* oa = out + (oka ? int64_t(rp.ya) * a.out_sy + int64_t(rp.za) * a.out_sz : 0);
  cpx<T>* ob = out + (okb ? int64_t(rp.yb) * a.out_sy + int64_t(rp.zb) * a.out_sz : 0);
  for (int k = f.t; k <= L / 2; k += kGenTPL) {
    cpx<T> A, B;
    r2c_unpack(lds[f.lidx(f.b, k, buf)], lds[f.lidx(f.b, (L - k) % L, buf)], A, B);
    if (oka) oa[k] = A;
    if (okb) ob[k] = B;
  }
}

// ---------------------------------------------------------------------------
// inverse x pass: pairs of half spectra -> real rows, with FP64 epilogue
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ cpx<T> c2r_pack(const cpx<T>* ra, const cpx<T>* rb, bool okb, int p,
                                           int L) {
  // Z[p] = A[p] + i B[p] with the Hermitian extension A[L-k] = conj A[k];
  // the imaginary parts of the self-conjugate bins (0 and L/2) are dropped.
  const bool hi = 2 * p > L;
  const int k = hi ? L - p : p;
  cpx<T> A = ra[k];
  cpx<T> B = okb ? rb[k] : mk<T>(0, 0);
  if (k == 0 || 2 * k == L) {
    A.y = 0;
    B.y = 0;
  }
  if (hi) {
    A.y = -A.y;
    B.y = -B.y;
  }
  return mk(A.x - B.y, A.y + B.x);
}

__device__ __forceinline__ void epi_store(const Epilogue& e, int64_t off, double val) {
  const double v = e.s1 * val;
  if (e.dst1) e.dst1[off] = v;
  if (e.dst2) e.dst2[off] = e.bmode == 0 ? e.s2 * v : ((e.s2 * v) * e.beta) / e.c;
}

template <typename T, int L>
__global__ void __launch_bounds__((XGeo<T, L>::NT))
    k_xc2r(const cpx<T>* __restrict__ in, C2RIO<T> a, const cpx<T>* __restrict__ tw) {
  constexpr int BX = XGeo<T, L>::BX;
  using F = FftTile<T, L, 1, BX, true>;
  __shared__ cpx<T> lds[F::LDS];
  F f;
  const int o = blockIdx.y;
  const int64_t ra = 2 * (int64_t(blockIdx.x) * BX + f.b), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const RowPair rp = row_pair(ra, a.ny_rows);
  const cpx<T>* base = in + o * a.in_so;
  const cpx<T>* pa = base + (oka ? int64_t(rp.ya) * a.in_sy + int64_t(rp.za) * a.in_sz : 0);
  const cpx<T>* pb = base + (okb ? int64_t(rp.yb) * a.in_sy + int64_t(rp.zb) * a.in_sz : 0);
  f.load([&](int, int p) { return oka ? c2r_pack(pa, pb, okb, p, L) : mk<T>(0, 0); });
  f.run(lds, tw);
  const Epilogue e = a.ep[o];
  const int64_t offa = int64_t(rp.ya) * a.out_sy + int64_t(rp.za) * a.out_sz;
  const int64_t offb = int64_t(rp.yb) * a.out_sy + int64_t(rp.zb) * a.out_sz;
  f.store([&](int, int p, cpx<T> v) {
    const int x = p - a.w0;
    if (x < 0 || x >= a.n_out) return;
    if (oka) epi_store(e, offa + x, double(v.x));
    if (okb) epi_store(e, offb + x, double(v.y));
  });
}

template <typename T>
__global__ void __launch_bounds__(kGenBX* kGenTPL)
    k_xc2r_gen(const cpx<T>* __restrict__ in, C2RIO<T> a, const cpx<T>* __restrict__ tw,
               GenericPlan pl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  FftGeneric<T, 1, kGenBX, kGenTPL> f(pl.L);
  const int L = pl.L;
  const int o = blockIdx.y;
  const int64_t ra = 2 * (int64_t(blockIdx.x) * kGenBX + f.b), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const RowPair rp = row_pair(ra, a.ny_rows);
  const cpx<T>* base = in + o * a.in_so;
  const cpx<T>* pa = base + (oka ? int64_t(rp.ya) * a.in_sy + int64_t(rp.za) * a.in_sz : 0);
  const cpx<T>* pb = base + (okb ? int64_t(rp.yb) * a.in_sy + int64_t(rp.zb) * a.in_sz : 0);
  f.load(lds, [&](int, int p) { return oka ? c2r_pack(pa, pb, okb, p, L) : mk<T>(0, 0); });
  f.run(pl, tw);
  const Epilogue e = a.ep[o];
  const int64_t offa = int64_t(rp.ya) * a.out_sy + int64_t(rp.za) * a.out_sz;
  const int64_t offb = int64_t(rp.yb) * a.out_sy + int64_t(rp.zb) * a.out_sz;
  f.store([&](int, int p, cpx<T> v) {
    const int x = p - a.w0;
    if (x < 0 || x >= a.n_out) return;
    if (oka) epi_store(e, offa + x, double(v.x));
    if (okb) epi_store(e, offb + x, double(v.y));
  });
}

// ---------------------------------------------------------------------------
// spectral combine + inverse z pass
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ cpx<T> combine(const cpx<T>* __restrict__ rho, const ZinvIO<T>& a,
                                          const OutTerms<T>& ot, int kx, int ky, int kz) {
  cpx<T> acc = mk<T>(0, 0);
  const int64_t gi = kx + ky * a.sy + kz * a.sz;
  for (int t = 0; t < ot.nt; ++t) {
    const Term<T> tm = ot.t[t];
    bool my, mz, cj;
    if (tm.mir & 1) {  // rho_hat(-kx,ky',kz') = conj(rho_hat(kx,-ky',-kz'))
      my = !(tm.mir & 2);
      mz = !(tm.mir & 4);
      cj = true;
    } else {
      my = tm.mir & 2;
      mz = tm.mir & 4;
      cj = false;
    }
    const int yy = my ? (ky == 0 ? 0 : a.Ly - ky) : ky;
    const int zz = mz ? (kz == 0 ? 0 : a.Lz - kz) : kz;
    cpx<T> r = rho[kx + yy * a.sy + zz * a.sz];
    if (cj) r.y = -r.y;
    cpx<T> p = mul(r, tm.g[gi]);
    acc = add(acc, mk(tm.sgn * p.x, tm.sgn * p.y));
  }
  return acc;
}

template <typename T, int L>
__global__ void __launch_bounds__((Geo<T, L>::NT))
    k_zinv(const cpx<T>* __restrict__ rho, cpx<T>* __restrict__ out, ZinvIO<T> a,
           const cpx<T>* __restrict__ tw) {
  constexpr int BX = Geo<T, L>::BX;
  using F = FftTile<T, L, 1, BX, false>;
  __shared__ cpx<T> lds[F::LDS];
  F f;
  // one output per workgroup: measured faster than one workgroup doing all
  // outputs of a column.  The outputs of a column block are consecutive
  // logical workgroups, and each XCD runs a contiguous range of them
  // (xcd_block), so a column's rho_hat lines, read once per output and term,
  // and the lines neighbouring column blocks share are fetched into one L2
  const uint32_t gx = gridDim.x, gy = gridDim.y, no = gridDim.z;
  const uint32_t lw = xcd_block(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * no);
  const int o = int(lw % no);
  const uint32_t u = lw / no;
  const int kx = int(u % gx) * BX + f.b, ky = int(u / gx);
  const bool ok = kx < a.nb;
  const OutTerms<T>& ot = a.out[o];
  f.load([&](int, int p) { return ok ? combine(rho, a, ot, kx, ky, p) : mk<T>(0, 0); });
  f.run(lds, tw);
  cpx<T>* ob = out + o * a.o_so + kx + ky * a.o_sy;
  f.store([&](int, int p, cpx<T> v) {
    const int q = p - a.w0;
    if (ok && q >= 0 && q < a.n_out) ob[q * a.o_sz] = v;
  });
}

// element kz of the rho_hat factor of a term (the mirror and conjugation
// rules of combine)
template <typename T>
__device__ __forceinline__ cpx<T> rho_term(const cpx<T>* __restrict__ rho, const ZinvIO<T>& a,
                                           const Term<T>& tm, int kx, int ky, int kz) {
  bool my, mz, cj;
  if (tm.mir & 1) {
    my = !(tm.mir & 2);
    mz = !(tm.mir & 4);
    cj = true;
  } else {
    my = tm.mir & 2;
    mz = tm.mir & 4;
    cj = false;
  }
  const int yy = my ? (ky == 0 ? 0 : a.Ly - ky) : ky;
  const int zz = mz ? (kz == 0 ? 0 : a.Lz - kz) : kz;
  cpx<T> r = rho[kx + yy * a.sy + zz * a.sz];
  if (cj) r.y = -r.y;
  return r;
}

// element p of a z-real Green line (Term::zmh mirror rule)
template <typename T, int L>
__device__ __forceinline__ cpx<T> zreal_in(const cpx<T>* gl, int64_t sz, const Term<T>& tm, int p) {
  if (tm.zmh > 0 && p >= tm.zmh) {
    if (p == L - 1) return mk<T>(0, 0);
    const cpx<T> v = gl[(L - 2 - p) * sz];
    return tm.zms < 0 ? mk<T>(-v.x, -v.y) : v;
  }
  return gl[p * sz];
}

// The multiply and inverse z pass for outputs of one term whose Green array
// is z-real: per (kx, ky) line the Green line is transformed along z here (it
// is never stored as a full spectrum), multiplied by its rho_hat factor into
// LDS and transformed back.
//
// One-dimensional grid, XCD-aware: output o of column unit u is logical
// workgroup 8*nout*(u/8) + 8*o + u%8 (xcd_block), so every XCD runs all
// outputs of a contiguous range of units: the second and later reads of a
// column's rho_hat line, and the neighbouring units' reads of the same
// 128-byte lines, hit that XCD's L2.  Units are (x block, ky) pairs, nxb x
// blocks per ky.
// lines per workgroup of k_zinv_g: in FP32 half the pass rule's, so that
// twice the workgroups fit the LDS of a CU (measured faster: 0.17 -> 0.155 ms
// at 256^3; in FP64 the half-line column reads cost more than that gains)
template <typename T, int L>
struct ZGeo {
  static constexpr int BX =
      sizeof(T) == 4 && Geo<T, L>::BX >= 4 ? Geo<T, L>::BX / 2 : Geo<T, L>::BX;
  static constexpr int NT = BX * Geo<T, L>::TPL;
};

template <typename T, int L>
__global__ void __launch_bounds__((ZGeo<T, L>::NT))
    k_zinv_g(const cpx<T>* __restrict__ rho, cpx<T>* __restrict__ out, ZinvIO<T> a,
             const cpx<T>* __restrict__ tw, int nxb) {
  constexpr int BX = ZGeo<T, L>::BX;
  using FF = FftTile<T, L, -1, BX, false>;
  using FI = FftTile<T, L, 1, BX, false>;
  static_assert(FF::LDS == FI::LDS, "forward and inverse tiles differ");
  __shared__ cpx<T> lds[FF::LDS];
  const int lw = int(xcd_block(blockIdx.x, gridDim.x));
  const int chunk = lw / (8 * a.nout), r = lw % (8 * a.nout);
  const int o = r / 8, u = chunk * 8 + r % 8;
  if (u >= nxb * a.Ly) return;  // whole workgroup: past the last unit
  const int xb = u % nxb, ky = u / nxb;
  const Term<T> tm = a.out[o].t[0];
  {
    FF ff;
    const int kx = xb * BX + ff.b;
    const bool ok = kx < a.nb;
    const cpx<T>* gl = tm.g + (ok ? kx + int64_t(ky) * a.sy : 0);
    ff.load([&](int, int p) { return ok ? zreal_in<T, L>(gl, a.sz, tm, p) : mk<T>(0, 0); });
    ff.run(lds, tw);
    __syncthreads();  // the transform's last reads of lds
    ff.store([&](int b, int p, cpx<T> g) {
      cpx<T> q = mk<T>(0, 0);
      if (ok) {
        q = mul(rho_term(rho, a, tm, kx, ky, p), g);
        q = mk(tm.sgn * q.x, tm.sgn * q.y);
      }
      lds[FF::lidx(b, p)] = q;
    });
  }
  __syncthreads();
  FI fi;
  const int kx = xb * BX + fi.b;
  const bool ok = kx < a.nb;
  fi.load([&](int b, int p) { return lds[FI::lidx(b, p)]; });
  __syncthreads();
  fi.run(lds, tw);
  cpx<T>* ob = out + o * a.o_so + kx + ky * a.o_sy;
  fi.store([&](int, int p, cpx<T> v) {
    const int q = p - a.w0;
    if (ok && q >= 0 && q < a.n_out) ob[q * a.o_sz] = v;
  });
}

// element at uniform byte address base plus a per-thread 32-bit byte offset
// (the global_load/store form with a scalar base)
template <typename T>
__device__ __forceinline__ const cpx<T>& at_off(const char* base, uint32_t vo) {
  return *reinterpret_cast<const cpx<T>*>(base + vo);
}

// The forward z pass of rho_hat and, for every output, the z transform of its
// z-real Green line, the multiply and the inverse z pass, per (kx, ky) column:
// rho_hat's column (ZinvIO::zin stored planes) is read once and its spectrum
// kept in registers across the outputs, and the forward and Green spectra
// meet in registers (same tile, same element positions).  Every output is a
// single term without mirrors.  One workgroup per (x block, ky) unit.
//
// Addressing: element (line b, position t + c) of a column block is at a
// uniform base (block, c) plus the per-thread byte offset of (b, t), computed
// once; the host checks that those offsets fit in 32 bits.
//
// Measured against variants at 256^3: handing the product to the inverse
// transform by a register permutation instead of through LDS (a thread's
// output positions are its input positions), with or without reading the
// next Green line during the inverse transform, needs about 150 VGPRs
// (3 waves per SIMD instead of 4) and is 5-10 % slower.
template <typename T, int L>
__global__ void __launch_bounds__((ZGeo<T, L>::NT))
    k_zconv(const cpx<T>* __restrict__ rho, cpx<T>* __restrict__ out, ZinvIO<T> a,
            const cpx<T>* __restrict__ tw, int nxb) {
  constexpr int BX = ZGeo<T, L>::BX;
  using FF = FftTile<T, L, -1, BX, false>;
  using FI = FftTile<T, L, 1, BX, false>;
  constexpr int TPL = FF::TPL;
  constexpr int64_t CB = sizeof(cpx<T>);
  static_assert(FF::LDS == FI::LDS, "forward and inverse tiles differ");
  __shared__ cpx<T> lds[FF::LDS];
  const int lw = int(xcd_block(blockIdx.x, gridDim.x));
  const int xb = lw % nxb, ky = lw / nxb;
  FF fr;
  const int b = fr.b, t = fr.t;
  const bool ok = xb * BX + b < a.nb;
  const int64_t col0 = xb * BX + int64_t(ky) * a.sy;  // column of line 0 (uniform)
  // per-thread byte offsets: (b, t) direct, (b, TPL-1-t) for mirrored reads
  const uint32_t vo = uint32_t((b + t * a.sz) * CB);
  const uint32_t vm = uint32_t((b + (TPL - 1 - t) * a.sz) * CB);
  const uint32_t vout = uint32_t((b + t * a.o_sz) * CB);
  const char* rb = reinterpret_cast<const char*>(rho + col0);
  fr.load_c([&](int c) {
    return (ok && t + c < a.zin) ? at_off<T>(rb + c * a.sz * CB, vo) : mk<T>(0, 0);
  });
  fr.run(lds, tw);
#pragma clang loop unroll(disable)
  for (int o = 0; o < a.nout; ++o) {
    const Term<T> tm = a.out[o].t[0];
    const char* gb = reinterpret_cast<const char*>(tm.g + col0);
    // the tiles' LDS index arithmetic stays inside the loop (hoisted, it would
    // hold two transforms' worth of addresses in registers across it)
    FF fg;
    asm volatile("" : "+v"(fg.b), "+v"(fg.t));
    // element p = t + c of a z-real Green line (zreal_in): mirrored at
    // p >= zmh to L-2-p = (L-1-TPL-c) + (TPL-1-t), zero at p = L-1
    fg.load_c([&](int c) {
      const int p = t + c;
      if (!ok) return mk<T>(0, 0);
      if (tm.zmh > 0 && p >= tm.zmh) {
        if (p == L - 1) return mk<T>(0, 0);
        const cpx<T> v = at_off<T>(gb + (L - 1 - TPL - c) * a.sz * CB, vm);
        return tm.zms < 0 ? mk<T>(-v.x, -v.y) : v;
      }
      return at_off<T>(gb + c * a.sz * CB, vo);
    });
    __syncthreads();  // the previous transform's last reads of lds
    fg.run(lds, tw);
    // the term's sign: a multiplication by +1 is the identity (uniform
    // branch; the same values as sgn * q).  FP32 only: in FP64 the second
    // copy of the loop costs registers (124 -> 138 VGPRs, a wave per SIMD)
    if (sizeof(T) == 4 && tm.sgn == T(1)) {
#pragma unroll
      for (int i = 0; i < FF::E; ++i) fg.v[i] = mul(fr.v[i], fg.v[i]);
    } else {
#pragma unroll
      for (int i = 0; i < FF::E; ++i) {
        const cpx<T> q = mul(fr.v[i], fg.v[i]);
        fg.v[i] = mk(tm.sgn * q.x, tm.sgn * q.y);
      }
    }
    __syncthreads();  // the transform's last reads of lds
    fg.dump(lds);
    __syncthreads();
    FI fi;
    asm volatile("" : "+v"(fi.b), "+v"(fi.t));
    fi.load([&](int bb, int p) { return lds[FI::lidx(bb, p)]; });
    __syncthreads();
    fi.run(lds, tw);
    char* ob = reinterpret_cast<char*>(out + o * a.o_so + col0 - int64_t(ky) * a.sy +
                                       int64_t(ky) * a.o_sy);
    fi.store_c([&](int c, cpx<T> v) {
      const int q = t + c - a.w0;
      if (ok && q >= 0 && q < a.n_out)
        *reinterpret_cast<cpx<T>*>(ob + (c - a.w0) * a.o_sz * CB + vout) = v;
    });
  }
}

template <typename T>
__global__ void __launch_bounds__(kGenBX* kGenTPL)
    k_zinv_gen(const cpx<T>* __restrict__ rho, cpx<T>* __restrict__ out, ZinvIO<T> a,
               const cpx<T>* __restrict__ tw, GenericPlan pl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  FftGeneric<T, 1, kGenBX, kGenTPL> f(pl.L);
  const int kx = blockIdx.x * kGenBX + f.b, ky = blockIdx.y, o = blockIdx.z;
  const bool ok = kx < a.nb;
  const OutTerms<T>& ot = a.out[o];
  f.load(lds, [&](int, int p) { return ok ? combine(rho, a, ot, kx, ky, p) : mk<T>(0, 0); });
  f.run(pl, tw);
  cpx<T>* ob = out + o * a.o_so + kx + ky * a.o_sy;
  f.store([&](int, int p, cpx<T> v) {
    const int q = p - a.w0;
    if (ok && q >= 0 && q < a.n_out) ob[q * a.o_sz] = v;
  });
}

}  // namespace osc
