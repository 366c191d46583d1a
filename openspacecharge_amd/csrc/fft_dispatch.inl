// Definitions of the launchers for one precision; included by fft_f64.hip and
// fft_f32.hip with OSC_REAL defined.
#include "fft_launch.h"

namespace osc {

namespace {
using T = OSC_REAL;

inline dim3 grid3(int64_t x, int64_t y, int64_t z) {
  return dim3((unsigned)x, (unsigned)y, (unsigned)z);
}
inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

size_t gen_lds_bytes(int L) { return size_t(LdsPad<T>::size(2 * kGenBX * L)) * sizeof(cpx<T>); }

template <int L, int V>
void axis_tv(int dir, const cpx<T>* in, cpx<T>* out, const AxisIO& a, int g1, int g2,
             const cpx<T>* tw, hipStream_t s) {
  constexpr int BX = AxisTiles<T, L>::bx(V), NT = BX * Geo<T, L>::TPL;
  dim3 g = grid3(cdiv(a.nb, BX), g1, g2);
  if (dir < 0 && a.mh > 0)
    launch((k_axis<T, L, -1, true, BX>), g, dim3(NT), 0, s, in, out, a, tw);
  else if (dir < 0)
    launch((k_axis<T, L, -1, false, BX>), g, dim3(NT), 0, s, in, out, a, tw);
  else
    launch((k_axis<T, L, 1, false, BX>), g, dim3(NT), 0, s, in, out, a, tw);
}
// the tile width the tuner picked for this pass (fft_tune_*)
template <int L>
void axis_t(int dir, const cpx<T>* in, cpx<T>* out, const AxisIO& a, int g1, int g2,
            const cpx<T>* tw, hipStream_t s) {
  const int kind = dir > 0 ? 2 : (a.mh > 0 ? 1 : 0);
  const int64_t lines = int64_t(a.nb) * g1 * g2;
  bool distinct[3];
  for (int v = 0; v < 3; ++v) distinct[v] = v == 1 || AxisTiles<T, L>::bx(v) != AxisTiles<T, L>::bx(1);
  FftTune tune(int(sizeof(T)), L, kind, lines, distinct, s);
  switch (tune.variant()) {
    case 0:
      axis_tv<L, 0>(dir, in, out, a, g1, g2, tw, s);
      break;
    case 2:
      axis_tv<L, 2>(dir, in, out, a, g1, g2, tw, s);
      break;
    default:
      axis_tv<L, 1>(dir, in, out, a, g1, g2, tw, s);
  }
  tune.done(s);
}
template <int L>
void xr2c_t(const double* in, cpx<T>* out, const R2CIO& a, const cpx<T>* tw, hipStream_t s,
            const CornerDiff* cd) {
  constexpr int BX = XGeo<T, L>::BX, NT = XGeo<T, L>::NT;
  dim3 g = grid3(cdiv(cdiv(a.nrows, 2), BX), 1, 1);
  if (cd)
    launch((k_xr2c<T, L, true>), g, dim3(NT), 0, s, in, out, a, tw, *cd);
  else
    launch((k_xr2c<T, L, false>), g, dim3(NT), 0, s, in, out, a, tw, CornerDiff{});
}
template <int L>
void xc2r_t(const cpx<T>* in, const C2RIO<T>& a, int nout, const cpx<T>* tw, hipStream_t s) {
  constexpr int BX = XGeo<T, L>::BX, NT = XGeo<T, L>::NT;
  dim3 g = grid3(cdiv(cdiv(a.nrows, 2), BX), nout, 1);
  launch((k_xc2r<T, L>), g, dim3(NT), 0, s, in, a, tw);
}
// k_zconv's per-thread byte offsets (line b < BX, position t < TPL of a
// column block) must fit in 32 bits
template <int L>
bool zconv_fits_t(const ZinvIO<T>& a) {
  const int64_t s = a.sz > a.o_sz ? a.sz : a.o_sz;
  const int64_t most = (ZGeo<T, L>::BX - 1 + int64_t(PlanInfo<L>::TPL - 1) * s) *
                       int64_t(sizeof(cpx<T>));
  return most < (int64_t(1) << 32);
}

template <int L>
void zinv_t(const cpx<T>* rho, cpx<T>* out, const ZinvIO<T>& a, int Ly, int nout, const cpx<T>* tw,
            hipStream_t s) {
  constexpr int BX = Geo<T, L>::BX, NT = BX * Geo<T, L>::TPL;
  dim3 g = grid3(cdiv(a.nb, BX), Ly, nout);
  if (a.zreal && a.zin > 0) {
    OSC_REQUIRE(zconv_fits_t<L>(a), OSC_ERR_ARG, "fused z convolution: offsets exceed 32 bits");
    const int nxb = int(cdiv(a.nb, ZGeo<T, L>::BX));
    launch((k_zconv<T, L>), grid3(int64_t(nxb) * Ly, 1, 1), dim3(ZGeo<T, L>::NT), 0,
                       s, rho, out, a, tw, nxb);
  } else if (a.zreal) {
    const int nxb = int(cdiv(a.nb, ZGeo<T, L>::BX));
    const int64_t units = int64_t(nxb) * Ly;
    const dim3 g1 = grid3(cdiv(units, 8) * 8 * nout, 1, 1);
    launch((k_zinv_g<T, L>), g1, dim3(ZGeo<T, L>::NT), 0, s, rho, out, a, tw, nxb);
  } else
    launch((k_zinv<T, L>), g, dim3(NT), 0, s, rho, out, a, tw);
}
}  // namespace

template <>
void launch_axis<T>(int L, int dir, const cpx<T>* in, cpx<T>* out, const AxisIO& a, int g1, int g2,
                    const cpx<T>* tw, hipStream_t s) {
  OSC_REQUIRE(a.mh == 0 || (dir < 0 && 2 * a.mh == L), OSC_ERR_ARG,
              "half-stored input lines are for forward passes over even lengths");
  switch (L) {
#define OSC_CASE(LL)                           \
  case LL:                                     \
    axis_t<LL>(dir, in, out, a, g1, g2, tw, s); \
    break;
    OSC_FOR_EACH_L(OSC_CASE)
#undef OSC_CASE
    default: {
      GenericPlan pl = generic_plan(L);
      dim3 g = grid3(cdiv(a.nb, kGenBX), g1, g2);
      if (dir < 0)
        launch((k_axis_gen<T, -1>), g, dim3(kGenBX * kGenTPL), gen_lds_bytes(L), s, in,
                           out, a, tw, pl);
      else
        launch((k_axis_gen<T, 1>), g, dim3(kGenBX * kGenTPL), gen_lds_bytes(L), s, in,
                           out, a, tw, pl);
    }
  }
  OSC_HIP(hipGetLastError());
}

template <>
void launch_xr2c<T>(int L, const double* in, cpx<T>* out, const R2CIO& a, const cpx<T>* tw,
                    hipStream_t s) {
  OSC_REQUIRE(a.nrows < (int64_t(1) << 31), OSC_ERR_ARG, "x pass: too many rows for 32-bit indices");
  switch (L) {
#define OSC_CASE(LL)                         \
  case LL:                                   \
    xr2c_t<LL>(in, out, a, tw, s, nullptr);  \
    break;
    OSC_FOR_EACH_L(OSC_CASE)
#undef OSC_CASE
    default: {
      GenericPlan pl = generic_plan(L);
      dim3 g = grid3(cdiv(cdiv(a.nrows, 2), kGenBX), 1, 1);
      launch((k_xr2c_gen<T>), g, dim3(kGenBX * kGenTPL), gen_lds_bytes(L), s, in, out,
                         a, tw, pl);
    }
  }
  OSC_HIP(hipGetLastError());
}

template <>
void launch_xr2c_diff<T>(int L, const double* F, cpx<T>* out, const R2CIO& a,
                         const CornerDiff& cd, const cpx<T>* tw, hipStream_t s) {
  OSC_REQUIRE(a.nrows < (int64_t(1) << 31), OSC_ERR_ARG, "x pass: too many rows for 32-bit indices");
  switch (L) {
#define OSC_CASE(LL)                     \
  case LL:                               \
    xr2c_t<LL>(F, out, a, tw, s, &cd);   \
    break;
    OSC_FOR_EACH_L(OSC_CASE)
#undef OSC_CASE
    default:
      OSC_REQUIRE(false, OSC_ERR_FFT_LENGTH, "corner-difference x pass needs a planned length");
  }
  OSC_HIP(hipGetLastError());
}

template <>
void launch_xc2r<T>(int L, const cpx<T>* in, const C2RIO<T>& a, int nout, const cpx<T>* tw,
                    hipStream_t s) {
  OSC_REQUIRE(a.nrows < (int64_t(1) << 31), OSC_ERR_ARG, "x pass: too many rows for 32-bit indices");
  switch (L) {
#define OSC_CASE(LL)                  \
  case LL:                            \
    xc2r_t<LL>(in, a, nout, tw, s);   \
    break;
    OSC_FOR_EACH_L(OSC_CASE)
#undef OSC_CASE
    default: {
      GenericPlan pl = generic_plan(L);
      dim3 g = grid3(cdiv(cdiv(a.nrows, 2), kGenBX), nout, 1);
      launch((k_xc2r_gen<T>), g, dim3(kGenBX * kGenTPL), gen_lds_bytes(L), s, in, a,
                         tw, pl);
    }
  }
  OSC_HIP(hipGetLastError());
}

template <>
bool zconv_fits<T>(int L, const ZinvIO<T>& a) {
  switch (L) {
#define OSC_CASE(LL) \
  case LL:           \
    return zconv_fits_t<LL>(a);
    OSC_FOR_EACH_L(OSC_CASE)
#undef OSC_CASE
    default:
      return false;
  }
}

template <>
void launch_zinv<T>(int L, const cpx<T>* rho, cpx<T>* out, const ZinvIO<T>& a, int Ly, int nout,
                    const cpx<T>* tw, hipStream_t s) {
  switch (L) {
#define OSC_CASE(LL)                              \
  case LL:                                        \
    zinv_t<LL>(rho, out, a, Ly, nout, tw, s);     \
    break;
    OSC_FOR_EACH_L(OSC_CASE)
#undef OSC_CASE
    default: {
      OSC_REQUIRE(!a.zreal && a.zin == 0, OSC_ERR_FFT_LENGTH,
                  "z-real Green arrays need a planned length");
      GenericPlan pl = generic_plan(L);
      dim3 g = grid3(cdiv(a.nb, kGenBX), Ly, nout);
      launch((k_zinv_gen<T>), g, dim3(kGenBX * kGenTPL), gen_lds_bytes(L), s, rho, out,
                         a, tw, pl);
    }
  }
  OSC_HIP(hipGetLastError());
}

}  // namespace osc
