// Integrated-Green-function primitives, FP64, in the upstream operation order
// (the library is compiled with -ffp-contract=off so no FMA is formed).
#pragma once
#include <hip/hip_runtime.h>

namespace osc {

// xlafun2 (open_spacecharge_mod.f90:721-726):
//   x*atan(yz/(r x)) - z*log(r+y) + y*log((r-z)/(r+z))/2
__host__ __device__ inline double xlafun2(double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return x * atan((y * z) / (r * x)) - z * log(r + y) + y * log((r - z) / (r + z)) / 2.0;
}

// lafun2 (open_spacecharge_mod.f90:741-747)
__host__ __device__ inline double lafun2(double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return -(z * z * atan(x * y / (z * r)) / 2.0) - y * y * atan(x * z / (y * r)) / 2.0 -
         x * x * atan(y * z / (x * r)) / 2.0 + y * z * log(x + r) + x * z * log(y + r) +
         x * y * log(z + r);
}

// lafun (core:229-238)
__host__ __device__ inline double lafun(double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return -(0.5 * (z * z) * atan(x * y / (z * r))) - 0.5 * (y * y) * atan(x * z / (y * r)) -
         0.5 * (x * x) * atan(y * z / (x * r)) + y * z * log(x + r) + x * z * log(y + r) +
         x * y * log(z + r);
}

// xlafun / ylafun / zlafun (core:377-428), with the upstream log guards
__host__ __device__ inline double xlafun(double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  double res = z - x * atan(z / x) + x * atan(y * z / (x * r));
  if (y + r != 0) res = res - z * log(y + r);
  if (z + r != 0) res = res - y * log(z + r);
  return res;
}
__host__ __device__ inline double ylafun(double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  double res = x - y * atan(x / y) + y * atan(z * x / (y * r));
  if (z + r != 0) res = res - x * log(z + r);
  if (x + r != 0) res = res - z * log(x + r);
  return res;
}
__host__ __device__ inline double zlafun(double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  double res = y - z * atan(y / z) + z * atan(x * y / (z * r));
  if (x + r != 0) res = res - y * log(x + r);
  if (y + r != 0) res = res - x * log(y + r);
  return res;
}

// coulombfun (core:189-200)
__host__ __device__ inline double coulombfun(double u, double v, double w, double gam) {
  if (u == 0.0 && v == 0.0 && w == 0.0) return 0.0;
  return 1.0 / sqrt(u * u + v * v + (gam * w) * (gam * w));
}

template <int IC>
__host__ __device__ inline double lafun_c(double x, double y, double z) {
  if constexpr (IC == 0) return lafun(x, y, z);
  else if constexpr (IC == 1) return xlafun(x, y, z);
  else if constexpr (IC == 2) return ylafun(x, y, z);
  else return zlafun(x, y, z);
}

template <int IC>
__host__ __device__ inline double igf8(double x1, double x2, double y1, double y2, double z1,
                                       double z2) {
  return lafun_c<IC>(x2, y2, z2) - lafun_c<IC>(x1, y2, z2) - lafun_c<IC>(x2, y1, z2) -
         lafun_c<IC>(x2, y2, z1) - lafun_c<IC>(x1, y1, z1) + lafun_c<IC>(x1, y1, z2) +
         lafun_c<IC>(x1, y2, z1) + lafun_c<IC>(x2, y1, z1);
}

// igfcoulombfun / igfexfun / igfeyfun / igfezfun (core:205-372); igf == 0 ->
// coulombfun (phi only; other components are never requested upstream)
__host__ __device__ inline double igf_legacy(int igf, int icomp, double u, double v, double w,
                                             double gam, double dx, double dy, double dz) {
  if (igf == 0) return icomp == 0 ? coulombfun(u, v, w, gam) : 0.0;
  double x1 = u - 0.5 * dx, x2 = u + 0.5 * dx;
  double y1 = v - 0.5 * dy, y2 = v + 0.5 * dy;
  double z1 = (w - 0.5 * dz) * gam, z2 = (w + 0.5 * dz) * gam;
  if (icomp == 0) return igf8<0>(x1, x2, y1, y2, z1, z2) / (dx * dy * dz * gam);
  double res;
  if (x1 < 0.0 && x2 > 0.0 && y1 < 0.0 && y2 > 0.0 && z1 < 0.0 && z2 > 0.0) {
    res = 0.0;
  } else if (icomp == 1) {
    res = igf8<1>(x1, x2, y1, y2, z1, z2);
  } else if (icomp == 2) {
    res = igf8<2>(x1, x2, y1, y2, z1, z2);
  } else {
    res = igf8<3>(x1, x2, y1, y2, z1, z2);
  }
  if (icomp == 3) return res / (dx * dy * dz * gam);
  return res / (dx * dy * dz);
}

// rfun (core:723-755): 5x5 mode sum of the rectangular-pipe IGF
__host__ __device__ inline double rfun(double u, double v, double w, double gam, double a,
                                       double b, double hz, int i, int j) {
  const double pi = 3.141592653589793115997963468544185161590576171875;  // acos(-1d0)
  double ainv = 1.0 / a, binv = 1.0 / b;
  double piainv = pi * ainv, pibinv = pi * binv;
  double res = 0.0;
  for (int m = 1; m <= 5; ++m)
    for (int n = 1; n <= 5; ++n) {
      double km = m * pi * ainv, kn = n * pi * binv;
      double kapmn = sqrt(km * km + kn * kn);
      double zfun = (exp(-(kapmn * fabs(gam * w - hz))) - 2.0 * exp(-(kapmn * fabs(gam * w))) +
                     exp(-(kapmn * fabs(gam * w + hz)))) /
                    (hz * hz * (kapmn * kapmn));
      if (w == 0) zfun = zfun + 2.0 / (hz * kapmn);
      int im = 1, jn = 1;
      for (int t = 0; t < m; ++t) im *= i;
      for (int t = 0; t < n; ++t) jn *= j;
      double term = (double)(im * jn) * cos(m * u * piainv) * cos(n * v * pibinv) * zfun / kapmn;
      res = res + term;
    }
  res = res * 2.0 * pi * ainv * binv;
  return res;
}

}  // namespace osc
