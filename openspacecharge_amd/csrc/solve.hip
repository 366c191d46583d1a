// Doubled-grid / period-L convolution pipeline and the solver assemblies.
//
// Every upstream solver is expressed as: forward R2C of rho once, then for
// each output field one inverse transform of a sum of spectral terms
// sgn * rho_hat(mirror(k)) * G_hat(k).  The upstream per-axis FFT sign flips
// (correlations, core:682-713, 790-806, 1228, 1254) become frequency mirrors
// of rho_hat, and its z-flipped image charges (open_spacecharge_mod.f90:475,
// core:1047-1057) become a mirror plus a circular shift of the Green array.
#include "osc_internal.h"

namespace osc {

// speed of light of the legacy B-field formula (core:51,555,929)
constexpr double kClightDp = 299792458.0;

namespace {
constexpr int kThreads = 256;
inline int64_t nblk(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  return b > 65536 * 8 ? 65536 * 8 : b;
}

__global__ void k_fill(double* p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// finite-difference gradient with one-sided x2 edges (core:87-104)
__global__ void k_fd(const double* __restrict__ phi, int nx, int ny, int nz, double dx, double dy,
                     double dz, double gam0, double* __restrict__ ef) {
  const int64_t n = int64_t(nx) * ny * nz;
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < n;
       c += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(c % nx), j = int((c / nx) % ny), k = int(c / (int64_t(nx) * ny));
    const int im1 = i - 1 < 0 ? i : i - 1, ip1 = i + 1 > nx - 1 ? i : i + 1;
    const int jm1 = j - 1 < 0 ? j : j - 1, jp1 = j + 1 > ny - 1 ? j : j + 1;
    const int km1 = k - 1 < 0 ? k : k - 1, kp1 = k + 1 > nz - 1 ? k : k + 1;
    const double xfac = (im1 == i || ip1 == i) ? 2.0 : 1.0;
    const double yfac = (jm1 == j || jp1 == j) ? 2.0 : 1.0;
    const double zfac = (km1 == k || kp1 == k) ? 2.0 : 1.0;
    const int64_t sy = nx, sz = int64_t(nx) * ny;
    ef[c] = -((phi[ip1 + j * sy + k * sz] - phi[im1 + j * sy + k * sz]) / (2.0 * dx) * gam0 * xfac);
    ef[n + c] =
        -((phi[i + jp1 * sy + k * sz] - phi[i + jm1 * sy + k * sz]) / (2.0 * dy) * gam0 * yfac);
    ef[2 * n + c] =
        -((phi[i + j * sy + kp1 * sz] - phi[i + j * sy + km1 * sz]) / (2.0 * dz) / gam0 * zfac);
  }
}

// magnetic field of the legacy solvers (core:127-136, 613-622)
__global__ void k_bcore(const double* __restrict__ ef, int64_t n, double beta0, double cl,
                        double* __restrict__ bf) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    bf[i] = -ef[n + i] * beta0 / cl;
    bf[n + i] = ef[i] * beta0 / cl;
    bf[2 * n + i] = 0.0;
  }
}

// rho(:,:,k) <-> rho(:,:,nz+1-k) (core:1047-1057)
__global__ void k_flipz(double* rho, int nx, int ny, int nz) {
  const int64_t plane = int64_t(nx) * ny;
  const int64_t n = plane * (nz / 2);
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < n;
       c += int64_t(gridDim.x) * blockDim.x) {
    const int64_t ij = c % plane, k = c / plane;
    const int64_t a = ij + k * plane, b = ij + (nz - 1 - k) * plane;
    const double r1 = rho[a];
    rho[a] = rho[b];
    rho[b] = r1;
  }
}
}  // namespace

void fill_f64(osc_ctx* c, double* p, int64_t n, double v) {
  if (n <= 0) return;
  launch(k_fill, dim3(nblk(n)), dim3(kThreads), 0, c->stream, p, n, v);
  OSC_HIP(hipGetLastError());
}

void fd_gradient(osc_ctx* c, const double* phi, const int n[3], const double delta[3], double gam,
                 double* ef) {
  const int64_t nt = int64_t(n[0]) * n[1] * n[2];
  ProfScope ps(c, ST_FD_B);
  launch(k_fd, dim3(nblk(nt)), dim3(kThreads), 0, c->stream, phi, n[0], n[1], n[2],
                     delta[0], delta[1], delta[2], gam, ef);
  OSC_HIP(hipGetLastError());
}

void bfield_core(osc_ctx* c, const double* ef, int64_t n, double gam, double* bf) {
  const double beta0 = sqrt(1.0 - 1.0 / (gam * gam));
  ProfScope ps(c, ST_FD_B);
  launch(k_bcore, dim3(nblk(n)), dim3(kThreads), 0, c->stream, ef, n, beta0,
                     kClightDp, bf);
  OSC_HIP(hipGetLastError());
}

void flip_z(osc_ctx* c, double* rho, const int n[3]) {
  const int64_t nt = int64_t(n[0]) * n[1] * (n[2] / 2);
  if (nt == 0) return;
  launch(k_flipz, dim3(nblk(nt)), dim3(kThreads), 0, c->stream, rho, n[0], n[1], n[2]);
  OSC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// the pipeline
// ---------------------------------------------------------------------------
template <typename T>
void conv_solve(osc_ctx* c, const double* rho, const int n[3], const SpecDims& d, const int win[3],
                const std::vector<OutSpec>& outs) {
  const int nx = n[0], ny = n[1], nz = n[2];
  const int L0 = d.L[0], L1 = d.L[1], L2 = d.L[2];
  const int64_t KXP = d.KXP;
  OSC_REQUIRE(outs.size() >= 1 && outs.size() <= size_t(kMaxOut), OSC_ERR_ARG, "bad output count");
  for (int a = 0; a < 3; ++a)
    OSC_REQUIRE(fft_len_supported(d.L[a]), OSC_ERR_FFT_LENGTH,
                "transform length " + std::to_string(d.L[a]) + " unsupported");
  hipStream_t s = c->stream;
  DevBuf S(c->pool, sizeof(cpx<T>) * d.size());
  cpx<T>* sp = S.as<cpx<T>>();
  // forward: x (rows y<ny, z<nz), y (planes z<nz), z (all rows)
  R2CIO r{};
  r.in_sy = nx;
  r.in_sz = int64_t(nx) * ny;
  r.out_sy = KXP;
  r.out_sz = KXP * L1;
  r.nrows = int64_t(ny) * nz;
  r.ny_rows = ny;
  r.n_in = nx;
  {
    ProfScope ps(c, ST_FWD_X);
    launch_xr2c<T>(L0, rho, sp, r, twiddles<T>(c, L0), s);
  }
  AxisIO ay{};
  ay.s_b = ay.d_b = 1;
  ay.s_p = ay.d_p = KXP;
  ay.s_o1 = ay.d_o1 = KXP * L1;
  ay.nb = d.KX;
  ay.n_in = ny;
  ay.w0 = 0;
  ay.n_out = L1;
  {
    ProfScope ps(c, ST_FWD_Y);
    launch_axis<T>(L1, -1, sp, sp, ay, nz, 1, twiddles<T>(c, L1), s);
  }
  // outputs that are each one unmirrored term over a z-real Green array take
  // rho_hat's z pass inside the multiply (once per column for all outputs)
  ZinvIO<T> zs{};
  zs.sz = zs.o_sz = KXP * L1;
  bool zconv = fft_has_plan(L2) && zconv_fits<T>(L2, zs);
  for (const OutSpec& os : outs)
    zconv = zconv && os.nt == 1 && os.t[0].mir == 0 && os.t[0].g.zr;
  if (!zconv) {
    AxisIO az{};
    az.s_b = az.d_b = 1;
    az.s_p = az.d_p = KXP * L1;
    az.s_o1 = az.d_o1 = KXP;
    az.nb = d.KX;
    az.n_in = nz;
    az.w0 = 0;
    az.n_out = L2;
    ProfScope ps(c, ST_FWD_Z);
    launch_axis<T>(L2, -1, sp, sp, az, L1, 1, twiddles<T>(c, L2), s);
  }
  // inverse: combine + z, y (window), x (window + epilogue)
  const int nout = int(outs.size());
  const int64_t so = KXP * L1 * nz;
  DevBuf Tb(c->pool, sizeof(cpx<T>) * so * nout);
  cpx<T>* tp = Tb.as<cpx<T>>();
  ZinvIO<T> zi{};
  zi.sy = KXP;
  zi.sz = KXP * L1;
  zi.o_sy = KXP;
  zi.o_sz = KXP * L1;
  zi.o_so = so;
  zi.Ly = L1;
  zi.Lz = L2;
  zi.nb = d.KX;
  zi.w0 = win[2];
  zi.n_out = nz;
  zi.nout = nout;
  zi.zreal = outs[0].t[0].g.zr;
  zi.zin = zconv ? nz : 0;
  bool multi = false;
  for (int o = 0; o < nout; ++o) {
    zi.out[o].nt = outs[o].nt;
    multi = multi || outs[o].nt > 1;
    for (int t = 0; t < outs[o].nt; ++t) {
      const GreenRef& gr = outs[o].t[t].g;
      OSC_REQUIRE(gr.zr == zi.zreal, OSC_ERR_ARG, "Green arrays of one solve differ in z form");
      zi.out[o].t[t].g = static_cast<const cpx<T>*>(gr.g);
      zi.out[o].t[t].sgn = T(outs[o].t[t].sgn);
      zi.out[o].t[t].mir = outs[o].t[t].mir;
      zi.out[o].t[t].zmh = gr.zmh;
      zi.out[o].t[t].zms = gr.zms;
    }
  }
  // outputs summing several terms take full spectra: the z pass of each
  // distinct z-real Green array into a scratch spectrum
  std::vector<DevBuf> zfull;
  if (zi.zreal && multi) {
    std::map<const void*, const cpx<T>*> done;
    ProfScope ps(c, ST_GREEN_FFT);
    for (int o = 0; o < nout; ++o)
      for (int t = 0; t < zi.out[o].nt; ++t) {
        Term<T>& tm = zi.out[o].t[t];
        auto it = done.find(tm.g);
        if (it == done.end()) {
          zfull.emplace_back(c->pool, sizeof(cpx<T>) * d.size());
          cpx<T>* full = zfull.back().template as<cpx<T>>();
          AxisIO az{};
          az.s_b = az.d_b = 1;
          az.s_p = az.d_p = KXP * L1;
          az.s_o1 = az.d_o1 = KXP;
          az.nb = d.KX;
          az.n_in = L2;
          az.w0 = 0;
          az.n_out = L2;
          az.mh = tm.zmh;
          az.ms = tm.zms;
          launch_axis<T>(L2, -1, tm.g, full, az, L1, 1, twiddles<T>(c, L2), s);
          it = done.emplace(tm.g, full).first;
        }
        tm.g = it->second;
        tm.zmh = tm.zms = 0;
      }
    zi.zreal = 0;
  }
  {
    ProfScope ps(c, ST_INV_ZMUL);
    launch_zinv<T>(L2, sp, tp, zi, L1, nout, twiddles<T>(c, L2), s);
  }
  AxisIO ai{};
  ai.s_b = ai.d_b = 1;
  ai.s_p = ai.d_p = KXP;
  ai.s_o1 = ai.d_o1 = KXP * L1;
  ai.s_o2 = ai.d_o2 = so;
  ai.nb = d.KX;
  ai.n_in = L1;
  ai.w0 = win[1];
  ai.n_out = ny;
  {
    ProfScope ps(c, ST_INV_Y);
    launch_axis<T>(L1, 1, tp, tp, ai, nz, nout, twiddles<T>(c, L1), s);
  }
  C2RIO<T> cr{};
  cr.in_sy = KXP;
  cr.in_sz = KXP * L1;
  cr.in_so = so;
  cr.out_sy = nx;
  cr.out_sz = int64_t(nx) * ny;
  cr.nrows = int64_t(ny) * nz;
  cr.ny_rows = ny;
  cr.w0 = win[0];
  cr.n_out = nx;
  for (int o = 0; o < nout; ++o) cr.ep[o] = outs[o].ep;
  {
    ProfScope ps(c, ST_INV_X);
    launch_xc2r<T>(L0, tp, cr, nout, twiddles<T>(c, L0), s);
  }
}
template void conv_solve<double>(osc_ctx*, const double*, const int[3], const SpecDims&,
                                 const int[3], const std::vector<OutSpec>&);
template void conv_solve<float>(osc_ctx*, const double*, const int[3], const SpecDims&,
                                const int[3], const std::vector<OutSpec>&);

}  // namespace osc
