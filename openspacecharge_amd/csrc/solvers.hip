// Solver assemblies: each upstream solver as a list of spectral outputs for
// conv_solve, plus the finite-difference / magnetic-field epilogues.
#include "osc_internal.h"
#include "solvers.h"

namespace osc {

namespace {
constexpr double kFpeiSolver2 = 8987551502.2918644;               // open_spacecharge_mod.f90:547
constexpr double kFpeiConv3d = 299792458.0 * 299792458.0 * 1e-7;  // core:515,1250
constexpr double kFpeiPipe = 8987551744.0;                        // core:783
constexpr double kClightSp = 299792448.0;                         // open_spacecharge_mod.f90:436
constexpr double kClightDp = 299792458.0;                         // core:51

OutSpec out1(GreenRef g, double s1, double* dst1) {
  OutSpec o;
  o.nt = 1;
  o.t[0] = {g, 1.0, 0};
  std::memset(&o.ep, 0, sizeof(o.ep));
  o.ep.dst1 = dst1;
  o.ep.s1 = s1;
  o.ep.c = 1.0;
  o.ep.beta = 1.0;
  return o;
}
OutSpec out2(GreenRef g, GreenRef h, double sh, int mir, double s1, double* dst1) {
  OutSpec o = out1(g, s1, dst1);
  o.nt = 2;
  o.t[1] = {h, sh, mir};
  return o;
}
void set_b(OutSpec& o, double* dst2, int bmode, double s2, double beta, double c) {
  o.ep.dst2 = dst2;
  o.ep.bmode = bmode;
  o.ep.s2 = s2;
  o.ep.beta = beta;
  o.ep.c = c;
}

GreenSpec offset_spec(const int L[3], const double delta[3], double gamma, int icomp,
                      const double off[3], int shz) {
  GreenSpec s = make_green_spec();
  s.kind = GK_OFFSET;
  s.icomp = icomp;
  for (int a = 0; a < 3; ++a) {
    s.L[a] = L[a];
    s.delta[a] = delta[a];
    s.offset[a] = off[a];
  }
  s.gamma = gamma;
  s.out_shift_z = shz;
  s.zero_raw = 1;
  return s;
}

// legacy period and index mapping (osc_alloc_freespace_array core:151-182,
// shift core:471)
GreenSpec legacy_spec(const int nlo[3], const int nhi[3], const int npad[3], const double delta[3],
                      double gamma, int icomp, int igf) {
  GreenSpec s = make_green_spec();
  s.kind = GK_LEGACY;
  s.icomp = icomp;
  s.igf = igf;
  for (int a = 0; a < 3; ++a) {
    s.lo[a] = nlo[a] - nhi[a];
    const int hi = nhi[a] - nlo[a] + npad[a];
    s.L[a] = hi - s.lo[a] + 1;
    s.shift[a] = s.L[a] / 2 - (npad[a] + 1) / 2;
    s.delta[a] = delta[a];
  }
  s.gamma = gamma;
  return s;
}

// the Green spectra of one solve: cached ones stay pinned in the context's
// cache until the solve's kernels are queued (the end of its scope)
template <typename T>
struct Specs {
  std::vector<DevBuf> keep;
  osc_ctx* pinned = nullptr;
  explicit Specs(size_t n) : keep(n) {}
  ~Specs() {
    if (pinned) pinned->gcache.unpin();
  }
  Specs(const Specs&) = delete;
  Specs& operator=(const Specs&) = delete;
  GreenRef get(osc_ctx* c, const GreenSpec& s, size_t slot) {
    if (!pinned) {
      pinned = c;
      c->gcache.pin();
    }
    return green_spectrum<T>(c, s, keep[slot]);
  }
};

void mesh_n(const osc_mesh& m, int n[3]) {
  for (int a = 0; a < 3; ++a) n[a] = m.nhi[a] - m.nlo[a] + 1;
}

// ---- osc_freespace_solver2 ------------------------------------------------------
template <typename T>
void solver2_t(osc_ctx* c, const int n[3], const double* rho, double gamma, const double delta[3],
               double* ef, double* phi, const double* offset) {
  int L[3], win[3];
  for (int a = 0; a < 3; ++a) {
    L[a] = 2 * n[a];
    win[a] = n[a] - 1;
  }
  const SpecDims d = SpecDims::of(L);
  const double off[3] = {offset ? offset[0] : 0.0, offset ? offset[1] : 0.0,
                         offset ? offset[2] : 0.0};
  const double s1 = kFpeiSolver2 / double(L[0] * L[1] * L[2]);  // fpei/(nx2*ny2*nz2) :587
  const int64_t nt = int64_t(n[0]) * n[1] * n[2];
  Specs<T> sp(4);
  std::vector<OutSpec> outs;
  if (phi) outs.push_back(out1(sp.get(c, offset_spec(L, delta, gamma, 0, off, 0), 0), s1, phi));
  if (ef)
    for (int ic = 1; ic <= 3; ++ic)
      outs.push_back(out1(sp.get(c, offset_spec(L, delta, gamma, ic, off, 0), ic), s1,
                          ef + (ic - 1) * nt));
  if (!outs.empty()) conv_solve<T>(c, rho, n, d, win, outs);
}

// ---- space_charge_3d --------------------------------------------------------------
template <typename T>
void sc3d_t(osc_ctx* c, const osc_mesh& m, const double* rho, const double* offset,
            int at_cathode, int calc_b, double* ef, double* bf) {
  int n[3], L[3], win[3];
  mesh_n(m, n);
  for (int a = 0; a < 3; ++a) {
    L[a] = 2 * n[a];
    win[a] = n[a] - 1;
  }
  const SpecDims d = SpecDims::of(L);
  const int64_t nt = int64_t(n[0]) * n[1] * n[2];
  double off[3] = {offset ? offset[0] : 0.0, offset ? offset[1] : 0.0, offset ? offset[2] : 0.0};
  const double s1 = kFpeiSolver2 / double(L[0] * L[1] * L[2]);
  double bc = 0.0;
  if (calc_b) {
    const double beta = sqrt(1 - 1 / (m.gamma * m.gamma));  // :457
    bc = beta / kClightSp;
  }
  Specs<T> sp(6);
  GreenRef G[3];
  for (int ic = 1; ic <= 3; ++ic)
    G[ic - 1] = sp.get(c, offset_spec(L, m.delta, m.gamma, ic, off, 0), ic - 1);
  std::vector<OutSpec> outs;
  if (!at_cathode) {
    for (int ic = 0; ic < 3; ++ic) outs.push_back(out1(G[ic], s1, ef + ic * nt));
    if (calc_b) {
      set_b(outs[0], bf + nt, 0, bc, 1.0, 1.0);   // By =  (beta/c) Ex  (:460)
      set_b(outs[1], bf, 0, -bc, 1.0, 1.0);       // Bx = -(beta/c) Ey  (:459)
    }
  } else {
    // image solve of -rho flipped in z with the offset of :472
    double offi[3] = {off[0], off[1], off[2]};
    offi[2] = offi[2] + 2 * m.min[2] + (m.max[2] - m.min[2]);
    GreenRef H[3];
    for (int ic = 1; ic <= 3; ++ic)
      H[ic - 1] = sp.get(c, offset_spec(L, m.delta, m.gamma, ic, offi, n[2] - 1), 3 + ic - 1);
    for (int ic = 0; ic < 3; ++ic) outs.push_back(out2(G[ic], H[ic], -1.0, 4, s1, ef + ic * nt));
    if (calc_b) {
      // Bx = -(beta/c)(Ey_free - Ey_img), By = (beta/c)(Ex_free - Ex_img) (:459-460, 483-484)
      OutSpec bx = out2(G[1], H[1], 1.0, 4, s1, nullptr);
      set_b(bx, bf, 0, -bc, 1.0, 1.0);
      OutSpec by = out2(G[0], H[0], 1.0, 4, s1, nullptr);
      set_b(by, bf + nt, 0, bc, 1.0, 1.0);
      outs.push_back(bx);
      outs.push_back(by);
    }
  }
  conv_solve<T>(c, rho, n, d, win, outs);
  if (calc_b) fill_f64(c, bf + 2 * nt, nt, 0.0);  // bfield = 0 (:458) for Bz
}

// ---- osc_freespace_solver (legacy) --------------------------------------------------
template <typename T>
int legacy_free_t(osc_ctx* c, const osc_mesh& m, const double* rho, int direct, int igf,
                  double* phi, double* ef, double* bf) {
  if (igf == 0 && direct == 1) return OSC_ERR_DIRECT_NEEDS_IGF;  // core:58-62
  int n[3];
  mesh_n(m, n);
  const int64_t nt = int64_t(n[0]) * n[1] * n[2];
  GreenSpec s0 = legacy_spec(m.nlo, m.nhi, m.npad, m.delta, m.gamma, 0, igf);
  const SpecDims d = SpecDims::of(s0.L);
  const int win[3] = {0, 0, 0};
  // factr = 1/((1*iperiod)*(1*jperiod)*(1*kperiod))*fpei*qtot (core:522)
  const double s1 =
      1.0 / ((1.0 * s0.L[0]) * (1.0 * s0.L[1]) * (1.0 * s0.L[2])) * kFpeiConv3d * 1.0;
  const double beta0 = sqrt(1.0 - 1.0 / (m.gamma * m.gamma));
  Specs<T> sp(3);
  std::vector<OutSpec> outs;
  if (direct == 0) {
    outs.push_back(out1(sp.get(c, s0, 0), s1, phi));
    conv_solve<T>(c, rho, n, d, win, outs);
    fd_gradient(c, phi, n, m.delta, m.gamma, ef);
    bfield_core(c, ef, nt, m.gamma, bf);
  } else {
    for (int ic = 1; ic <= 3; ++ic) {
      GreenSpec s = legacy_spec(m.nlo, m.nhi, m.npad, m.delta, m.gamma, ic, igf);
      outs.push_back(out1(sp.get(c, s, ic - 1), s1, ef + (ic - 1) * nt));
    }
    set_b(outs[0], bf + nt, 1, 1.0, beta0, kClightDp);   // By =  Ex*beta0/c
    set_b(outs[1], bf, 1, -1.0, beta0, kClightDp);       // Bx = -Ey*beta0/c
    conv_solve<T>(c, rho, n, d, win, outs);
    fill_f64(c, bf + 2 * nt, nt, 0.0);
  }
  return OSC_OK;
}

// ---- osc_cathodeimages_solver ----------------------------------------------------------
template <typename T>
int cathode_t(osc_ctx* c, const osc_mesh& m, double* rho, int direct, int igf, int method,
              double* phi, double* ef, double* bf) {
  if (igf == 0 && direct == 1) return OSC_ERR_DIRECT_NEEDS_IGF;  // core:935-939
  int n[3];
  mesh_n(m, n);
  const int64_t nt = int64_t(n[0]) * n[1] * n[2];
  GreenSpec base = legacy_spec(m.nlo, m.nhi, m.npad, m.delta, m.gamma, 0, igf);
  const SpecDims d = SpecDims::of(base.L);
  const int win[3] = {0, 0, 0};
  const double s1 =
      1.0 / ((1.0 * base.L[0]) * (1.0 * base.L[1]) * (1.0 * base.L[2])) * kFpeiConv3d * 1.0;
  const double beta0 = sqrt(1.0 - 1.0 / (m.gamma * m.gamma));
  const double umax3 = m.min[2] + (m.nhi[2] - m.nlo[2]) * m.delta[2];  // core:1059
  const double zshft = m.min[2] + umax3;
  auto image_spec = [&](int ic) {
    GreenSpec s = legacy_spec(m.nlo, m.nhi, m.npad, m.delta, m.gamma, ic, igf);
    if (method == 1) {
      // osc_alloc_image_array z range 2*nlo .. 2*nhi+npad (core:1104-1105)
      s.kind = GK_IMG_CONVCORR;
      s.lo[2] = 2 * m.nlo[2];
      s.nlo3 = m.nlo[2];
      s.umin[2] = m.min[2];
    } else {
      // conv(flip_z(rho), g) = conv(reverse(rho), g shifted by nz-1)
      s.kind = GK_IMG_SHIFT;
      s.zshft = zshft;
      s.out_shift_z = n[2] - 1;
    }
    return s;
  };
  Specs<T> sp(8);
  std::vector<OutSpec> outs;
  if (direct == 0) {
    const GreenRef g = sp.get(c, legacy_spec(m.nlo, m.nhi, m.npad, m.delta, m.gamma, 0, igf), 0);
    const GreenRef h = sp.get(c, image_spec(0), 1);
    DevBuf phs(c->pool, sizeof(double) * nt), es(c->pool, sizeof(double) * 3 * nt);
    outs.push_back(out2(g, h, -1.0, 4, s1, phi));                 // phi = phi_free - phi_img
    outs.push_back(out2(g, h, 1.0, 4, s1, phs.as<double>()));     // phi_free + phi_img
    conv_solve<T>(c, rho, n, d, win, outs);
    fd_gradient(c, phi, n, m.delta, m.gamma, ef);
    fd_gradient(c, phs.as<double>(), n, m.delta, m.gamma, es.as<double>());
    // B from E + 2 E_img = grad(phi_free + phi_img) (core:1076-1077)
    bfield_core(c, es.as<double>(), nt, m.gamma, bf);
  } else {
    GreenRef G[3];
    GreenRef H[3];
    for (int ic = 1; ic <= 3; ++ic) {
      G[ic - 1] = sp.get(c, legacy_spec(m.nlo, m.nhi, m.npad, m.delta, m.gamma, ic, igf), ic - 1);
      H[ic - 1] = sp.get(c, image_spec(ic), 3 + ic - 1);
    }
    for (int ic = 0; ic < 3; ++ic) outs.push_back(out2(G[ic], H[ic], -1.0, 4, s1, ef + ic * nt));
    OutSpec bx = out2(G[1], H[1], 1.0, 4, s1, nullptr);
    set_b(bx, bf, 1, -1.0, beta0, kClightDp);  // -(Ey + 2 Ey_img)*beta0/c
    OutSpec by = out2(G[0], H[0], 1.0, 4, s1, nullptr);
    set_b(by, bf + nt, 1, 1.0, beta0, kClightDp);
    outs.push_back(bx);
    outs.push_back(by);
    conv_solve<T>(c, rho, n, d, win, outs);
    fill_f64(c, bf + 2 * nt, nt, 0.0);
  }
  if (method == 2) flip_z(c, rho, n);  // upstream leaves rho flipped (core:1047-1057)
  return OSC_OK;
}

// ---- osc_rectpipe_solver ------------------------------------------------------------------
template <typename T>
int pipe_t(osc_ctx* c, const osc_mesh& m, int direct, const double* rho, double* phi, double* ef,
           double* bf) {
  int n[3];
  mesh_n(m, n);
  const int64_t nt = int64_t(n[0]) * n[1] * n[2];
  PipeState& p = c->pipe;
  OSC_REQUIRE(p.built, OSC_ERR_NO_GREEN,
              "rectpipe green functions should have been computed during initialization");
  for (int a = 0; a < 3; ++a)
    OSC_REQUIRE(p.nhi[a] - p.nlo[a] + 1 == n[a], OSC_ERR_ARG,
                "mesh size differs from the pipe Green-function grid");
  if (direct == 1) {
    bfield_core(c, ef, nt, m.gamma, bf);  // E left unset upstream (core:608-622)
    return OSC_ERR_PIPE_DIRECT;
  }
  const SpecDims d = SpecDims::of(p.L);
  GreenRef g[4];  // full half spectra
  for (int t = 0; t < 4; ++t) {
    if constexpr (sizeof(T) == 4) {
      if (!p.spec32[t].p) {
        p.spec32[t] = DevBuf(c->pool, sizeof(cpx<float>) * d.size());
        convert_spectrum(c, p.spec64[t].as<cpx<double>>(), p.spec32[t].as<cpx<float>>(),
                         d.size());
      }
      g[t].g = p.spec32[t].p;
    } else {
      g[t].g = p.spec64[t].p;
    }
  }
  // ccon = t1 - t2 - t3 + t4 with correlations in y, x, xy (core:790-806)
  OutSpec o = out1(g[0], 1.0 / ((1.0 * p.L[0]) * (1.0 * p.L[1]) * (1.0 * p.L[2])) * kFpeiPipe * 1.0,
                   phi);
  o.nt = 4;
  o.t[1] = {g[1], -1.0, 2};
  o.t[2] = {g[2], -1.0, 1};
  o.t[3] = {g[3], 1.0, 3};
  const int win[3] = {0, 0, 0};
  conv_solve<T>(c, rho, n, d, win, {o});
  fd_gradient(c, phi, n, m.delta, m.gamma, ef);
  bfield_core(c, ef, nt, m.gamma, bf);
  return OSC_OK;
}
}  // namespace

void solver2(osc_ctx* c, const int n[3], const double* rho, double gamma, const double delta[3],
             double* ef, double* phi, const double* offset) {
  if (c->prec == OSC_PREC_FP32) solver2_t<float>(c, n, rho, gamma, delta, ef, phi, offset);
  else solver2_t<double>(c, n, rho, gamma, delta, ef, phi, offset);
}
void space_charge_3d(osc_ctx* c, const osc_mesh& m, const double* rho, const double* offset,
                     int at_cathode, int calc_b, double* ef, double* bf) {
  if (c->prec == OSC_PREC_FP32) sc3d_t<float>(c, m, rho, offset, at_cathode, calc_b, ef, bf);
  else sc3d_t<double>(c, m, rho, offset, at_cathode, calc_b, ef, bf);
}
// Solves that take E as the finite-difference gradient of phi (legacy with
// direct_field_calc = 0, the rectangular pipe) run the convolution in FP64
// in either precision mode: the differences of neighbouring phi values
// cancel, amplifying phi's relative rounding by about the bunch size in
// cells, which in FP32 exceeds the 1e-5 field tolerance.
int legacy_free(osc_ctx* c, const osc_mesh& m, const double* rho, int direct, int igf,
                double* phi, double* ef, double* bf) {
  if (c->prec == OSC_PREC_FP32 && direct) return legacy_free_t<float>(c, m, rho, direct, igf, phi, ef, bf);
  return legacy_free_t<double>(c, m, rho, direct, igf, phi, ef, bf);
}
int cathode_legacy(osc_ctx* c, const osc_mesh& m, double* rho, int direct, int igf, int method,
                   double* phi, double* ef, double* bf) {
  if (c->prec == OSC_PREC_FP32 && direct)
    return cathode_t<float>(c, m, rho, direct, igf, method, phi, ef, bf);
  return cathode_t<double>(c, m, rho, direct, igf, method, phi, ef, bf);
}
int rectpipe(osc_ctx* c, const osc_mesh& m, int direct, const double* rho, double* phi, double* ef,
             double* bf) {
  return pipe_t<double>(c, m, direct, rho, phi, ef, bf);
}

// osc_getgrnpipe (core:632-715): the half spectra of the four real arrays on
// the pipe grid, built from their separable form (pipe_spectrum; upstream
// signs are applied only on file I/O)
void pipe_build(osc_ctx* c, double gamma, double a, double b, const double delta[3],
                const double umin[3], const int npad[3]) {
  PipeState& p = c->pipe;
  OSC_REQUIRE(p.alloc, OSC_ERR_NO_GREEN, "osc_alloc_rectpipe_arrays has not been called");
  GreenSpec s = make_green_spec();
  s.kind = GK_PIPE;
  for (int d = 0; d < 3; ++d) {
    s.lo[d] = p.nlo[d] - p.nhi[d];
    s.L[d] = p.L[d];
    s.shift[d] = p.L[d] / 2 - (npad[d] + 1) / 2;
    s.delta[d] = delta[d];
    s.umin[d] = umin[d];
  }
  s.gamma = gamma;
  s.a = a;
  s.b = b;
  const SpecDims d = SpecDims::of(p.L);
  for (int t = 0; t < 4; ++t) {
    s.term = t;
    if (!p.spec64[t].p) p.spec64[t] = DevBuf(c->pool, sizeof(cpx<double>) * d.size());
    pipe_spectrum(c, s, d, p.spec64[t].as<cpx<double>>());
    p.spec32[t].reset();
  }
  p.built = true;
  p.a = a;
  p.b = b;
  p.gamma = gamma;
  for (int k = 0; k < 3; ++k) {
    p.delta[k] = delta[k];
    p.umin[k] = umin[k];
  }
}

// upstream FFT sign patterns of the pipe arrays (core:666,682,698,713)
const int kPipeSigns[4][3] = {{1, 1, 1}, {1, -1, 1}, {-1, 1, 1}, {-1, -1, 1}};

}  // namespace osc
