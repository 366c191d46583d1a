/*
 * osc_oracle -- CPU restatement (plain C, FP64) of the upstream OpenSC
 * space-charge path, used ONLY as test infrastructure: the parity checker for
 * the HIP library and a CPU baseline.  Nothing on the product path links it.
 *
 * Every routine follows the upstream routine it names (file:line in the .c
 * file), including its floating-point operation order, so that it can be
 * compared bit-for-bit with the compiled upstream library.
 *
 * Conventions: arrays are Fortran column-major (x fastest, lower bound 1 in
 * the upstream code, 0 here); complex arrays are interleaved (re, im) doubles;
 * field arrays are (nx, ny, nz, 3) with the component slowest.
 */
#ifndef OSC_ORACLE_H
#define OSC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* physical constants exactly as the upstream code folds them */
extern const double oc_fpei_solver2;   /* open_spacecharge_mod.f90:547 */
extern const double oc_fpei_conv3d;    /* open_spacecharge_core_mod.f90:515,1250 */
extern const double oc_fpei_pipe;      /* open_spacecharge_core_mod.f90:783 */
extern const double oc_clight_sp;      /* open_spacecharge_mod.f90:436,546 */
extern const double oc_clight_dp;      /* open_spacecharge_core_mod.f90:51,555,929 */

/* ---- test knobs (defaults restate upstream exactly) ------------------------ */
/* perturb every IGF primitive result by a pseudo-random -1/0/+1 ulp (0 = off) */
void oc_set_perturb(uint64_t seed);
/* zero the raw, never-used last planes of the offset-grid Green function */
void oc_set_zero_unused_planes(int on);

/* ---- 3-D complex FFT (fft_interface_mod.f90:96-143) ---------------------- */
/* in place, unnormalised, axis d uses exp(idir[d] * 2*pi*i*j*k/n).           */
/* returns 0, or -1 if a length has a prime factor > 23 (upstream stops).    */
int oc_ccfft3d(double *a, int n1, int n2, int n3, const int idir[3]);

/* ---- deposition ----------------------------------------------------------- */
/* deposit_particles (open_spacecharge_mod.f90:156-288).
 * q: per-particle charges or NULL; total_charge: used when q == NULL
 * (has_total != 0).  resize != 0 recomputes min/max/delta from the bunch.
 * Particles whose 8 nodes are not all on the mesh are skipped and counted in
 * *n_dropped (upstream has no check and writes out of bounds).
 * Returns 0, or -1 when neither q nor total_charge is given (upstream prints
 * an error and returns). */
int oc_deposit_particles(int64_t n, const double *x, const double *y, const double *z,
                         const double *q, int has_total, double total_charge, int resize,
                         const int nlo[3], const int nhi[3], double mn[3], double mx[3],
                         double delta[3], double *charge, double *rho, int64_t *n_dropped);

/* get_mesh_quantities (test_mod.f90:288-358) */
void oc_get_mesh_quantities(int64_t n, const double *x, const double *y, const double *z,
                            const int nlo[3], const int nhi[3], int isetzmin,
                            double delta[3], double umin[3]);
/* depose_rho_scalar (test_mod.f90:410-474); returns ifail (dropped count) */
int64_t oc_depose_rho_scalar(int64_t n, const double *x, const double *y, const double *z,
                             const int nlo[3], const int nhi[3], double chrgpermacro,
                             const double delta[3], const double umin[3], double *rho);

/* ---- Green functions ------------------------------------------------------ */
/* osc_get_cgrn_freespace (open_spacecharge_mod.f90:633-707): complex array
 * (isize,jsize,ksize), real-space (not transformed). offset may be NULL. */
void oc_get_cgrn_freespace(double *cgrn, int isize, int jsize, int ksize,
                           const double delta[3], double gamma, int icomp,
                           const double *offset);
/* primitive functions (for tests) */
double oc_xlafun2(double x, double y, double z);
double oc_lafun2(double x, double y, double z);
double oc_xlafun(double x, double y, double z);
double oc_ylafun(double x, double y, double z);
double oc_zlafun(double x, double y, double z);
double oc_lafun(double x, double y, double z);
double oc_igf(int igfflag, int icomp, double u, double v, double w, double gam,
              double dx, double dy, double dz);
double oc_rfun(double u, double v, double w, double gam, double a, double b, double hz,
               int i, int j);

/* osc_getgrnfree (core:454-497): transformed legacy Green function, array of
 * size (2n-1+npad) per axis.  shift variant = osc_getgrnimageshift
 * (core:1135-1178); convcorr = osc_getgrnimageconvcorr (core:1185-1230). */
void oc_getgrnfree(double *cgrn, const int nlo[3], const int nhi[3], const int npad[3],
                   const double delta[3], double gamma, int icomp, int igfflag);
void oc_getgrnimageshift(double *cgrn, const int nlo[3], const int nhi[3], const int npad[3],
                         const double delta[3], double gamma, double zshft, int icomp,
                         int igfflag);
void oc_getgrnimageconvcorr(double *cgrn, const int nlo[3], const int nhi[3],
                            const int npad[3], const double umin[3], const double delta[3],
                            double gamma, int icomp, int igfflag);
/* osc_getgrnpipe (core:632-715): four transformed arrays g[0..3] */
void oc_getgrnpipe(double *g[4], const int nlo[3], const int nhi[3], const int npad[3],
                   double gamma, double a, double b, const double delta[3],
                   const double umin[3]);

/* ---- solvers -------------------------------------------------------------- */
/* osc_freespace_solver2 (open_spacecharge_mod.f90:535-597); efield/phi/offset
 * may be NULL. rho is (nx,ny,nz). */
void oc_freespace_solver2(int nx, int ny, int nz, const double *rho, double gamma,
                          const double delta[3], double *efield, double *phi,
                          const double *offset);
/* space_charge_3d (open_spacecharge_mod.f90:431-489) */
void oc_space_charge_3d(const int nlo[3], const int nhi[3], const double mn[3],
                        const double mx[3], const double delta[3], double gamma,
                        const double *rho, const double *offset, int at_cathode,
                        int calc_bfield, double *efield, double *bfield);
/* space_charge_freespace -> osc_freespace_solver (core:30-138).
 * Returns -1 (outputs untouched) for direct=1 with igf=0, as upstream. */
int oc_space_charge_freespace(const int nlo[3], const int nhi[3], const int npad[3],
                              const double delta[3], double gamma, const double *rho,
                              int direct, int igf, double *phi, double *efield,
                              double *bfield);
/* space_charge_cathodeimages -> osc_cathodeimages_solver (core:905-1083).
 * method 2 flips rho in z in place, as upstream. */
int oc_space_charge_cathodeimages(const int nlo[3], const int nhi[3], const int npad[3],
                                  const double umin[3], const double delta[3], double gamma,
                                  double *rho, int direct, int igf, int method, double *phi,
                                  double *efield, double *bfield);
/* space_charge_rectpipe -> osc_rectpipe_solver (core:536-624) with transformed
 * Green arrays g[0..3] from oc_getgrnpipe. direct=1: E untouched, B from E. */
int oc_space_charge_rectpipe(const int nlo[3], const int nhi[3], const int npad[3],
                             const double delta[3], double gamma, const double *rho,
                             double *const g[4], int direct, double *phi, double *efield,
                             double *bfield);

/* ---- gather --------------------------------------------------------------- */
/* interpolate_field (open_spacecharge_mod.f90:311-390) for n points.
 * E and B are (3, n) and may be NULL; clamped (optional, may be NULL)
 * receives 1 for points that were clamped onto the mesh. */
void oc_interpolate_field(const int nlo[3], const int nhi[3], const double mn[3],
                          const double delta[3], const double *efield, const double *bfield,
                          int64_t n, const double *x, const double *y, const double *z,
                          double *E, double *B, int8_t *clamped);

/* ---- pipe Green file (core:825-870) --------------------------------------- */
int oc_write_rectpipe_grn(const char *path, double a, double b, const double delta[3],
                          const double umin[3], const double umax[3], const int nlo[3],
                          const int nhi[3], const int period[3], double gamma,
                          double *const g[4]);
int oc_read_rectpipe_grn(const char *path, double hdr_real[12], int hdr_int[9],
                         double *g[4], int64_t n_each);

#ifdef __cplusplus
}
#endif
#endif
