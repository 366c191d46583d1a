/*
 * opensc_hip.h -- C-ABI of the MI355X (gfx950) OpenSC space-charge library.
 *
 * Every entry point replaces one routine of the upstream Fortran module
 * open_spacecharge_mod (code/open_spacecharge_mod.f90) or of the module it
 * re-exports, open_spacecharge_core_mod (code/open_spacecharge_core_mod.f90);
 * the routine is cited beside each declaration.  The Fortran drop-in module
 * (openspacecharge_amd/fortran/open_spacecharge_mod.f90) binds these through
 * ISO_C_BINDING; Python binds them through ctypes.
 *
 * Conventions
 *  - Arrays are Fortran column-major with x fastest: a grid is (nx,ny,nz),
 *    a field is (nx,ny,nz,3) with the component slowest, complex arrays are
 *    interleaved (re,im) doubles, particle E/B outputs are (3,n).
 *  - Any array may live in host memory or in device memory (HBM) of the
 *    context's GPU; the library detects which.  Host arrays are staged through
 *    device buffers and the call returns when results are back.  When every
 *    array of a call is device memory and osc_ctx_set_async(ctx, 1) was set,
 *    the call only enqueues work on the context stream.
 *  - Every function returns OSC_OK or an OSC_ERR_* code; osc_ctx_last_error
 *    gives the message.  Upstream error behaviour (print + return / stop) is
 *    reproduced by the Fortran module from these codes.
 *  - Results are deterministic: no floating-point atomics anywhere.
 *  - A context is not thread-safe; use one context per host thread.
 */
#ifndef OPENSC_HIP_H
#define OPENSC_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSC_ABI_VERSION 1

enum osc_status {
  OSC_OK = 0,
  OSC_ERR_ARG = 1,              /* invalid argument (null pointer, bad sizes)              */
  OSC_ERR_NO_CHARGE = 2,        /* neither qa nor total_charge (open_spacecharge_mod.f90:249-252) */
  OSC_ERR_FFT_LENGTH = 3,       /* transform length unsupported (fft_interface_mod.f90:158-162)  */
  OSC_ERR_HIP = 4,              /* HIP runtime error                                       */
  OSC_ERR_NOMEM = 5,            /* device memory exhausted                                 */
  OSC_ERR_DIRECT_NEEDS_IGF = 6, /* direct field calc with igf off (core:58-62, 935-939)    */
  OSC_ERR_PIPE_DIRECT = 7,      /* rect pipe needs direct_field_calc=.false. (core:608-610) */
  OSC_ERR_NO_GREEN = 8,         /* pipe Green functions not built or read (core:578-581)   */
  OSC_ERR_IO = 9,               /* Green-function file I/O (core:825-870)                  */
  OSC_ERR_NO_DEVICE = 10        /* no usable GPU                                           */
};

enum osc_precision {
  OSC_PREC_FP64 = 0, /* drop-in: FP64 convolution (upstream arithmetic)                 */
  OSC_PREC_FP32 = 1  /* FP32 convolution against the FP64-built Green spectrum          */
};

/* mesh3d_struct (open_spacecharge_mod.f90:14-27) without its allocatable arrays */
typedef struct osc_mesh {
  int32_t nlo[3];  /* lowest grid index (upstream default 1)  */
  int32_t nhi[3];  /* highest grid index (upstream default 64) */
  int32_t npad[3]; /* padding for the cyclic convolution (default 1) */
  double min[3];
  double max[3];
  double delta[3];
  double gamma;
  double charge;
} osc_mesh;

typedef struct osc_ctx osc_ctx;

/* ---- context ---------------------------------------------------------------
 * Holds the device, the stream, pooled work buffers, the twiddle tables and
 * the Green-function caches.  It also holds the pipe Green functions that the
 * upstream code keeps in module globals (core:22-23). */
int osc_ctx_create(int device, osc_ctx** ctx);
int osc_ctx_destroy(osc_ctx* ctx);
/* use an existing HIP stream (hipStream_t); NULL restores the context's own */
int osc_ctx_set_stream(osc_ctx* ctx, void* hip_stream);
int osc_ctx_set_async(osc_ctx* ctx, int async);
int osc_ctx_set_precision(osc_ctx* ctx, int precision);
/* Green-spectrum cache: enable (default 1) and byte budget (default 8 GiB).
 * A spectrum is kept from the second solve that misses on its mesh, spacing
 * and solver on; a first-seen one is built into a pooled buffer that later
 * solves reuse (no device allocation per new bunch shape). */
int osc_ctx_set_green_cache(osc_ctx* ctx, int enable, int64_t max_bytes);
/* Fused steps (osc_space_charge_step) with a resized (bunch-following) mesh
 * of at most 128^3 nodes and no offset record their grid part -- Green
 * build, spectral solve, B -- as a HIP graph after the first step of each
 * (shape, precision, solver, B, gamma), and replay it for later steps of
 * that shape whatever their mesh: the Green primitive reads the spacing from
 * the device, so the solve is queued without waiting for the mesh.  Results
 * are bit-identical to the eager launches.  Fixed meshes run eagerly and
 * reuse cached Green spectra.  Default off (measured on MI355X: replay is
 * 2-7 % slower than the eager launches from 32^3 to 128^3, whose host work
 * already overlaps the device); off clears the graphs.  Profiling
 * (osc_ctx_set_profiling) runs the eager launches. */
int osc_ctx_set_graphs(osc_ctx* ctx, int enable);
/* graphs held, graphs captured, replays since creation */
int osc_ctx_graph_stats(osc_ctx* ctx, int64_t* out3);
int osc_ctx_clear_caches(osc_ctx* ctx);
int osc_ctx_synchronize(osc_ctx* ctx);
const char* osc_ctx_last_error(const osc_ctx* ctx);
int osc_abi_version(void);

/* Per-stage timing with HIP events on the context stream (no counterpart
 * upstream, which only times the whole solve with cpu_time,
 * test_opensc.f90:175,219-220).  While enabled, every kernel launch is
 * bracketed by events of its stage; osc_ctx_profile_read synchronises, adds
 * the elapsed times (ms) and launch counts of each stage into the arrays
 * (OSC_NSTAGE entries) and clears the records. */
#define OSC_NSTAGE 20
int osc_ctx_set_profiling(osc_ctx* ctx, int on);
int osc_ctx_profile_read(osc_ctx* ctx, double* ms, int64_t* launches);
const char* osc_stage_name(int stage);
/* 1 if this transform length is supported (prime factors <= 23 and fits LDS) */
int osc_fft_length_supported(int n);
/* 1 if this length has a compile-time radix plan (LDS-tiled multi-line passes);
 * other supported lengths use the runtime-radix engine, one line per workgroup */
int osc_fft_length_planned(int n);
/* Tile width (lines per workgroup) of the strided FFT passes, chosen per
 * (precision, length, pass kind, line count) by an online tuner: the first
 * calls of each run the candidate widths (half, default, double) in turn,
 * timed by events, then the fastest is kept; results do not depend on the
 * width.  mode 1: tune (default); 0: default width; 2, 3, 4: force the half,
 * default or double width.  OSC_FFT_TUNE_FILE (environment) keeps the
 * decisions across processes.  Process-wide. */
int osc_fft_tuning(int mode);
/* lengths/kinds decided, still measuring */
int osc_fft_tune_stats(int64_t* out2);

/* ---- deposition --------------------------------------------------------------
 * deposit_particles (open_spacecharge_mod.f90:156-288).
 * qa (n charges) or *total_charge must be given (qa wins, as upstream).
 * resize_mesh != 0 sets mesh->min/max/delta from the bunch (the upstream
 * default).  mesh->charge is set.  rho is (nx,ny,nz).  Particles whose 8 CIC
 * nodes are not all on the mesh are skipped and counted in *n_dropped
 * (upstream has no check there and writes out of bounds).  Cell indices are
 * computed exactly as upstream (FP64, no contraction); the accumulation is a
 * deterministic cell sort + segmented reduction. */
int osc_deposit_particles(osc_ctx* ctx, int64_t n, const double* xa, const double* ya,
                          const double* za, const double* qa, const double* total_charge,
                          int resize_mesh, osc_mesh* mesh, double* rho, int64_t* n_dropped);

/* One tracking step: deposit -> space-charge solve -> gather at the same
 * particles.  No single upstream routine: it replaces the caller's sequence
 * deposit_particles (open_spacecharge_mod.f90:156-288), space_charge_3d
 * (:431-489) and interpolate_field per particle (:311-390), and gives the
 * same results as osc_deposit_particles + osc_space_charge_3d +
 * osc_interpolate_field.  The gather reuses the deposit's cell-sorted
 * particle order.  E and B are (3,n); B requires calc_bfield; mesh arrays
 * rho (nx,ny,nz), efield and bfield (nx,ny,nz,3) are outputs. */
int osc_space_charge_step(osc_ctx* ctx, int64_t n, const double* xa, const double* ya,
                          const double* za, const double* qa, const double* total_charge,
                          int resize_mesh, const double* offset, int at_cathode, int calc_bfield,
                          osc_mesh* mesh, double* rho, double* efield, double* bfield, double* E,
                          double* B, int64_t* n_dropped, int64_t* n_clamped);

/* get_mesh_quantities (code/test_mod.f90:288-358): harness mesh rule with one
 * guard node per side; isetzmin=1 puts zmin at 1e-9 (cathode runs). */
int osc_get_mesh_quantities(osc_ctx* ctx, int64_t n, const double* xa, const double* ya,
                            const double* za, const int32_t nlo[3], const int32_t nhi[3],
                            int isetzmin, double delta[3], double umin[3]);

/* depose_rho_scalar (code/test_mod.f90:410-474): unit weights scaled by
 * chrgpermacro at the end, out-of-range particles dropped and counted. */
int osc_depose_rho_scalar(osc_ctx* ctx, int64_t n, const double* xa, const double* ya,
                          const double* za, const int32_t nlo[3], const int32_t nhi[3],
                          double chrgpermacro, const double delta[3], const double umin[3],
                          double* rho, int64_t* ifail);

/* ---- free-space and cathode solve (IGF, offset grid) -------------------------
 * space_charge_3d (open_spacecharge_mod.f90:431-489).  offset may be NULL
 * (zero).  bfield may be NULL when calc_bfield == 0. */
int osc_space_charge_3d(osc_ctx* ctx, const osc_mesh* mesh, const double* rho,
                        const double* offset, int at_cathode, int calc_bfield, double* efield,
                        double* bfield);

/* osc_freespace_solver2 (open_spacecharge_mod.f90:535-597): efield (nx,ny,nz,3)
 * and/or phi (nx,ny,nz), either may be NULL; offset may be NULL. */
int osc_freespace_solver2(osc_ctx* ctx, int nx, int ny, int nz, const double* rho, double gamma,
                          const double delta[3], double* efield, double* phi,
                          const double* offset);

/* osc_get_cgrn_freespace (open_spacecharge_mod.f90:633-707): the real-space
 * Green array (isize,jsize,ksize) as complex with zero imaginary part. */
int osc_get_cgrn_freespace(osc_ctx* ctx, double* cgrn, int isize, int jsize, int ksize,
                           const double delta[3], double gamma, int icomp, const double* offset);

/* ccfft3d (fft_interface_mod.f90:96-143): complex (n1,n2,n3) transform with
 * exponent sign idir[d] on axis d, unnormalised; a may equal b. */
int osc_ccfft3d(osc_ctx* ctx, const double* a, double* b, const int32_t idir[3], int n1, int n2,
                int n3);

/* ---- legacy solvers (period 2n-1+npad, first-n extraction) -------------------
 * space_charge_freespace (open_spacecharge_mod.f90:60-75) ->
 * osc_freespace_solver (core:30-138). */
int osc_space_charge_freespace(osc_ctx* ctx, const osc_mesh* mesh, const double* rho,
                               int direct_field_calc, int integrated_green_function, double* phi,
                               double* efield, double* bfield);

/* space_charge_cathodeimages (open_spacecharge_mod.f90:83-103) ->
 * osc_cathodeimages_solver (core:905-1083).  image_method 2 flips rho in z in
 * place, as upstream (core:1047-1057). */
int osc_space_charge_cathodeimages(osc_ctx* ctx, const osc_mesh* mesh, double* rho,
                                   int direct_field_calc, int integrated_green_function,
                                   int image_method, double* phi, double* efield, double* bfield);

/* transformed Green functions in the upstream convention (full complex array
 * of the period, forward sign +1 on each axis except where upstream flips it):
 * osc_getgrnfree (core:454-497), osc_getgrnimageshift (core:1135-1178),
 * osc_getgrnimageconvcorr (core:1185-1230). */
int osc_getgrnfree(osc_ctx* ctx, double* cgrn, const int32_t nlo[3], const int32_t nhi[3],
                   const int32_t npad[3], const double delta[3], double gamma, int icomp,
                   int igfflag);
int osc_getgrnimageshift(osc_ctx* ctx, double* cgrn, const int32_t nlo[3], const int32_t nhi[3],
                         const int32_t npad[3], const double delta[3], double gamma, double zshft,
                         int icomp, int igfflag);
int osc_getgrnimageconvcorr(osc_ctx* ctx, double* cgrn, const int32_t nlo[3],
                            const int32_t nhi[3], const int32_t npad[3], const double umin[3],
                            const double delta[3], double gamma, int icomp, int igfflag);

/* ---- rectangular pipe --------------------------------------------------------
 * osc_alloc_rectpipe_arrays (core:876-898): fixes the pipe grid in the context. */
int osc_alloc_rectpipe_arrays(osc_ctx* ctx, const int32_t nlo[3], const int32_t nhi[3],
                              const int32_t npad[3]);
/* osc_getgrnpipe (core:632-715): builds the four pipe Green spectra on the GPU */
int osc_getgrnpipe(osc_ctx* ctx, double gamma, double apipe, double bpipe, const double delta[3],
                   const double umin[3], const int32_t npad[3]);
/* osc_write_rectpipe_grn (core:851-870) / osc_read_rectpipe_grn (core:825-844):
 * upstream file format (Fortran unformatted sequential, 4-byte markers).
 * hdr_real (12: a,b,dx,dy,dz,xmin,ymin,zmin,xmax,ymax,zmax,gamma) and
 * hdr_int (9: nlo,nhi,period) receive the header and may be NULL. */
int osc_write_rectpipe_grn(osc_ctx* ctx, const char* path, double apipe, double bpipe,
                           const double delta[3], const double umin[3], const double umax[3],
                           const int32_t nlo[3], const int32_t nhi[3], double gamma);
int osc_read_rectpipe_grn(osc_ctx* ctx, const char* path, double* hdr_real, int32_t* hdr_int);
/* space_charge_rectpipe (open_spacecharge_mod.f90:110-127) ->
 * osc_rectpipe_solver (core:536-624).  direct_field_calc != 0 returns
 * OSC_ERR_PIPE_DIRECT with efield untouched and bfield computed from it. */
int osc_space_charge_rectpipe(osc_ctx* ctx, const osc_mesh* mesh, double apipe, double bpipe,
                              int direct_field_calc, int integrated_green_function,
                              const double* rho, double* phi, double* efield, double* bfield);

/* ---- gather ------------------------------------------------------------------
 * interpolate_field (open_spacecharge_mod.f90:311-390) for n points.  E and
 * B are (3,n); either field/output pair may be NULL.  Points off the mesh are
 * clamped to the edge cell as upstream; clamped (may be NULL) receives 1 for
 * those points and *n_clamped their count. */
int osc_interpolate_field(osc_ctx* ctx, const osc_mesh* mesh, const double* efield,
                          const double* bfield, int64_t n, const double* x, const double* y,
                          const double* z, double* E, double* B, int8_t* clamped,
                          int64_t* n_clamped);
/* No upstream counterpart: osc_interpolate_field for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results bit-identical;
 * every array must be device memory, and the mesh may have at most
 * OSC_NOALLOC_MAX_NODES nodes (else OSC_ERR_ARG).  Its scratch (the packed
 * node records, 64 MiB) is a fixed device array per GPU, held until the
 * process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race and must be serialised by the caller.  Points are gathered in
 * their own order (no cell sort), so very large batches are slower than
 * osc_interpolate_field.  n_clamped, when given, waits for the result. */
#define OSC_NOALLOC_MAX_NODES 1048576
int osc_interpolate_field_noalloc(osc_ctx* ctx, const osc_mesh* mesh, const double* efield,
                                  const double* bfield, int64_t n, const double* x,
                                  const double* y, const double* z, double* E, double* B,
                                  int8_t* clamped, int64_t* n_clamped);

/* ---- request loop (no upstream counterpart) -----------------------------------
 * A stream of independent solves -- free space, cathode images (the
 * space_charge_3d at_cathode path) and rectangular pipe -- each one
 * deposit -> solve -> gather of E (and B) at its particles, as a tracking code
 * would call deposit_particles, the solver and interpolate_field.  nworkers
 * host threads each own a context and HIP stream, so requests overlap on the
 * GPU.  Requests wait in one queue; an idle worker starts the oldest one that
 * fits the memory budget, preferring one whose geometry (solver, mesh shape,
 * fixed mesh, gamma, pipe size) it served recently, so that its cached Green
 * spectra are reused (green_cache != 0).  The budget (max_bytes, or 7/8 of the
 * free memory when max_bytes <= 0) covers the estimated bytes of the running
 * requests and what the workers' contexts hold between requests (Green
 * caches, capped at budget/(8 nworkers) each, pipe Green functions, pooled
 * buffers); a request larger than the budget runs alone.  Particle and output
 * arrays are device memory and must stay valid until osc_server_wait returns
 * for the request. */
enum osc_solver { OSC_SOLVE_FREE = 0, OSC_SOLVE_CATHODE = 1, OSC_SOLVE_PIPE = 2 };
typedef struct osc_request {
  int32_t solver;          /* osc_solver                                             */
  int32_t precision;       /* OSC_PREC_FP64 or OSC_PREC_FP32                          */
  int32_t nhi[3];          /* mesh nodes per axis (nlo = npad = 1)                    */
  int32_t resize_mesh;     /* 1: mesh from the bunch (deposit_particles rule); 0: min/max */
  int64_t n;               /* particles                                              */
  const double *x, *y, *z; /* coordinates                                            */
  double total_charge;
  double gamma;
  double min[3], max[3];   /* fixed mesh when resize_mesh == 0 (required by the pipe)  */
  double apipe, bpipe;     /* pipe cross-section (OSC_SOLVE_PIPE)                     */
  double *E, *B;           /* (3,n) outputs; B may be NULL                             */
} osc_request;
typedef struct osc_server osc_server;
int osc_server_create(int device, int nworkers, int64_t max_bytes, int green_cache,
                      osc_server** server);
int osc_server_submit(osc_server* server, const osc_request* req, int64_t* ticket);
/* blocks until the request is done; times_ms (3, may be NULL): queued, run,
 * total (submit to done); counts (2, may be NULL): particles dropped by the
 * deposit, points clamped by the gather.  Returns the request's status. */
int osc_server_wait(osc_server* server, int64_t ticket, double* times_ms, int64_t* counts);
int osc_server_drain(osc_server* server);
/* waits until no request is in flight, then drops every worker's cached Green
 * spectra and pipe Green functions into its buffer pool (a service after its
 * lattice or bunch set changes; a benchmark repeating a stream of new bunches
 * on a warm server).  No device memory is freed.  No upstream counterpart:
 * upstream keeps one set of pipe Green functions and no spectrum cache. */
int osc_server_clear_caches(osc_server* server);
/* counters since creation, the first n of: requests done, most requests
 * running at once, requests started on a worker that served the same geometry
 * before, requests started ahead of an older one that did not fit, Green
 * spectrum cache hits, misses (over all workers), bytes the idle workers'
 * contexts hold, the byte budget, times a worker gave pooled buffers back
 * (each stalls the device) and the bytes given back, device allocations of
 * the workers' buffer pools and the host microseconds they took */
#define OSC_SERVER_NSTATS 12
int osc_server_stats(osc_server* server, int64_t* out, int n);
const char* osc_server_last_error(const osc_server* server);
int osc_server_destroy(osc_server* server);

#ifdef __cplusplus
}
#endif
#endif /* OPENSC_HIP_H */
