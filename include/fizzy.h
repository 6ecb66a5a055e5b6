/*
 * fizzy.h -- C ABI of the MI355X-native CaNS-Fizzy hot path (libfizzy.so).
 *
 * Every entry point replaces one routine of the reference's Fortran module
 * interface; the reference file:line it stands in for is given next to each
 * declaration. The Fortran modules under cans-fizzy_amd/fortran/ bind these
 * entry points through iso_c_binding under the reference's module and
 * routine names.
 *
 * Conventions
 *  - Fields are device pointers (hipMalloc'd, e.g. through fz_malloc) to
 *    Fortran-ordered halo'd arrays: element (i,j,k), 0 <= i <= n1+1, lives at
 *    offset i + (n1+2)*(j + (n2+2)*k) (src/main.f90:162-204). AB2 right-hand
 *    sides are interior-only arrays, (i,j,k) at (i-1) + n1*((j-1) + n2*(k-1))
 *    (src/rk.f90:60-61).
 *  - Boundary-condition codes are 'P','D','N' laid out like the Fortran
 *    arrays cbc(0:1,3) (6 chars: x0 x1 y0 y1 z0 z1) and cbcvel(0:1,3,3)
 *    (18 chars: the 6 of u, then v, then w).
 *  - Scalars are passed by value; results the reference returns as scalars
 *    are written to host pointers (the call then synchronises its stream).
 *  - All work is enqueued on the context's HIP stream in call order, as the
 *    reference enqueues on OpenACC queue 1.
 *  - Return value: FZ_OK (0) or a negative FZ_ERR_* code; fz_last_error()
 *    gives a message. No exception crosses the ABI.
 *  - A context (and a case) is driven by one host thread; distinct contexts
 *    are independent.
 */
#ifndef FIZZY_H
#define FIZZY_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FZ_OK 0
#define FZ_ERR_ARG (-1)
#define FZ_ERR_HIP (-2)
#define FZ_ERR_UNSUPPORTED (-3)
#define FZ_ERR_STATE (-4)

typedef struct fz_ctx fz_ctx;
typedef struct fz_case fz_case;
typedef struct fz_rk_state fz_rk_state;

/* Case parameters: the &dns / &two_fluid namelist entries the hot path reads
 * (src/param.f90:91-115) and the build switches it depends on
 * (configs/flags.mk: ACDI or THINC/QQ, SDF normals, Brackbill normalisation,
 * constant-coefficient pressure splitting are the supported build). */
typedef struct {
  int ng[3];
  double l[3];
  int gtype;
  double gr;
  double cfl, dtmax, dt_f;
  int is_solve_ns, is_track_interface, is_forced_hit;
  int icheck;
  int vof;            /* 0: ACDI, 1: THINC/QQ (_INTERFACE_CAPTURING_VOF) */
  char cbcvel[18];
  char cbcpre[6];
  char cbcpsi[6];
  char cbcnor[18];
  double bcvel[18];
  double bcpre[6];
  double bcpsi[6];
  double bcnor[18];
  double bforce[3], gacc[3];
  double sigma;
  double rho12[2], mu12[2], beta12[2];
  double psi_thickness_factor, acdi_gam_factor, acdi_gam_min;
  int nspheres;       /* spheres.in records (x,y,z,r), src/two_fluid.f90:439-471 */
  int inipsi;         /* 0 uni,1 zer,2 bub3,3 bub2,4 bub1,5 drp3,6 drp2,7 drp1,8 flm */
  double spheres[4 * 64];
} fz_params;

const char *fz_last_error(void);

/* ---- context: grid, boundary conditions of the pressure, stream ---- */
/* replaces the set-up of src/main.f90:156-291 (initmpi, initgrid, inverse
 * spacings, workspaces); dzc..zf are host arrays of length n3+2 */
int fz_ctx_create(fz_ctx **ctx, const int n[3], const double l[3], const double *dzc, const double *dzf,
                  const double *zc, const double *zf, const char cbcpre[6], void *hip_stream);
int fz_ctx_destroy(fz_ctx *ctx);
int fz_ctx_sync(fz_ctx *ctx);
int fz_malloc(fz_ctx *ctx, size_t bytes, void **dptr);
int fz_free(fz_ctx *ctx, void *dptr);
int fz_upload(fz_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int fz_download(fz_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
int fz_copy(fz_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);
/* per-launch timing with HIP events on the context stream (no reference
 * counterpart; the reference times whole steps under _TIMING, src/main.f90:415-418) */
int fz_prof_enable(fz_ctx *ctx, int on);
/* one line per launch site, "name calls total_ms"; clears the records */
int fz_prof_report(fz_ctx *ctx, char *buf, size_t cap);

/* initgrid, src/initgrid.f90:14-78 (host routine, arrays of length n+2) */
int fz_initgrid(int gtype, int n, double gr, double lz, double *dzc, double *dzf, double *zc, double *zf);

/* ---- Poisson solver ---- */
/* initsolver, src/initsolver.f90:15-79; host outputs as in the reference
 * (lambdaxy n1*n2, a,b,c n3, rhsbx n2*n3*2, rhsby n1*n3*2, rhsbz n1*n2*2;
 * any of them may be NULL) and builds the device plan held by ctx */
int fz_initsolver(fz_ctx *ctx, const double bcpre[6], double *lambdaxy, double *a, double *b, double *c,
                  double *normfft, double *rhsbx, double *rhsby, double *rhsbz);
/* solver / solver_gpu, src/solver.f90:19-78, src/solver_gpu.f90:31-163 */
int fz_solver(fz_ctx *ctx, double *p);
/* solver with FP32 transforms and line solves (FP64 pivots) wrapped in
 * FP64-residual iterative refinement; p holds the right-hand side on entry
 * and the solution on exit, as for fz_solver. Refinements stop when
 * max|b - A x| <= tol*max|b|, after maxit of them, or when the residual
 * stops falling; *iters receives the refinements done and *relres (if not
 * null) the final max|b - A x|/max|b|. The null mode of an all-Neumann or
 * periodic operator is fixed by pinning one unknown of the mean column
 * (compare with the mean removed). No reference counterpart. */
int fz_solver_mixed(fz_ctx *ctx, double *p, double tol, int maxit, int *iters, double *relres);
/* fz_solver on n contexts of one shape (grid, spacing and pressure
 * conditions; each after fz_initsolver) with one launch per stage for all of
 * them: ps[q] is context q's field, solved in place as fz_solver does, bit
 * for bit. Runs on ctxs[0]'s stream after every context's stream has
 * drained, and returns when the solves are done. Periodic z is not
 * batched. No reference counterpart. */
int fz_solver_batched(fz_ctx *const *ctxs, double *const *ps, int n);
/* fft, src/fft.f90:176-187: an r2r transform (FFTW kind codes 0 R2HC, 1 HC2R,
 * 4 REDFT01, 5 REDFT10, 6 REDFT11, 8 RODFT01, 9 RODFT10, 10 RODFT11) over
 * `howmany` contiguous lines of length n, in place and unnormalised */
int fz_r2r(fz_ctx *ctx, int kind, int n, int howmany, double *x);
/* fillps, src/fillps.f90:13-51 */
int fz_fillps(fz_ctx *ctx, double dti, double rho0, const double *u, const double *v, const double *w, double *p);
/* updt_rhs_b, src/bound.f90:341-396 (boundary terms built by fz_initsolver) */
int fz_updt_rhs_b(fz_ctx *ctx, double *p);
/* correc, src/correc.f90:13-67 (constant-coefficient build; psi, rho12 unused) */
int fz_correc(fz_ctx *ctx, double rho0, const double rho12[2], double dt, const double *p, const double *psi,
              double *u, double *v, double *w);
/* updatep, src/updatep.f90:13-24 */
int fz_updatep(fz_ctx *ctx, const double *pp, double *p);
/* chk_poisson, src/debug.f90:81-114: max |L(fpp) - fp| */
int fz_chk_poisson(fz_ctx *ctx, const double *fp, const double *fpp, double *diffmax);

/* ---- boundary conditions ---- */
/* boundp, src/bound.f90:81-124 */
int fz_boundp(fz_ctx *ctx, const char cbc[6], const double bc[6], double *p);
/* bounduvw, src/bound.f90:15-79 */
int fz_bounduvw(fz_ctx *ctx, const char cbc[18], const double bc[18], int is_correc, double *u, double *v,
                double *w);

/* ---- phase field ---- */
/* acdi_transport_pf, src/acdi.f90:64-225 */
int fz_acdi_transport_pf(fz_ctx *ctx, double gam, double seps, const double *u, const double *v, const double *w,
                         const double *normx, const double *normy, const double *normz, const double *phi,
                         const double *psi, double *dpsidt, double *flux_x, double *flux_y, double *flux_z);
/* acdi_cmpt_phi, src/acdi.f90:227-248 */
int fz_acdi_cmpt_phi(fz_ctx *ctx, double seps, const double *psi, double *phi);
/* acdi_set_gamma, src/acdi.f90:32-62 (synchronises) */
int fz_acdi_set_gamma(fz_ctx *ctx, double gam_factor, const double *u, const double *v, const double *w,
                      double *gam);
/* acdi_set_epsilon, src/acdi.f90:16-30 (host routine) */
int fz_acdi_set_epsilon(fz_ctx *ctx, double seps_factor, double *seps);
/* vof_thinc_transport_psi, src/vof_thinc_qq.f90:39-133 */
int fz_vof_thinc_transport_psi(fz_ctx *ctx, double beta, const double *u, const double *v, const double *w,
                               const double *normx, const double *normy, const double *normz, const double *phi,
                               const double *psi, double *dpsidt, double *flux_x, double *flux_y, double *flux_z);
/* vof_thinc_cmpt_phi, src/vof_thinc_qq.f90:135-157 */
int fz_vof_thinc_cmpt_phi(fz_ctx *ctx, double beta, const double *psi, double *phi);
/* clip_field, src/two_fluid.f90:35-58 (nh = 1) */
int fz_clip_field(fz_ctx *ctx, const double minmax[2], double *p);
/* cmpt_norm_curv_youngs, src/two_fluid.f90:60-175 */
int fz_cmpt_norm_curv_youngs(fz_ctx *ctx, const double *psi, double *normx, double *normy, double *normz,
                             double *kappa);
/* init2fl, src/two_fluid.f90:247-437 (spheres taken from prm) */
int fz_init2fl(fz_ctx *ctx, const fz_params *prm, double seps, double *psi);

/* ---- momentum ---- */
/* mom_xyz_ad, src/mom.f90:505-851 */
int fz_mom_xyz_ad(fz_ctx *ctx, const double rho12[2], const double mu12[2], const double *psiflx_x,
                  const double *psiflx_y, const double *psiflx_z, const double *u, const double *v,
                  const double *w, const double *psi, double *dudt, double *dvdt, double *dwdt);
/* mom_xyz_oth, src/mom.f90:853-1039 (beta12 unused without the scalar) */
int fz_mom_xyz_oth(fz_ctx *ctx, double dt_r, const double rho12[2], const double beta12[2], const double bforce[3],
                   const double gacc[3], double sigma, double rho0, const double rho_av[3], const double *p,
                   const double *psi, const double *kappa, const double *pn, const double *po, double *dudt,
                   double *dvdt, double *dwdt);
/* lscale_forcing, src/forcing.f90:13-65 (ftype 1 ABC, 2 Taylor-Green) */
int fz_lscale_forcing(fz_ctx *ctx, int ftype, double alpha, double dt, double *u, double *v, double *w);
/* bulk_mean_12_stag, src/utils.f90:68-136 (synchronises) */
int fz_bulk_mean_12_stag(fz_ctx *ctx, int idir, const double *psi, const double p12[2], double *mean);

/* ---- time integration: rk and rk_2fl with their AB2 history held in a
 * state object instead of `save` variables (src/rk.f90:44-70, :214-228) ---- */
int fz_rk_state_create(fz_ctx *ctx, fz_rk_state **st);
int fz_rk_state_destroy(fz_rk_state *st);
/* rk, src/rk.f90:16-131 */
int fz_rk(fz_ctx *ctx, fz_rk_state *st, const double rkpar[2], double dt, double dt_r, const double bforce[3],
          const double gacc[3], double sigma, const double rho_av[3], const double rho12[2], const double mu12[2],
          const double beta12[2], double rho0, const double *psi, const double *kappa, const double *p,
          const double *pn, const double *po, const double *psio, const double *psiflx_x, const double *psiflx_y,
          const double *psiflx_z, double *u, double *v, double *w);
/* rk_2fl, src/rk.f90:193-254 (vof selects THINC/QQ, else ACDI) */
int fz_rk_2fl(fz_ctx *ctx, fz_rk_state *st, const double rkpar[2], double dt, double gam, double seps, double beta,
              int vof, const double *u, const double *v, const double *w, const double *normx, const double *normy,
              const double *normz, const double *phi, double *psi, double *psiflx_x, double *psiflx_y,
              double *psiflx_z);

/* ---- checks ---- */
/* chkdiv, src/chkdiv.f90:15-49 (synchronises) */
int fz_chkdiv(fz_ctx *ctx, const double *u, const double *v, const double *w, double *divtot, double *divmax);
/* chkdiv into device memory without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops). The same kernels and summation order as
 * fz_chkdiv; d_out[0] = divtot, d_out[1] = divmax, written asynchronously on
 * the context's stream, with no synchronisation. The per-block partials live
 * in a static device array of each GPU (1024 blocks x 2 doubles, 16 KiB,
 * held until the process ends) that every call on that GPU shares: calls
 * that overlap on one GPU race. */
int fz_chkdiv_noalloc(fz_ctx *ctx, const double *u, const double *v, const double *w, double *d_out);
/* chkdt, src/chkdt.f90:15-94 (synchronises) */
int fz_chkdt(fz_ctx *ctx, int is_solve_ns, int is_track_interface, int vof, const double mu12[2],
             const double rho12[2], double sigma, const double gacc[3], const double *u, const double *v,
             const double *w, double gam, double seps, double *dtmax);

/* ---- device math (tests) ---- */
/* The library's FP64 routines evaluated on the device over n elements:
 * fn 0 exp, 1 expm1, 2 log, 3 tanh, 4 sin, 5 cos (the routines of the ACDI,
 * THINC/QQ and forcing kernels), 6 the in-range division x/y of the tanh,
 * 7 the compiler's division x/y. No counterpart in the reference: for tests
 * of the device routines against the oracle's copies (asynchronous). */
int fz_math_eval(fz_ctx *ctx, int fn, long n, const double *x, const double *y, double *out);
/* acdi_cmpt_phi (src/acdi.f90:227-248) then cmpt_norm_curv_youngs
 * (src/two_fluid.f90:81-174) as the FP64 ACDI step runs them: phi from psi,
 * raising the device word *flag (cleared first) when some phi is neither
 * zero nor of a magnitude in [2^-100, 2^20], then the Youngs kernels that
 * flag selects (the check-free one when it is clear). For tests of both
 * paths against the oracle (asynchronous). */
int fz_phase_geometry_eval(fz_ctx *ctx, double seps, const double *psi, double *phi, double *normx, double *normy,
                           double *normz, double *kappa, unsigned *flag);

/* ---- field files ---- */
/* load_one / io_field, src/load.f90:19-168: io 'w' writes the interior of the
 * halo'd device field p followed by [time, istep]; io 'r' reads such a file
 * into the interior of p (a file of the wrong size is refused, as the
 * reference aborts on it, src/load.f90:41-51) */
int fz_load_one(fz_ctx *ctx, char io, const char *filename, double *p, double *time, int *istep);
/* out3d, src/output.f90:185-236: the interior every nskip points from index 1 */
int fz_out3d(fz_ctx *ctx, const char *filename, const int nskip[3], const double *p);
/* grid.bin, grid.out and geometry.out of src/main.f90:230-243 into datadir
 * (a prefix, e.g. "data/") */
int fz_write_grid(fz_ctx *ctx, const char *datadir);

/* ---- a whole case: the time loop of src/main.f90:414-562 on the device ---- */
/* field ids: 0 u,1 v,2 w,3 p,4 pn,5 po,6 psi,7 phi,8 kappa,9 normx,10 normy,11 normz,
 * 12 psiflx_x,13 psiflx_y,14 psiflx_z (halo'd); 15 dpsidt_old,16 dudt_old,17 dvdt_old,
 * 18 dwdt_old (interior) */
/* scalars: 0 time,1 dt,2 dto,3 gam,4 seps,5 rho0,6 divtot,7 divmax,8 dt_cfl,9 istep,10 is_first */
int fz_case_create(fz_case **cs, const fz_params *prm, void *hip_stream);
/* precision modes of a case (no reference counterpart: the reference builds
 * all-FP64 or, with -D_SINGLE_PRECISION, all-FP32) */
#define FZ_PREC_FP64 0
/* FP32 phase field (psi, phi, normals, curvature, phase fluxes) and FP32 AB2
 * histories with FP32 arithmetic in the phase-field kernels; velocity,
 * pressure, the momentum arithmetic and the global sums in FP64; the
 * projection solves with FP32 transforms and line solves (FP64 pivots)
 * followed by nref FP64-residual refinements (nref = -1: the plain
 * step's FP64 projection instead). Two-fluid cases (ACDI or THINC/QQ) with
 * homogeneous pressure boundary values. */
#define FZ_PREC_MIXED 1
int fz_case_create_ex(fz_case **cs, const fz_params *prm, void *hip_stream, int precision, int nref);
int fz_case_precision(fz_case *cs, int *precision, int *nref);
/* the FP32 fields of a mixed-precision case (fz_case_field refuses them) */
int fz_case_field_f32(fz_case *cs, int id, float **dptr, size_t *count);
int fz_case_destroy(fz_case *cs);
int fz_case_ctx(fz_case *cs, fz_ctx **ctx);
int fz_case_field(fz_case *cs, int id, double **dptr, size_t *count);
int fz_case_get_scalars(fz_case *cs, double out[12]);
int fz_case_set_scalars(fz_case *cs, const double in[12]);
/* initflow 'zer' + init2fl + boundary conditions + phi/normals + gamma + dt
 * (src/main.f90:300-407) */
int fz_case_init(fz_case *cs);
/* nsteps time steps; *killed = 1 when a divergence or time-step check
 * stopped the run as src/main.f90:530-562 does */
int fz_case_step(fz_case *cs, int nsteps, int *killed);
/* fz_case_step(cs, 1, killed) in two halves, for a caller that steps many
 * cases: begin enqueues the step (and its checks when the step number is a
 * multiple of icheck, *check_pending = 1) without waiting; end waits for the
 * checks, if any, and sets the next dt and *killed. The same results as
 * fz_case_step; nothing else may be called on the case in between. */
int fz_case_step_begin(fz_case *cs, int *check_pending);
int fz_case_step_end(fz_case *cs, int *killed);
/* replay each time step after the first from a captured HIP graph (one per
 * parity of the buffer rotation; dt, its AB2 factors and gamma are read
 * from device memory, so the graphs serve every step). Applies to ACDI
 * two-fluid cases whose step needs no host read-back; other cases, and
 * steps with per-launch profiling on, run as before. Results are the same
 * bit for bit. No reference counterpart. */
int fz_case_set_graphs(fz_case *cs, int on);
/* pick the case's launch parameters (stencil block order, momentum-kernel
 * occupancy, Thomas rows per chunk) by timing trial_steps steps per
 * candidate on a scratch case of the same parameters; applies the choice to
 * the case and caches it per shape for later cases. Results are the same
 * bit for bit under every choice. report (if not null) receives the choice
 * and its time. No reference counterpart. */
int fz_case_autotune(fz_case *cs, int trial_steps, char *report, size_t cap);
/* batched steps of same-shaped cases (no reference counterpart): one launch
 * per stage of the step for all the cases, each block taking its case's
 * buffers and scalars from a device table. Cases qualify when their first
 * step is taken and they share the grid, every boundary condition and
 * precision: ACDI two-fluid cases whose step needs no host read-back, FP64 or
 * mixed precision with the FP64 projection (nref = -1), a non-periodic z and
 * homogeneous pressure boundary values; create refuses others with
 * FZ_ERR_UNSUPPORTED and the reason. A step of the batch gives every case
 * the fields, scalars and checks of its own fz_case_step bit for bit; the
 * work runs on the first case's stream after each case's earlier work, and
 * each case's later work waits for it. nsteps steps, or fewer when a check
 * stops a case (killed[q] = 1; killed may be null); cases must not be
 * stepped otherwise while in a batch, and a batch must be made again once
 * its cases have been stepped alone. */
typedef struct fz_case_batch fz_case_batch;
int fz_case_batch_create(fz_case *const *cases, int n, fz_case_batch **batch);
int fz_case_batch_step(fz_case_batch *batch, int nsteps, int *killed);
/* one batched step in two halves, as fz_case_step_begin/_end: begin enqueues
 * it (and the checks due, *check_pending = 1) without waiting; end reads the
 * checks back, if any, and sets killed[] (n entries, or null) */
int fz_case_batch_step_begin(fz_case_batch *batch, int *check_pending);
int fz_case_batch_step_end(fz_case_batch *batch, int *killed);
int fz_case_batch_destroy(fz_case_batch *batch);
/* the checkpoint of src/main.f90:597-631: prefix_u.bin, _v, _w, _p, _psi
 * in the load_one format, with the case's time and step */
int fz_case_save(fz_case *cs, const char *prefix);
/* restart, src/main.f90:312-407 with restart = T: the five fields, time and
 * step from such files, then the same boundary conditions, phi/normals,
 * gamma and dt as a fresh start; the AB2 histories start over */
int fz_case_restart(fz_case *cs, const char *prefix);

/* ---- serving: many cases on one device (no reference counterpart; the
 * footprint model follows device_memory_footprint, src/utils.f90:155-219,
 * for this library's buffers) ---- */
typedef struct fz_server fz_server;
/* device bytes fz_case_create + fz_case_init allocate for prm */
int fz_case_footprint(const fz_params *prm, size_t *bytes);
/* the same for a precision mode of fz_case_create_ex */
int fz_case_footprint_ex(const fz_params *prm, int precision, size_t *bytes);
/* free and total memory of the current device */
int fz_device_memory(size_t *free_bytes, size_t *total_bytes);
int fz_server_create(fz_server **s, size_t hbm_budget);
int fz_server_destroy(fz_server *s);
/* queue a case for nsteps time steps; its footprint must fit the budget */
int fz_server_submit(fz_server *s, const fz_params *prm, int nsteps, int *id);
/* the same in a precision mode of fz_case_create_ex */
int fz_server_submit_ex(fz_server *s, const fz_params *prm, int nsteps, int precision, int nref, int *id);
/* admit queued cases in arrival order while their footprints fit (the
 * server's set-up thread creates and initialises them while other cases
 * run; a round with no running case waits for one), then one time step
 * of every running case on its own stream (with FZ_SERVE_BATCH=b in the
 * environment, up to b steps for cases whose steps are shorter than the
 * slowest one's, fewer when a check step comes first); waits for the
 * round. n_queued counts the requests not yet running. */
int fz_server_round(fz_server *s, int *n_active, int *n_queued);
/* 0 queued, 1 running, 2 finished (killed: a check stopped it), 3 failed
 * (its set-up or a step returned an error: that case alone is dropped, its
 * memory returned, and the other cases go on), -1 unknown */
int fz_server_status(fz_server *s, int id, int *state, int *killed);
/* the error message of a failed request (an empty string for any other) */
int fz_server_error(fz_server *s, int id, char *buf, size_t cap);
/* a finished case (fields via fz_case_field) until it is released; a
 * failed request has none (FZ_ERR_STATE), and its release only forgets it */
int fz_server_case(fz_server *s, int id, fz_case **cs);
int fz_server_release(fz_server *s, int id);
/* durations (ms, HIP events on the case's stream) of the steps of request
 * id, or of every step (id 0); *count is the total, at most cap copied */
int fz_server_step_times(fz_server *s, int id, float *ms, int cap, int *count);
int fz_server_memory(fz_server *s, size_t *in_use, size_t *budget);

#ifdef __cplusplus
}
#endif
#endif
