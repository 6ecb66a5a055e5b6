/*
 * CPU oracle for the CaNS-Fizzy per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C restatement of the
 * reference Fortran routines; it is used by tests/, by __graft_entry__.smoke()
 * as the checker and by bench.py's cpu_baseline leg. The product (the HIP
 * library under cans-fizzy_amd/) never links, loads or calls it.
 *
 * Conventions (same as the reference, src/main.f90:162-204):
 *  - halo'd fields are column-major with a one-cell halo, element (i,j,k),
 *    0 <= i <= n1+1, at offset i + (n1+2)*(j + (n2+2)*k);
 *  - interior-only arrays (AB2 right-hand sides, src/rk.f90:60-61) hold
 *    (i,j,k), 1 <= i <= n1, at offset (i-1) + n1*((j-1) + n2*(k-1));
 *  - grid vectors dzc,dzf,zc,zf,dzci,dzfi have n3+2 entries (k = 0..n3+1);
 *  - boundary-condition codes are single characters 'P','D','N' laid out as
 *    the Fortran arrays cbc(0:1,3) (6 chars) and cbcvel(0:1,3,3) (18 chars).
 *
 * Pinning status of every function is listed in the header comment of
 * fz_oracle.c next to the function.
 */
#ifndef FZ_ORACLE_H
#define FZ_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* FFTW r2r kind codes, as in the reference binding src/fftw.f90:42-62 */
enum { OR_R2HC = 0, OR_HC2R = 1, OR_REDFT00 = 3, OR_REDFT01 = 4, OR_REDFT10 = 5,
       OR_REDFT11 = 6, OR_RODFT00 = 7, OR_RODFT01 = 8, OR_RODFT10 = 9, OR_RODFT11 = 10 };

/* case parameters: the &dns/&two_fluid namelist entries the hot path reads
 * (src/param.f90:91-115), plus the compile-time switches it depends on. */
typedef struct {
  int ng[3];
  double l[3];
  int gtype;
  double gr;
  double cfl, dtmax, dt_f;
  int is_solve_ns, is_track_interface, is_forced_hit;
  int icheck;
  int vof;            /* 0: ACDI (default build), 1: THINC/QQ (_INTERFACE_CAPTURING_VOF) */
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
  int nspheres;       /* spheres for inipsi (x,y,z,r) */
  int inipsi;         /* 0 'uni', 1 'zer', 2 'bub3', 3 'bub2', 4 'bub1', 5 'drp3', 6 'drp2', 7 'drp1', 8 'flm' */
  double spheres[4 * 64];
} or_params;

/* --- r2r transforms (FFTW definitions, unnormalised) --- */
void or_r2r(int kind, int n, int howmany, double *x);   /* lines contiguous, back to back */

/* --- grid and solver set-up --- */
void or_initgrid(int gtype, int n, double gr, double lz, double *dzc, double *dzf, double *zc, double *zf);
void or_initsolver(const int ng[3], const double dli[3], const double *dzci, const double *dzfi,
                   const char cbc[6], const double bc[6], double *lambdaxy, double *a, double *b, double *c,
                   double *normfft, double *rhsbx, double *rhsby, double *rhsbz);
void or_solver(const int n[3], const char cbc[6], double normfft, const double *lambdaxy,
               const double *a, const double *b, const double *c, double *p);
double or_chk_poisson(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                      const double *fp, const double *fpp);

/* --- projection --- */
void or_fillps(const int n[3], const double dli[3], const double *dzfi, double dti, double rho0,
               const double *u, const double *v, const double *w, double *p);
void or_correc(const int n[3], const double dli[3], const double *dzci, double rho0, double dt,
               const double *p, double *u, double *v, double *w);
void or_updatep(const int n[3], const double *pp, double *p);
void or_updt_rhs_b(const int n[3], const char cbc[6], const double *rhsbx, const double *rhsby,
                   const double *rhsbz, double *p);

/* --- boundary conditions --- */
void or_boundp(const int n[3], const char cbcpre[6], const char cbc[6], const double bc[6],
               const double dl[3], const double *dzc, double *p);
void or_bounduvw(const int n[3], const char cbcpre[6], const char cbc[18], const double bc[18], int is_correc,
                 const double dl[3], const double *dzc, const double *dzf, double *u, double *v, double *w);

/* --- phase field --- */
void or_acdi_transport_pf(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                          double gam, double seps, const double *u, const double *v, const double *w,
                          const double *normx, const double *normy, const double *normz, const double *phi,
                          const double *psi, double *dpsidt, double *flux_x, double *flux_y, double *flux_z);
void or_acdi_cmpt_phi(const int n[3], double seps, const double *psi, double *phi);
double or_acdi_set_gamma(const int n[3], double gam_factor, const double *u, const double *v, const double *w);
double or_acdi_set_epsilon(const double dl[3], const double *dzfi, int n3, double seps_factor);
void or_vof_thinc_transport_psi(const int n[3], const double dli[3], const double *dzfi, double beta,
                                const double *u, const double *v, const double *w, const double *normx,
                                const double *normy, const double *normz, const double *phi,
                                double *dpsidt, double *flux_x, double *flux_y, double *flux_z);
void or_vof_thinc_cmpt_phi(const int n[3], double beta, const double *psi, double *phi);
void or_clip_field(const int n[3], double rmin, double rmax, double *p);
void or_cmpt_norm_curv_youngs(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                              const double *psi, double *normx, double *normy, double *normz, double *kappa);

/* --- momentum --- */
void or_mom_xyz_ad(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                   const double rho12[2], const double mu12[2], const double *psiflx_x,
                   const double *psiflx_y, const double *psiflx_z, const double *u, const double *v,
                   const double *w, const double *psi, double *dudt, double *dvdt, double *dwdt);
void or_mom_xyz_oth(const int n[3], const double dli[3], const double *dzci, const double *dzfi, double dt_r,
                    const double rho12[2], const double bforce[3], const double gacc[3], double sigma,
                    double rho0, const double rho_av[3], const double *p, const double *psi,
                    const double *kappa, const double *pn, const double *po,
                    double *dudt, double *dvdt, double *dwdt);
void or_rk_mom_ab2(const int n[3], double factor1, double factor2, const double rho12[2], const double *psio,
                   const double *psi, const double *dudt, const double *dvdt, const double *dwdt,
                   const double *dudto, const double *dvdto, const double *dwdto,
                   double *u, double *v, double *w);
void or_rk_mom_oth(const int n[3], double factor12, const double rho12[2], const double *psi,
                   const double *dudt, const double *dvdt, const double *dwdt, double *u, double *v, double *w);
void or_rk_psi_ab2(const int n[3], double factor1, double factor2, const double *dpsidt,
                   const double *dpsidto, double *psi);
void or_lscale_forcing_tg(const int n[3], double alpha, double dt, const double l[3], const double dl[3],
                          const double *zc, const double *zf, double *u, double *v);
double or_bulk_mean_12_stag(const int n[3], int idir, const double *grid_vol_ratio, const double *psi,
                            const double p12[2]);

/* --- checks --- */
void or_chkdiv(const int n[3], const double dli[3], const double *dzfi, const double *u, const double *v,
               const double *w, double out[2]);
double or_chkdt(const int n[3], const double dl[3], const double *dzci, const double *dzfi, int is_solve_ns,
                int is_track_interface, int vof, const double mu12[2], const double rho12[2], double sigma,
                const double gacc[3], const double *u, const double *v, const double *w, double gam, double seps);

/* --- initial condition --- */
void or_init2fl(const or_params *prm, double seps, const double *zc, double *psi);

void or_math_eval(int fn, long n, const double *x, double *y);

/* --- a whole case: the step loop of src/main.f90:414-562 --- */
typedef struct or_case or_case;
or_case *or_case_create(const or_params *prm);
void or_case_destroy(or_case *c);
/* field handles: 0 u,1 v,2 w,3 p,4 pn,5 po,6 psi,7 phi,8 kappa,9 normx,10 normy,11 normz,
 * 12 psiflx_x,13 psiflx_y,14 psiflx_z (halo'd); 15 dpsidt_old,16 dudt_old,17 dvdt_old,
 * 18 dwdt_old (interior) */
double *or_case_field(or_case *c, int id);
/* scalars: 0 time,1 dt,2 dto,3 gam,4 seps,5 rho0,6 divtot,7 divmax,8 dt_cfl,9 istep,10 is_first */
double *or_case_scalars(or_case *c);
void or_case_init_fields(or_case *c);       /* initflow 'zer' + init2fl + BCs + phi/normals + gamma + dt */
void or_case_finish_init(or_case *c);       /* the same from the fields as they are (restart) */
int or_case_step(or_case *c, int nsteps);   /* returns 0, or 1 when a divergence/dt check killed the run */

#ifdef __cplusplus
}
#endif
#endif
