// A whole case on the device: the time loop of src/main.f90:414-562 and the
// AB2 integrators rk / rk_2fl of src/rk.f90, with the history arrays that the
// reference keeps in `save` variables held per case (or per fz_rk_state).
//
// The reference's full-array copies psio = psi, pp = p and po = pn; pn = p
// (src/main.f90:430-432, :488-490, :505-508) become pointer rotations:
//  - the phase-field update writes psi(n+1) into the buffer that held
//    psi(n-1), so the previous buffer is psio with its halos intact;
//  - fillps/solver/correc work on a separate correction buffer dp, and
//    updatep writes p + dp into the buffer that held the pressure of the
//    previous step, which then becomes po; pn is p itself (pn == p whenever
//    mom_xyz_oth reads it).
#include <cmath>
#include <cstring>
#include <string>

#include "fz_internal.h"

int fz_launch_rk_psi(fz_ctx *c, double f1, double f2, const double *dpsidt, const double *dpsidto, double *psi);
int fz_launch_rk_mom(fz_ctx *c, double f1, double f2, const double rho12[2], const double *psio, const double *psi,
                     const double *du, const double *dv, const double *dw, const double *duo, const double *dvo,
                     const double *dwo, double *u, double *v, double *w);
int fz_launch_rk_oth(fz_ctx *c, double f12, const double rho12[2], const double *psi, const double *du,
                     const double *dv, const double *dw, double *u, double *v, double *w);
template <typename PT, typename HT>
int fz_launch_mom_fused(fz_ctx *c, const double rho12[2], const double mu12[2], double factor1, double factor2,
                        int first, double factor12, double dt_r, const double *stc, double sigma, double rho0,
                        const double bforce[3],
                        const double gacc[3], const double rho_av[3], const PT *fx, const PT *fy, const PT *fz,
                        const double *u, const double *v, const double *w, const PT *psi, const PT *psio,
                        const double *p, const double *po, const PT *kappa, HT *duo, HT *dvo, HT *dwo, double *un,
                        double *vn, double *wn, int xper);
int fz_launch_fill(fz_ctx *c, long n, double val, double *p);
int fz_launch_correc_step(fz_ctx *c, double rho0, double dt, const double *stc, const double *dp, const double *p,
                          const double *u, const double *v, const double *w, double *un, double *vn, double *wn,
                          double *p_out, unsigned long long *slots, int xper, unsigned rlo);
int fz_launch_gamma_fin(fz_ctx *c, unsigned long long *slots, double factor, double gmin, double *out,
                        unsigned *clear);
int fz_acdi_cmpt_phi_rf(fz_ctx *c, double seps, const double *psi, double *phi, unsigned *rflag);
template <typename T>
int fz_launch_acdi_step(fz_ctx *c, const double *d_gam, double gam, double seps, double factor1, double factor2,
                        const double *stc, int first, const double *u, const double *v, const double *w,
                        const T *normx, const T *normy, const T *normz, const T *phi, const T *psio, T *dpsio,
                        T *psi_out, T *flux_x, T *flux_y, T *flux_z, int xper);
template <typename T>
int fz_youngs_step(fz_ctx *c, const T *psi, T *normx, T *normy, T *normz, T *kappa, int xper, const unsigned *rflag);
int fz_chkdiv_dev(fz_ctx *c, const double *u, const double *v, const double *w, double *d_out);
int fz_chkdt_adv_dev(fz_ctx *c, const double *u, const double *v, const double *w, double *d_out);
double fz_chkdt_finish(const fz_ctx *c, double dti, int is_solve_ns, int is_track_interface, int vof,
                       const double mu12[2], const double rho12[2], double sigma, const double gacc[3], double gam,
                       double seps);
template <typename PT>
int fz_bulk_mean_dev(fz_ctx *c, int idir, const PT *psi, const double p12[2], const double *gvr, double *d_out);
int fz_acdi_cmpt_phi_f32(fz_ctx *c, double seps, const float *psi, float *phi);
int fz_vof_phi_f32(fz_ctx *c, double beta, const float *psi, float *phi);
template <typename T>
int fz_launch_vof_step(fz_ctx *c, double beta, double factor1, double factor2, const double *stc, int first,
                       const double *u, const double *v, const double *w, const T *normx, const T *normy,
                       const T *normz, const T *phi, const T *psio, T *dpsio, T *psi_out, T *flux_x, T *flux_y,
                       T *flux_z);
int fz_launch_forcing(fz_ctx *c, int ftype, double alpha, double dt, const double *stc, double *u, double *v,
                      double *w);
int fz_youngs_f32(fz_ctx *c, const float *psi, float *normx, float *normy, float *normz, float *kappa);
int fz_boundp_f32(fz_ctx *c, const char cbc[6], const double bc[6], float *p);
int fz_bounduvw_f32(fz_ctx *c, const char cbc[18], const double bc[18], int is_correc, float *u, float *v, float *w);

// out-of-place AB2 phase-field update (psi_out may alias psi_in)

enum { F_U, F_V, F_W, F_P, F_PN, F_PO, F_PSI, F_PHI, F_KAPPA, F_NX, F_NY, F_NZ, F_FX, F_FY, F_FZ, F_DPSIO, F_DUO,
       F_DVO, F_DWO, F_N };
enum { S_TIME, S_DT, S_DTO, S_GAM, S_SEPS, S_RHO0, S_DIVTOT, S_DIVMAX, S_DTCFL, S_ISTEP, S_FIRST, S_N };

struct fz_case {
  fz_params prm;
  fz_ctx *ctx = nullptr;
  double *f[F_N] = {nullptr};
  double *psio = nullptr, *dp = nullptr;
  double *un = nullptr, *vn = nullptr, *wn = nullptr;  // the next velocity (rotated with u, v, w)
  double *dsc = nullptr;  // device scalars: 0 gam, 1-2 div, 3 adv dt, 4-6 rho_av
  unsigned long long *gslots = nullptr;  // maxima of the squared cell speed (gamma), zero between steps
  double *gvr_c = nullptr, *gvr_f = nullptr;
  double sc[12] = {0};
  char cbccur[6];
  double bccur[6];
  bool gam_on_device = false;
  // mixed precision: the phase field, its geometry, fluxes and the AB2
  // histories in FP32 (g[id], f[id] null); velocity, pressure and the
  // projection stay FP64
  int precision = FZ_PREC_FP64;
  int nref = 1;
  float *g[F_N] = {nullptr};
  float *psio32 = nullptr;
  // per-step scalars in device memory and the captured step graphs
  double *dstc = nullptr;
  double stc_host[FZ_STC_N] = {0};
  bool stc_valid = false;
  bool use_graphs = false;
  bool check_pending = false;  // a step's check is enqueued and not yet read
  bool capture_blocked = false;  // another thread may be making HIP calls: no graph capture now
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    const void *key[2] = {nullptr, nullptr};
  } graph[2];
};

// fields held in FP32 by the mixed-precision mode
static bool is_f32_field(int id) {
  return id == F_PSI || id == F_PHI || id == F_KAPPA || (id >= F_NX && id <= F_FZ) || id >= F_DPSIO;
}

// grouped boundary updates of the step (one gather launch per group of up
// to four fields; each field's values as its own boundp/bounduvw call gives)
template <typename T>
// (xdone: the ACDI kernel wrote the periodic x halos of the interior rows)
static int bc_fluxes_psi(fz_ctx *c, const fz_params &P, T *fx, T *fy, T *fz, T *psi, int xdone = 0) {
  const FzBcItem it[2] = {{1, P.cbcvel, P.bcvel, 0, {fx, fy, fz}, xdone},
                          {0, P.cbcpsi, P.bcpsi, 0, {psi, nullptr, nullptr}, xdone}};
  return fz_bounds_group<T>(c, 2, it);
}
// (xdone: the producing kernel wrote the periodic x halos of the interior rows)
template <typename T>
static int bc_geometry(fz_ctx *c, const fz_params &P, const char *cbccur, const double *bccur, T *kappa, T *nx, T *ny,
                       T *nz, int xdone) {
  const FzBcItem it[4] = {{0, cbccur, bccur, 0, {kappa, nullptr, nullptr}, xdone},
                          {0, P.cbcnor + 0, P.bcnor + 0, 0, {nx, nullptr, nullptr}, xdone},
                          {0, P.cbcnor + 6, P.bcnor + 6, 0, {ny, nullptr, nullptr}, xdone},
                          {0, P.cbcnor + 12, P.bcnor + 12, 0, {nz, nullptr, nullptr}, xdone}};
  return fz_bounds_group<T>(c, 4, it);
}

__global__ void k_to_f32(long n, const double *__restrict__ a, float *__restrict__ b) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) b[q] = (float)a[q];
}
__global__ void k_to_f64(long n, const float *__restrict__ a, double *__restrict__ b) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) b[q] = (double)a[q];
}
static int to_f32(fz_ctx *c, long n, const double *a, float *b) {
  FZ_LAUNCH(c, "to_f32", k_to_f32, dim3((unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096)), dim3(256), 0, n, a, b);
  FZ_LAUNCHED();
  return FZ_OK;
}
static int to_f64(fz_ctx *c, long n, const float *a, double *b) {
  FZ_LAUNCH(c, "to_f64", k_to_f64, dim3((unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096)), dim3(256), 0, n, a, b);
  FZ_LAUNCHED();
  return FZ_OK;
}

static int case_alloc(fz_case *cs) {
  fz_ctx *c = cs->ctx;
  const bool mixed = cs->precision == FZ_PREC_MIXED;
  for (int q = 0; q < F_N; q++) {
    if (q == F_PN) continue;  // pn is p (see above)
    const long n = q < F_DPSIO ? c->nh : c->ni;
    if (mixed && is_f32_field(q)) {
      FZ_HIP(hipMalloc((void **)&cs->g[q], sizeof(float) * n));
      FZ_HIP(hipMemsetAsync(cs->g[q], 0, sizeof(float) * n, c->st));
    } else {
      FZ_HIP(hipMalloc((void **)&cs->f[q], sizeof(double) * n));
      FZ_HIP(hipMemsetAsync(cs->f[q], 0, sizeof(double) * n, c->st));
    }
  }
  double **hal[5] = {mixed ? nullptr : &cs->psio, &cs->dp, &cs->un, &cs->vn, &cs->wn};
  for (int q = 0; q < 5; q++) {
    if (!hal[q]) continue;
    FZ_HIP(hipMalloc((void **)hal[q], sizeof(double) * c->nh));
    FZ_HIP(hipMemsetAsync(*hal[q], 0, sizeof(double) * c->nh, c->st));
  }
  if (mixed) {
    FZ_HIP(hipMalloc((void **)&cs->psio32, sizeof(float) * c->nh));
    FZ_HIP(hipMemsetAsync(cs->psio32, 0, sizeof(float) * c->nh, c->st));
  }
  FZ_HIP(hipMalloc((void **)&cs->gslots, sizeof(unsigned long long) * FZ_GAMMA_SLOTS));
  FZ_HIP(hipMemsetAsync(cs->gslots, 0, sizeof(unsigned long long) * FZ_GAMMA_SLOTS, c->st));
  FZ_HIP(hipMalloc((void **)&cs->dstc, sizeof(double) * FZ_STC_N));
  FZ_HIP(hipMalloc((void **)&cs->dsc, sizeof(double) * 16));
  FZ_HIP(hipMemsetAsync(cs->dsc, 0, sizeof(double) * 16, c->st));
  const int nz = c->d.n3 + 2;
  std::vector<double> gc(nz), gf(nz);
  const double vol = c->l[0] * c->l[1] * c->l[2];
  for (int k = 0; k < nz; k++) {  // src/main.f90:268-269
    gc[k] = c->dl[0] * c->dl[1] * c->h_dzc[k] / vol;
    gf[k] = c->dl[0] * c->dl[1] * c->h_dzf[k] / vol;
  }
  FZ_HIP(hipMalloc((void **)&cs->gvr_c, sizeof(double) * nz));
  FZ_HIP(hipMalloc((void **)&cs->gvr_f, sizeof(double) * nz));
  FZ_HIP(hipMemcpy(cs->gvr_c, gc.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
  FZ_HIP(hipMemcpy(cs->gvr_f, gf.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
  return FZ_OK;
}

extern "C" int fz_case_destroy(fz_case *cs) {
  if (!cs) return FZ_OK;
  if (cs->ctx) (void)hipStreamSynchronize(cs->ctx->st);
  for (int q = 0; q < F_N; q++) {
    if (cs->f[q]) (void)hipFree(cs->f[q]);
    if (cs->g[q]) (void)hipFree(cs->g[q]);
  }
  if (cs->psio32) (void)hipFree(cs->psio32);
  if (cs->dstc) (void)hipFree(cs->dstc);
  for (int q = 0; q < 2; q++)
    if (cs->graph[q].exec) (void)hipGraphExecDestroy(cs->graph[q].exec);
  double *o[10] = {cs->psio, nullptr, cs->un, cs->vn, cs->wn, cs->dp, cs->dsc, cs->gvr_c, cs->gvr_f,
                   (double *)cs->gslots};
  for (int q = 0; q < 10; q++)
    if (o[q]) (void)hipFree(o[q]);
  if (cs->ctx) fz_ctx_destroy(cs->ctx);
  delete cs;
  return FZ_OK;
}

extern "C" int fz_case_create(fz_case **out, const fz_params *prm, void *hip_stream) {
  return fz_case_create_ex(out, prm, hip_stream, FZ_PREC_FP64, 0);
}

extern "C" int fz_case_create_ex(fz_case **out, const fz_params *prm, void *hip_stream, int precision, int nref) {
  if (!out || !prm) return fz_set_error(FZ_ERR_ARG, "fz_case_create: null argument");
  if (precision != FZ_PREC_FP64 && precision != FZ_PREC_MIXED)
    return fz_set_error(FZ_ERR_ARG, "fz_case_create: unknown precision %d", precision);
  if (precision == FZ_PREC_MIXED) {
    if (nref < -1 || nref > 8)
      return fz_set_error(FZ_ERR_ARG, "fz_case_create: refinements %d not in [-1, 8]", nref);
    if (!prm->is_track_interface || !prm->is_solve_ns)
      return fz_set_error(FZ_ERR_UNSUPPORTED,
                          "mixed precision is implemented for the two-fluid step (interface and Navier-Stokes on)");
  }
  const char *codes[4] = {prm->cbcvel, prm->cbcpre, prm->cbcpsi, prm->cbcnor};
  const int lens[4] = {18, 6, 6, 18};
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < lens[a]; b++)
      if (codes[a][b] != 'P' && codes[a][b] != 'D' && codes[a][b] != 'N')
        return fz_set_error(FZ_ERR_ARG, "fz_case_create: bad boundary code '%c'", codes[a][b]);
  fz_case *cs = new fz_case();
  cs->prm = *prm;
  cs->precision = precision;
  cs->nref = nref;
  const int n3 = prm->ng[2];
  std::vector<double> dzc(n3 + 2), dzf(n3 + 2), zc(n3 + 2), zf(n3 + 2);
  int rc = fz_initgrid(prm->gtype, n3, prm->gr, prm->l[2], dzc.data(), dzf.data(), zc.data(), zf.data());
  if (rc == FZ_OK) rc = fz_ctx_create(&cs->ctx, prm->ng, prm->l, dzc.data(), dzf.data(), zc.data(), zf.data(),
                                      prm->cbcpre, hip_stream);
  if (rc == FZ_OK) rc = fz_initsolver(cs->ctx, prm->bcpre, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                      nullptr, nullptr);
  if (rc == FZ_OK && precision == FZ_PREC_MIXED && !cs->ctx->pois->rhsb_zero)
    rc = fz_set_error(FZ_ERR_UNSUPPORTED, "mixed precision needs homogeneous pressure boundary values");
  if (rc == FZ_OK) rc = case_alloc(cs);
  if (rc != FZ_OK) {
    fz_case_destroy(cs);
    return rc;
  }
  memcpy(cs->cbccur, prm->cbcpsi, 6);  // cbccur = cbcpsi, bccur = bcpre, src/param.f90:172-173
  memcpy(cs->bccur, prm->bcpre, sizeof(double) * 6);
  cs->sc[S_RHO0] = fmin(prm->rho12[0], prm->rho12[1]);  // src/param.f90:174-177
  cs->sc[S_FIRST] = 1.0;
  *out = cs;
  return FZ_OK;
}

extern "C" int fz_case_ctx(fz_case *cs, fz_ctx **ctx) {
  if (!cs || !ctx) return fz_set_error(FZ_ERR_ARG, "fz_case_ctx: null argument");
  *ctx = cs->ctx;
  return FZ_OK;
}

extern "C" int fz_case_field(fz_case *cs, int id, double **dptr, size_t *count) {
  if (!cs || !dptr || id < 0 || id >= F_N) return fz_set_error(FZ_ERR_ARG, "fz_case_field: bad argument");
  if (cs->precision == FZ_PREC_MIXED && is_f32_field(id))
    return fz_set_error(FZ_ERR_ARG, "fz_case_field: field %d is FP32 in this case (fz_case_field_f32)", id);
  *dptr = (id == F_PN) ? cs->f[F_P] : cs->f[id];
  if (count) *count = (size_t)(id < F_DPSIO ? cs->ctx->nh : cs->ctx->ni);
  return FZ_OK;
}

extern "C" int fz_case_field_f32(fz_case *cs, int id, float **dptr, size_t *count) {
  if (!cs || !dptr || id < 0 || id >= F_N) return fz_set_error(FZ_ERR_ARG, "fz_case_field_f32: bad argument");
  if (!cs->g[id]) return fz_set_error(FZ_ERR_ARG, "fz_case_field_f32: field %d is not FP32 in this case", id);
  *dptr = cs->g[id];
  if (count) *count = (size_t)(id < F_DPSIO ? cs->ctx->nh : cs->ctx->ni);
  return FZ_OK;
}

extern "C" int fz_case_precision(fz_case *cs, int *precision, int *nref) {
  if (!cs || !precision) return fz_set_error(FZ_ERR_ARG, "fz_case_precision: null argument");
  *precision = cs->precision;
  if (nref) *nref = cs->nref;
  return FZ_OK;
}

static void drop_graphs(fz_case *cs);

static int sync_gam(fz_case *cs) {
  if (!cs->gam_on_device) return FZ_OK;
  fz_ctx *c = cs->ctx;
  FZ_HIP(hipMemcpyAsync(c->h_small, cs->dsc, sizeof(double), hipMemcpyDeviceToHost, c->st));
  FZ_HIP(hipStreamSynchronize(c->st));
  cs->sc[S_GAM] = c->h_small[0];
  return FZ_OK;
}

extern "C" int fz_case_get_scalars(fz_case *cs, double out[12]) {
  if (!cs || !out) return fz_set_error(FZ_ERR_ARG, "fz_case_get_scalars: null argument");
  FZ_TRY(sync_gam(cs));
  memcpy(out, cs->sc, sizeof(cs->sc));
  return FZ_OK;
}

extern "C" int fz_case_set_scalars(fz_case *cs, const double in[12]) {
  if (!cs || !in) return fz_set_error(FZ_ERR_ARG, "fz_case_set_scalars: null argument");
  memcpy(cs->sc, in, sizeof(cs->sc));
  drop_graphs(cs);  // seps is an argument of the captured kernels
  fz_ctx *c = cs->ctx;
  c->h_small[0] = cs->sc[S_GAM];
  FZ_HIP(hipMemcpyAsync(cs->dsc, c->h_small, sizeof(double), hipMemcpyHostToDevice, c->st));
  FZ_HIP(hipStreamSynchronize(c->st));
  cs->gam_on_device = true;
  return FZ_OK;
}

// the FP64 phase field's range flag (k_acdi_phi, k_youngs): a word of the
// case's device scalars, zero at creation
static unsigned *range_flag(fz_case *cs) { return (unsigned *)(cs->dsc + 15); }

// The directions (bit d) in which the pressure and all three velocity
// components are periodic: there the boundary conditions after the
// correction overwrite every halo cell of u, v, w and p (with is_correc only
// a non-periodic direction keeps the correction's normal velocity in its
// halo, src/bound.f90:46-68), so the correction kernel may leave those halo
// cells out.
static unsigned interior_dirs(const fz_params &P) {
  unsigned m = 0;
  for (int d = 0; d < 3; d++) {
    bool per = P.cbcpre[2 * d] == 'P' && P.cbcpre[2 * d + 1] == 'P';
    for (int comp = 0; comp < 3; comp++)
      per = per && P.cbcvel[6 * comp + 2 * d] == 'P' && P.cbcvel[6 * comp + 2 * d + 1] == 'P';
    if (per) m |= 1u << d;
  }
  return m;
}

// (the Youngs kernel writes the periodic x halos of its four fields when x
// is periodic, and their boundary update leaves those cells alone)
static int phase_geometry(fz_case *cs) {
  fz_ctx *c = cs->ctx;
  const fz_params &P = cs->prm;
  double **f = cs->f;
  const int xp = c->per[0] ? 1 : 0;
  if (cs->precision == FZ_PREC_MIXED) {
    float **g = cs->g;
    if (P.vof) FZ_TRY(fz_vof_phi_f32(c, 1.0 / (2.0 * P.psi_thickness_factor), g[F_PSI], g[F_PHI]));
    else FZ_TRY(fz_acdi_cmpt_phi_f32(c, cs->sc[S_SEPS], g[F_PSI], g[F_PHI]));
    FZ_TRY(fz_youngs_step<float>(c, g[F_PHI], g[F_NX], g[F_NY], g[F_NZ], g[F_KAPPA], xp, nullptr));
    return bc_geometry<float>(c, P, cs->cbccur, cs->bccur, g[F_KAPPA], g[F_NX], g[F_NY], g[F_NZ], xp);
  }
  // (ACDI: the phase field's range flag, cleared by the step's gamma_fin)
  unsigned *rf = P.vof ? nullptr : range_flag(cs);
  if (!P.vof) FZ_TRY(fz_acdi_cmpt_phi_rf(c, cs->sc[S_SEPS], f[F_PSI], f[F_PHI], rf));
  else FZ_TRY(fz_vof_thinc_cmpt_phi(c, 1.0 / (2.0 * P.psi_thickness_factor), f[F_PSI], f[F_PHI]));
  FZ_TRY(fz_youngs_step<double>(c, f[F_PHI], f[F_NX], f[F_NY], f[F_NZ], f[F_KAPPA], xp, rf));
  return bc_geometry<double>(c, P, cs->cbccur, cs->bccur, f[F_KAPPA], f[F_NX], f[F_NY], f[F_NZ], xp);
}

// chkdt + the dt selection of src/main.f90:396-404 and :532-550
static int new_dt(fz_case *cs, double *dt, int *kill) {
  fz_ctx *c = cs->ctx;
  const fz_params &P = cs->prm;
  FZ_TRY(fz_chkdt_adv_dev(c, cs->f[F_U], cs->f[F_V], cs->f[F_W], cs->dsc + 3));
  FZ_HIP(hipMemcpyAsync(c->h_small, cs->dsc, sizeof(double) * 4, hipMemcpyDeviceToHost, c->st));
  FZ_HIP(hipStreamSynchronize(c->st));
  if (cs->gam_on_device) cs->sc[S_GAM] = c->h_small[0];
  const double dti = c->h_small[3];
  const double dt_cfl = fz_chkdt_finish(c, dti, P.is_solve_ns, P.is_track_interface, P.vof, P.mu12, P.rho12,
                                        P.sigma, P.gacc, cs->sc[S_GAM], cs->sc[S_SEPS]);
  cs->sc[S_DTCFL] = dt_cfl;
  double d = fmin(P.cfl * dt_cfl, P.dtmax);
  if (P.dt_f > 0.0) {
    // at start-up a fixed dt above the stable one only warns (src/main.f90:397-403);
    // in the loop it stops the run and keeps the stable dt (:534-541)
    if (P.dt_f > d && kill)
      *kill = 1;
    else
      d = P.dt_f;
  }
  *dt = d;
  return FZ_OK;
}

// the initialisation after the fields are set, src/main.f90:333-407:
// boundary conditions, pn = po = p, phi/normals/curvature, gamma and dt
static int finish_init(fz_case *cs) {
  fz_ctx *c = cs->ctx;
  const fz_params &P = cs->prm;
  double **f = cs->f;
  FZ_TRY(fz_bounduvw(c, P.cbcvel, P.bcvel, 0, f[F_U], f[F_V], f[F_W]));
  FZ_TRY(fz_boundp(c, P.cbcpre, P.bcpre, f[F_P]));
  FZ_TRY(fz_copy(c, f[F_PO], f[F_P], sizeof(double) * c->nh));
  if (cs->precision == FZ_PREC_MIXED) FZ_TRY(fz_boundp_f32(c, P.cbcpsi, P.bcpsi, cs->g[F_PSI]));
  else FZ_TRY(fz_boundp(c, P.cbcpsi, P.bcpsi, f[F_PSI]));
  FZ_TRY(phase_geometry(cs));
  if (!P.vof) {
    FZ_TRY(fz_acdi_gamma_dev(c, P.acdi_gam_factor, P.acdi_gam_min, f[F_U], f[F_V], f[F_W], cs->dsc));
    cs->gam_on_device = true;
  }
  double dt;
  FZ_TRY(new_dt(cs, &dt, nullptr));
  cs->sc[S_DT] = dt;
  cs->sc[S_DTO] = dt;
  cs->sc[S_FIRST] = 1.0;  // the AB2 histories start over (src/rk.f90:76-83)
  FZ_HIP(hipStreamSynchronize(c->st));
  return FZ_OK;
}

static int set_epsilon(fz_case *cs) {
  drop_graphs(cs);
  double seps;
  FZ_TRY(fz_acdi_set_epsilon(cs->ctx, cs->prm.psi_thickness_factor, &seps));  // src/main.f90:301
  cs->sc[S_SEPS] = seps;
  return FZ_OK;
}

extern "C" int fz_case_init(fz_case *cs) {
  if (!cs) return fz_set_error(FZ_ERR_ARG, "fz_case_init: null argument");
  fz_ctx *c = cs->ctx;
  double **f = cs->f;
  FZ_TRY(set_epsilon(cs));
  for (int q = F_U; q <= F_P; q++) FZ_TRY(fz_launch_fill(c, c->nh, 0.0, f[q]));  // initflow 'zer'
  if (cs->precision == FZ_PREC_MIXED) {  // the FP64 initial field, rounded
    FZ_TRY(fz_init2fl(c, &cs->prm, cs->sc[S_SEPS], cs->dp));
    FZ_TRY(to_f32(c, c->nh, cs->dp, cs->g[F_PSI]));
  } else {
    FZ_TRY(fz_init2fl(c, &cs->prm, cs->sc[S_SEPS], f[F_PSI]));
  }
  cs->sc[S_TIME] = 0.0;
  cs->sc[S_ISTEP] = 0.0;
  return finish_init(cs);
}

static const char *kFieldExt[5] = {"u", "v", "w", "p", "psi"};
static const int kFieldId[5] = {F_U, F_V, F_W, F_P, F_PSI};

int fz_load_one(fz_ctx *c, char io, const char *filename, double *p, double *time, int *istep);

extern "C" int fz_case_save(fz_case *cs, const char *prefix) {
  if (!cs || !prefix) return fz_set_error(FZ_ERR_ARG, "fz_case_save: null argument");
  double time = cs->sc[S_TIME];
  int istep = (int)cs->sc[S_ISTEP];
  for (int q = 0; q < 5; q++) {
    const std::string name = std::string(prefix) + "_" + kFieldExt[q] + ".bin";
    double *src = cs->f[kFieldId[q]];
    if (!src) {  // an FP32 field, written in the file's FP64 format
      FZ_TRY(to_f64(cs->ctx, cs->ctx->nh, cs->g[kFieldId[q]], cs->dp));
      src = cs->dp;
    }
    FZ_TRY(fz_load_one(cs->ctx, 'w', name.c_str(), src, &time, &istep));
  }
  return FZ_OK;
}

extern "C" int fz_case_restart(fz_case *cs, const char *prefix) {
  if (!cs || !prefix) return fz_set_error(FZ_ERR_ARG, "fz_case_restart: null argument");
  FZ_TRY(set_epsilon(cs));
  double time = 0.0;
  int istep = 0;
  for (int q = 0; q < 5; q++) {
    const std::string name = std::string(prefix) + "_" + kFieldExt[q] + ".bin";
    double *dst = cs->f[kFieldId[q]] ? cs->f[kFieldId[q]] : cs->dp;
    FZ_TRY(fz_load_one(cs->ctx, 'r', name.c_str(), dst, &time, &istep));
    if (!cs->f[kFieldId[q]]) FZ_TRY(to_f32(cs->ctx, cs->ctx->nh, cs->dp, cs->g[kFieldId[q]]));
  }
  cs->sc[S_TIME] = time;
  cs->sc[S_ISTEP] = (double)istep;
  return finish_init(cs);
}

template <typename T>
static void swap_ptr(T *&a, T *&b) {
  T *t = a;
  a = b;
  b = t;
}

// The device work of one time step (src/main.f90:414-529); pointer
// rotations happen on the host as the work is enqueued. Host
// synchronisation only for a bulk mean that the momentum needs by value.
struct StepArgs {
  double dt, factor1, factor2, factor12, dt_r;
  bool first;
  const double *stc;  // the same scalars in device memory, or null
};

static int enqueue_step(fz_case *cs, const StepArgs &a) {
  fz_ctx *c = cs->ctx;
  const fz_params &P = cs->prm;
  double **f = cs->f;
  float **g = cs->g;
  const bool mixed = cs->precision == FZ_PREC_MIXED;
  const double rho0 = cs->sc[S_RHO0];
  const double dt = a.dt, dt_r = a.dt_r, factor1 = a.factor1, factor2 = a.factor2, factor12 = a.factor12;
  const bool is_first = a.first;
  // x periodic: the correction kernel and the FP32 ACDI kernel write the x
  // halos of the interior rows themselves (the boundary update leaves those
  // cells alone), as does the Youngs kernel of phase_geometry (the FP64 ACDI
  // kernel measured 40, later 50 us slower with them, against 31 us of boundary
  // work)
  const int xp = c->per[0] ? 1 : 0;
  // psio <- psi by rotation: the update writes into the older buffer
  swap_ptr(cs->psio, f[F_PSI]);
  swap_ptr(cs->psio32, g[F_PSI]);
  if (mixed && P.vof) {  // the VOF sequence below on the FP32 phase field
    FZ_TRY(fz_launch_vof_step<float>(c, 1.0 / (2.0 * P.psi_thickness_factor), factor1, factor2, a.stc,
                                     is_first ? 1 : 0,
                                     f[F_U], f[F_V], f[F_W], g[F_NX], g[F_NY], g[F_NZ], g[F_PHI], cs->psio32,
                                     g[F_DPSIO], g[F_PSI], g[F_FX], g[F_FY], g[F_FZ]));
    FZ_TRY(bc_fluxes_psi<float>(c, P, g[F_FX], g[F_FY], g[F_FZ], g[F_PSI]));
    FZ_TRY(phase_geometry(cs));
  } else if (mixed) {  // the same sequence on the FP32 phase field
    FZ_TRY(fz_launch_acdi_step<float>(c, cs->gam_on_device ? cs->dsc : nullptr, cs->sc[S_GAM], cs->sc[S_SEPS],
                                      factor1, factor2, a.stc, is_first ? 1 : 0, f[F_U], f[F_V], f[F_W], g[F_NX], g[F_NY],
                                      g[F_NZ], g[F_PHI], cs->psio32, g[F_DPSIO], g[F_PSI], g[F_FX], g[F_FY],
                                      g[F_FZ], xp));
    FZ_TRY(bc_fluxes_psi<float>(c, P, g[F_FX], g[F_FY], g[F_FZ], g[F_PSI], xp));
    FZ_TRY(phase_geometry(cs));
  } else if (P.is_track_interface) {
    if (!P.vof) {
      // transport, AB2 update and clip in one pass; the history is updated in place
      FZ_TRY(fz_launch_acdi_step<double>(c, cs->gam_on_device ? cs->dsc : nullptr, cs->sc[S_GAM], cs->sc[S_SEPS],
                                           factor1, factor2, a.stc, is_first ? 1 : 0, f[F_U], f[F_V], f[F_W], f[F_NX], f[F_NY], f[F_NZ],
                                 f[F_PHI], cs->psio, f[F_DPSIO], f[F_PSI], f[F_FX], f[F_FY], f[F_FZ], 0));
    } else {
      // vof_thinc_transport_psi, the AB2 update and clip in one pass (src/rk.f90:226-252)
      FZ_TRY(fz_launch_vof_step<double>(c, 1.0 / (2.0 * P.psi_thickness_factor), factor1, factor2, a.stc,
                                        is_first ? 1 : 0, f[F_U], f[F_V], f[F_W], f[F_NX], f[F_NY], f[F_NZ],
                                        f[F_PHI], cs->psio, f[F_DPSIO], f[F_PSI], f[F_FX], f[F_FY], f[F_FZ]));
    }
    FZ_TRY(bc_fluxes_psi<double>(c, P, f[F_FX], f[F_FY], f[F_FZ], f[F_PSI]));
    FZ_TRY(phase_geometry(cs));
  } else {
    FZ_HIP(hipMemcpyAsync(f[F_PSI], cs->psio, sizeof(double) * c->nh, hipMemcpyDeviceToDevice, c->st));
    for (int q = F_FX; q <= F_FZ; q++) FZ_TRY(fz_launch_fill(c, c->nh, 0.0, f[q]));
  }
  if (P.is_solve_ns) {
    double rho_av[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; d++) {
      if (fabs(P.gacc[d]) > 0.0 && c->per[d]) {  // src/main.f90:470-477
        if (mixed)
          FZ_TRY(fz_bulk_mean_dev(c, d + 1, g[F_PSI], P.rho12, d < 2 ? cs->gvr_c : cs->gvr_f, cs->dsc + 4 + d));
        else
          FZ_TRY(fz_bulk_mean_dev(c, d + 1, f[F_PSI], P.rho12, d < 2 ? cs->gvr_c : cs->gvr_f, cs->dsc + 4 + d));
        FZ_HIP(hipMemcpyAsync(c->h_small + 8, cs->dsc + 4 + d, sizeof(double), hipMemcpyDeviceToHost, c->st));
        FZ_HIP(hipStreamSynchronize(c->st));
        rho_av[d] = c->h_small[8];
      }
    }
    // rk, src/rk.f90:72-130, in one pass into the spare velocity buffers
    // (x periodic and no forcing after it: the FP64 kernel writes the x halos
    // of the interior rows, and the boundary update leaves them alone: 1387
    // -> 1397 us, the gather 36 -> 14 us at 256x256x512; the FP32-phase-field
    // kernel measured 1150 -> 1280 us with them and leaves them to the gather)
    const int xm = xp && !P.is_forced_hit && !mixed ? 1 : 0;
    if (mixed)
      FZ_TRY((fz_launch_mom_fused<float, float>(c, P.rho12, P.mu12, factor1, factor2, is_first ? 1 : 0, factor12,
                                               dt_r, a.stc, P.sigma, rho0, P.bforce, P.gacc, rho_av, g[F_FX], g[F_FY],
                                               g[F_FZ], f[F_U], f[F_V], f[F_W], g[F_PSI], cs->psio32, f[F_P],
                                               f[F_PO], g[F_KAPPA], g[F_DUO], g[F_DVO], g[F_DWO], cs->un, cs->vn,
                                               cs->wn, xm)));
    else
      FZ_TRY((fz_launch_mom_fused<double, double>(c, P.rho12, P.mu12, factor1, factor2, is_first ? 1 : 0, factor12,
                                                 dt_r, a.stc, P.sigma, rho0, P.bforce, P.gacc, rho_av, f[F_FX], f[F_FY],
                                                 f[F_FZ], f[F_U], f[F_V], f[F_W], f[F_PSI], cs->psio, f[F_P],
                                                 f[F_PO], f[F_KAPPA], f[F_DUO], f[F_DVO], f[F_DWO], cs->un, cs->vn,
                                                 cs->wn, xm)));
    swap_ptr(f[F_U], cs->un);
    swap_ptr(f[F_V], cs->vn);
    swap_ptr(f[F_W], cs->wn);
    if (P.is_forced_hit) FZ_TRY(fz_launch_forcing(c, 2, 0.5, dt, a.stc, f[F_U], f[F_V], f[F_W]));
    {
      const FzBcItem it[1] = {{1, P.cbcvel, P.bcvel, 0, {f[F_U], f[F_V], f[F_W]}, xm}};
      FZ_TRY(fz_bounds_group<double>(c, 1, it));
    }
    // projection on the correction buffer dp, src/main.f90:488-502
    // (nref = -1: the mixed step with the FP64 projection of the plain step)
    if (mixed && cs->nref >= 0)
      FZ_TRY(fz_projection_rhs_solve_mixed(c, 1.0 / dt, rho0, f[F_U], f[F_V], f[F_W], cs->dp, cs->nref, a.stc));
    else FZ_TRY(fz_projection_rhs_solve(c, 1.0 / dt, rho0, f[F_U], f[F_V], f[F_W], cs->dp, a.stc));
    FZ_TRY(fz_boundp(c, P.cbcpre, P.bcpre, cs->dp));
    // correc into the spare velocity buffers, updatep into the buffer of the
    // previous pressure and the speeds of acdi_set_gamma, in one pass
    FZ_TRY(fz_launch_correc_step(c, rho0, dt, a.stc, cs->dp, f[F_P], f[F_U], f[F_V], f[F_W], cs->un, cs->vn, cs->wn,
                                 f[F_PO], P.vof ? nullptr : cs->gslots, xp, interior_dirs(P)));
    swap_ptr(f[F_U], cs->un);
    swap_ptr(f[F_V], cs->vn);
    swap_ptr(f[F_W], cs->wn);
    swap_ptr(f[F_P], f[F_PO]);  // p <- p + dp, po <- p of the step before (= pn)
    {
      const FzBcItem it[2] = {{1, P.cbcvel, P.bcvel, 1, {f[F_U], f[F_V], f[F_W]}, xp},
                              {0, P.cbcpre, P.bcpre, 0, {f[F_P], nullptr, nullptr}, xp}};
      FZ_TRY(fz_bounds_group<double>(c, 2, it));
    }
  } else {
    FZ_HIP(hipMemcpyAsync(f[F_PO], f[F_P], sizeof(double) * c->nh, hipMemcpyDeviceToDevice, c->st));
  }
  if (!P.vof) {
    if (P.is_solve_ns)
      FZ_TRY(fz_launch_gamma_fin(c, cs->gslots, P.acdi_gam_factor, P.acdi_gam_min, cs->dsc, range_flag(cs)));
    else FZ_TRY(fz_acdi_gamma_dev(c, P.acdi_gam_factor, P.acdi_gam_min, f[F_U], f[F_V], f[F_W], cs->dsc));
  }
  return FZ_OK;
}

// Steps after the first replay a captured graph of enqueue_step when the
// case qualifies: every kernel argument that changes between steps is read
// from device memory (gamma, and dt with its AB2 factors, also in the
// THINC/QQ step and the forcing), the step has no host synchronisation and
// no memory-copy nodes, and per-launch profiling is off. The pointer
// rotations give two graphs, one per parity, keyed by the buffers that
// hold p and psi when the step starts.
static bool graph_ok(const fz_case *cs) {
  const fz_params &P = cs->prm;
  if (!cs->use_graphs || cs->ctx->prof_on) return false;
  if (!P.is_track_interface || !P.is_solve_ns) return false;
  for (int d = 0; d < 3; d++)
    if (fabs(P.gacc[d]) > 0.0 && cs->ctx->per[d]) return false;  // bulk mean read back by the host
  return fz_projection_device_scalars_ok(cs->ctx, cs->precision == FZ_PREC_MIXED && cs->nref >= 0) != 0;
}

__global__ void k_set_stc(double *__restrict__ stc, double dt, double dti, double f1, double f2, double f12,
                          double dtr) {
  stc[FZ_STC_DT] = dt;
  stc[FZ_STC_DTI] = dti;
  stc[FZ_STC_F1] = f1;
  stc[FZ_STC_F2] = f2;
  stc[FZ_STC_F12] = f12;
  stc[FZ_STC_DTR] = dtr;
}

static int set_stc(fz_case *cs, const StepArgs &a) {
  const double v[FZ_STC_N] = {a.dt, 1.0 / a.dt, a.factor1, a.factor2, a.factor12, a.dt_r};
  if (cs->stc_valid && memcmp(v, cs->stc_host, sizeof(v)) == 0) return FZ_OK;
  FZ_LAUNCH(cs->ctx, "set_stc", k_set_stc, dim3(1), dim3(1), 0, cs->dstc, v[0], v[1], v[2], v[3], v[4], v[5]);
  FZ_LAUNCHED();
  memcpy(cs->stc_host, v, sizeof(v));
  cs->stc_valid = true;
  return FZ_OK;
}

static void drop_graphs(fz_case *cs) {
  for (int q = 0; q < 2; q++) {
    if (cs->graph[q].exec) (void)hipGraphExecDestroy(cs->graph[q].exec);
    cs->graph[q] = fz_case::StepGraph();
  }
}

// one step through a captured graph (captured on the first use per parity)
static int graph_step(fz_case *cs, const StepArgs &a) {
  fz_ctx *c = cs->ctx;
  const void *k0 = cs->f[F_P];
  const void *k1 = cs->precision == FZ_PREC_MIXED ? (const void *)cs->g[F_PSI] : (const void *)cs->f[F_PSI];
  int slot = -1;
  for (int q = 0; q < 2; q++)
    if (cs->graph[q].exec && cs->graph[q].key[0] == k0 && cs->graph[q].key[1] == k1) slot = q;
  if (slot < 0 && cs->capture_blocked) return enqueue_step(cs, a);  // the same work, launched directly
  if (slot < 0) {
    slot = cs->graph[0].exec ? 1 : 0;
    if (cs->graph[slot].exec) drop_graphs(cs), slot = 0;
    // capture from a copy of the host-side pointer state; the launches
    // then rotate the real state exactly as the replay will
    FZ_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step(cs, a);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c->st, &gr);
    if (rc != FZ_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fz_set_error(FZ_ERR_HIP, "step graph capture: %s", hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess) return fz_set_error(FZ_ERR_HIP, "step graph instantiate: %s", hipGetErrorString(ei));
    cs->graph[slot].exec = ex;
    cs->graph[slot].key[0] = k0;
    cs->graph[slot].key[1] = k1;
    // the capture recorded the work without running it
    FZ_HIP(hipGraphLaunch(ex, c->st));
    return FZ_OK;
  }
  // replay: rotate the host-side pointers as enqueue_step does
  FZ_HIP(hipGraphLaunch(cs->graph[slot].exec, c->st));
  double **f = cs->f;
  swap_ptr(cs->psio, f[F_PSI]);
  swap_ptr(cs->psio32, cs->g[F_PSI]);
  swap_ptr(f[F_P], f[F_PO]);
  return FZ_OK;
}

const fz_params *fz_case_prm(const fz_case *cs) { return &cs->prm; }
void fz_case_block_capture(fz_case *cs, bool on) { cs->capture_blocked = on; }
void fz_case_drop_graphs(fz_case *cs) { drop_graphs(cs); }

extern "C" int fz_case_set_graphs(fz_case *cs, int on) {
  if (!cs) return fz_set_error(FZ_ERR_ARG, "fz_case_set_graphs: null case");
  if (!on) drop_graphs(cs);
  cs->use_graphs = on != 0;
  return FZ_OK;
}

// The checks of src/main.f90:530-562 at the end of a step whose number is a
// multiple of icheck: chkdt and chkdiv enqueued with one read-back of their
// device scalars (gamma, divergence sum and maximum, 1/dt) into h_small[16..19]
static int check_enqueue(fz_case *cs) {
  fz_ctx *c = cs->ctx;
  double **f = cs->f;
  FZ_TRY(fz_chkdt_adv_dev(c, f[F_U], f[F_V], f[F_W], cs->dsc + 3));
  FZ_TRY(fz_chkdiv_dev(c, f[F_U], f[F_V], f[F_W], cs->dsc + 1));
  FZ_HIP(hipMemcpyAsync(c->h_small + 16, cs->dsc, sizeof(double) * 4, hipMemcpyDeviceToHost, c->st));
  cs->check_pending = true;
  return FZ_OK;
}

// the host side of the checks once the read-back has landed: the new dt
// (src/main.f90:532-550) and the stop conditions (small dt, divergence)
static void check_finish(fz_case *cs, int *kill) {
  fz_ctx *c = cs->ctx;
  const fz_params &P = cs->prm;
  const double *h = c->h_small + 16;
  if (cs->gam_on_device) cs->sc[S_GAM] = h[0];
  const double dt_cfl = fz_chkdt_finish(c, h[3], P.is_solve_ns, P.is_track_interface, P.vof, P.mu12, P.rho12,
                                        P.sigma, P.gacc, cs->sc[S_GAM], cs->sc[S_SEPS]);
  cs->sc[S_DTCFL] = dt_cfl;
  double d = fmin(P.cfl * dt_cfl, P.dtmax);
  if (P.dt_f > 0.0) {
    // a fixed dt above the stable one stops the run and keeps the stable dt
    if (P.dt_f > d)
      *kill = 1;
    else
      d = P.dt_f;
  }
  if (dt_cfl < FZ_EPS * 1e7) *kill = 1;  // small, src/param.f90:21
  cs->sc[S_DT] = d;
  cs->sc[S_DIVTOT] = h[1];
  cs->sc[S_DIVMAX] = h[2];
  if (cs->sc[S_DIVMAX] > FZ_EPS * 1e7 || std::isnan(cs->sc[S_DIVTOT])) *kill = 1;
  cs->check_pending = false;
}

// one step enqueued (src/main.f90:414-529), with its checks when due
static int step_enqueue(fz_case *cs) {
  const fz_params &P = cs->prm;
  StepArgs a;
  a.dt = cs->sc[S_DT];
  const double dto = cs->sc[S_DTO];
  a.first = cs->sc[S_FIRST] != 0.0;
  cs->sc[S_ISTEP] += 1.0;
  cs->sc[S_TIME] += a.dt;
  a.dt_r = a.dt / dto;  // src/main.f90:422-424, src/rk.f90:53-55
  const double tm1 = 1.0 + 0.5 * a.dt_r, tm2 = -0.5 * a.dt_r;
  a.factor1 = tm1 * a.dt;
  a.factor2 = tm2 * a.dt;
  a.factor12 = a.factor1 + a.factor2;
  const bool use_graph = !a.first && graph_ok(cs);
  a.stc = nullptr;
  if (use_graph) {
    a.stc = cs->dstc;
    FZ_TRY(set_stc(cs, a));
    FZ_TRY(graph_step(cs, a));
  } else {
    FZ_TRY(enqueue_step(cs, a));
  }
  cs->sc[S_FIRST] = 0.0;
  cs->sc[S_DTO] = a.dt;  // dto = dt, src/main.f90:523
  if (!P.vof) cs->gam_on_device = true;
  const int istep = (int)cs->sc[S_ISTEP];
  if (P.icheck > 0 && istep % P.icheck == 0) FZ_TRY(check_enqueue(cs));  // src/main.f90:530-562
  return FZ_OK;
}

extern "C" int fz_case_step(fz_case *cs, int nsteps, int *killed) {
  if (!cs || nsteps < 0) return fz_set_error(FZ_ERR_ARG, "fz_case_step: bad argument");
  if (cs->check_pending) return fz_set_error(FZ_ERR_STATE, "fz_case_step: fz_case_step_end has not been called");
  int kill = 0;
  for (int s = 0; s < nsteps && !kill; s++) {
    FZ_TRY(step_enqueue(cs));
    if (cs->check_pending) {
      FZ_HIP(hipStreamSynchronize(cs->ctx->st));
      check_finish(cs, &kill);
    }
  }
  if (killed) *killed = kill;
  return FZ_OK;
}

// fz_case_step split for callers that step many cases: begin enqueues one
// step and returns without waiting; end, once the step's work is done (or
// after waiting for it), reads the check results that set the next dt.
extern "C" int fz_case_step_begin(fz_case *cs, int *check_pending) {
  if (!cs) return fz_set_error(FZ_ERR_ARG, "fz_case_step_begin: null case");
  if (cs->check_pending) return fz_set_error(FZ_ERR_STATE, "fz_case_step_begin: the previous step has not ended");
  FZ_TRY(step_enqueue(cs));
  if (check_pending) *check_pending = cs->check_pending ? 1 : 0;
  return FZ_OK;
}

extern "C" int fz_case_step_end(fz_case *cs, int *killed) {
  if (!cs) return fz_set_error(FZ_ERR_ARG, "fz_case_step_end: null case");
  int kill = 0;
  if (cs->check_pending) {
    FZ_HIP(hipStreamSynchronize(cs->ctx->st));
    check_finish(cs, &kill);
  }
  if (killed) *killed = kill;
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// rk / rk_2fl with an explicit history object
// ---------------------------------------------------------------------------
struct fz_rk_state {
  fz_ctx *ctx;
  double *dpsi = nullptr, *dpsio = nullptr, *du = nullptr, *dv = nullptr, *dw = nullptr, *duo = nullptr,
         *dvo = nullptr, *dwo = nullptr;
  bool first_mom = true, first_psi = true;
};

extern "C" int fz_rk_state_create(fz_ctx *c, fz_rk_state **st) {
  if (!c || !st) return fz_set_error(FZ_ERR_ARG, "fz_rk_state_create: null argument");
  fz_rk_state *s = new fz_rk_state();
  s->ctx = c;
  double **b[8] = {&s->dpsi, &s->dpsio, &s->du, &s->dv, &s->dw, &s->duo, &s->dvo, &s->dwo};
  for (int q = 0; q < 8; q++) {
    hipError_t e = hipMalloc((void **)b[q], sizeof(double) * c->ni);
    if (e != hipSuccess) {
      fz_rk_state_destroy(s);
      return fz_set_error(FZ_ERR_HIP, "fz_rk_state_create: %s", hipGetErrorString(e));
    }
  }
  *st = s;
  return FZ_OK;
}

extern "C" int fz_rk_state_destroy(fz_rk_state *s) {
  if (!s) return FZ_OK;
  if (s->ctx) (void)hipStreamSynchronize(s->ctx->st);
  double *b[8] = {s->dpsi, s->dpsio, s->du, s->dv, s->dw, s->duo, s->dvo, s->dwo};
  for (int q = 0; q < 8; q++)
    if (b[q]) (void)hipFree(b[q]);
  delete s;
  return FZ_OK;
}

extern "C" int fz_rk(fz_ctx *c, fz_rk_state *s, const double rkpar[2], double dt, double dt_r, const double bforce[3],
                     const double gacc[3], double sigma, const double rho_av[3], const double rho12[2],
                     const double mu12[2], const double beta12[2], double rho0, const double *psi,
                     const double *kappa, const double *p, const double *pn, const double *po, const double *psio,
                     const double *psiflx_x, const double *psiflx_y, const double *psiflx_z, double *u, double *v,
                     double *w) {
  if (!c || !s || !rkpar || s->ctx != c) return fz_set_error(FZ_ERR_ARG, "fz_rk: bad argument");
  const double factor1 = rkpar[0] * dt, factor2 = rkpar[1] * dt, factor12 = factor1 + factor2;
  const long bi = sizeof(double) * c->ni;
  FZ_TRY(fz_mom_xyz_ad(c, rho12, mu12, psiflx_x, psiflx_y, psiflx_z, u, v, w, psi, s->du, s->dv, s->dw));
  if (s->first_mom) {
    FZ_HIP(hipMemcpyAsync(s->duo, s->du, bi, hipMemcpyDeviceToDevice, c->st));
    FZ_HIP(hipMemcpyAsync(s->dvo, s->dv, bi, hipMemcpyDeviceToDevice, c->st));
    FZ_HIP(hipMemcpyAsync(s->dwo, s->dw, bi, hipMemcpyDeviceToDevice, c->st));
    s->first_mom = false;
  }
  FZ_TRY(fz_launch_rk_mom(c, factor1, factor2, rho12, psio, psi, s->du, s->dv, s->dw, s->duo, s->dvo, s->dwo, u, v, w));
  swap_ptr(s->du, s->duo);
  swap_ptr(s->dv, s->dvo);
  swap_ptr(s->dw, s->dwo);
  FZ_TRY(fz_mom_xyz_oth(c, dt_r, rho12, beta12, bforce, gacc, sigma, rho0, rho_av, p, psi, kappa, pn, po, s->du, s->dv,
                        s->dw));
  return fz_launch_rk_oth(c, factor12, rho12, psi, s->du, s->dv, s->dw, u, v, w);
}

extern "C" int fz_rk_2fl(fz_ctx *c, fz_rk_state *s, const double rkpar[2], double dt, double gam, double seps,
                         double beta, int vof, const double *u, const double *v, const double *w, const double *normx,
                         const double *normy, const double *normz, const double *phi, double *psi, double *psiflx_x,
                         double *psiflx_y, double *psiflx_z) {
  if (!c || !s || !rkpar || s->ctx != c) return fz_set_error(FZ_ERR_ARG, "fz_rk_2fl: bad argument");
  const double factor1 = rkpar[0] * dt, factor2 = rkpar[1] * dt;
  if (!vof) FZ_TRY(fz_acdi_transport_pf(c, gam, seps, u, v, w, normx, normy, normz, phi, psi, s->dpsi, psiflx_x,
                                        psiflx_y, psiflx_z));
  else FZ_TRY(fz_vof_thinc_transport_psi(c, beta, u, v, w, normx, normy, normz, phi, psi, s->dpsi, psiflx_x, psiflx_y,
                                         psiflx_z));
  if (s->first_psi) {
    FZ_HIP(hipMemcpyAsync(s->dpsio, s->dpsi, sizeof(double) * c->ni, hipMemcpyDeviceToDevice, c->st));
    s->first_psi = false;
  }
  FZ_TRY(fz_launch_rk_psi(c, factor1, factor2, s->dpsi, s->dpsio, psi));
  const double mm[2] = {0.0, 1.0};
  FZ_TRY(fz_clip_field(c, mm, psi));
  swap_ptr(s->dpsi, s->dpsio);
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// Batched steps of same-shaped cases (no reference counterpart): the step of
// every case of the batch in one launch per stage. Each stage's kernel takes
// its block's case from the block index (the cases' grids stacked in z) and
// reads that case's buffers and scalars from a device table; the pointer
// rotations of a step become a per-case state in device memory that the
// step's last launch flips. The arithmetic is each case's own step's, so
// every case's fields equal its unbatched run bit for bit.
// ---------------------------------------------------------------------------
template <typename T>
int fz_batch_acdi_step(fz_ctx *c, fz_bref br, int ncase);
template <typename T>
int fz_batch_phase_geometry(fz_ctx *c, fz_bref br, int ncase);
template <typename PT, typename HT>
int fz_batch_mom(fz_ctx *c, fz_bref br, int ncase);
int fz_batch_correc_gamma(fz_ctx *c, fz_bref br, int ncase, bool gamma);
int fz_batch_fillps(fz_ctx *c, fz_bref br, int ncase);
int fz_batch_flip(fz_ctx *c, int *par, int ncase);
template <typename T>
int fz_bounds_group_batch(fz_ctx *c, int n, const FzBcItem *items, const int *roles, fz_bref br, int ncase);

struct fz_case_batch {
  std::vector<fz_case *> cs;
  fz_bcase *dtab = nullptr;  // the role table, one entry per case
  int *dpar = nullptr;       // each case's pointer state
  void **dsolve = nullptr;   // the correction buffers, work arrays and pivots of the cases
  std::vector<hipEvent_t> ev;
  std::vector<hipStream_t> own;  // the cases' own streams while the batch steps
  bool mixed = false;
  bool pending = false;  // a step begun and not ended
};

// the buffer of every role in the case's current pointer state (what
// enqueue_step hands to the kernels of the next step)
static void case_roles(const fz_case *cs, void *at[BR_N]) {
  const bool mixed = cs->precision == FZ_PREC_MIXED;
  auto ph = [&](int id) -> void * { return mixed ? (void *)cs->g[id] : (void *)cs->f[id]; };
  at[BR_U] = cs->f[F_U];
  at[BR_V] = cs->f[F_V];
  at[BR_W] = cs->f[F_W];
  at[BR_UN] = cs->un;
  at[BR_VN] = cs->vn;
  at[BR_WN] = cs->wn;
  at[BR_P] = cs->f[F_P];
  at[BR_PO] = cs->f[F_PO];
  at[BR_DP] = cs->dp;
  at[BR_PSI_OLD] = ph(F_PSI);  // the step starts with psio <- psi by rotation
  at[BR_PSI_NEW] = mixed ? (void *)cs->psio32 : (void *)cs->psio;
  at[BR_PHI] = ph(F_PHI);
  at[BR_NX] = ph(F_NX);
  at[BR_NY] = ph(F_NY);
  at[BR_NZ] = ph(F_NZ);
  at[BR_KAPPA] = ph(F_KAPPA);
  at[BR_FX] = ph(F_FX);
  at[BR_FY] = ph(F_FY);
  at[BR_FZ] = ph(F_FZ);
  at[BR_DPSIO] = ph(F_DPSIO);
  at[BR_DUO] = ph(F_DUO);
  at[BR_DVO] = ph(F_DVO);
  at[BR_DWO] = ph(F_DWO);
}

// why a case cannot be batched with the first one (null: it can)
static const char *batch_refusal(const fz_case *a, const fz_case *b) {
  const fz_params &P = b->prm, &Q = a->prm;
  if (b->sc[S_FIRST] != 0.0) return "its first step has not been taken";
  if (b->check_pending) return "a step has not ended";
  if (b->ctx->prof_on) return "per-launch profiling is on";
  if (P.vof || P.is_forced_hit) return "THINC/QQ and forced cases are not batched";
  if (!P.is_track_interface || !P.is_solve_ns) return "it does not solve both the flow and the interface";
  for (int d = 0; d < 3; d++)
    if (fabs(P.gacc[d]) > 0.0 && b->ctx->per[d]) return "its bulk density mean is read back by the host";
  if (b->precision == FZ_PREC_MIXED && b->nref >= 0) return "mixed precision with an FP32 projection is not batched";
  if (!b->ctx->pois || b->ctx->pois->periodic_z) return "a periodic z is not batched";
  if (!b->ctx->pois->rhsb_zero) return "non-homogeneous pressure boundary values";
  if (!b->gam_on_device) return "gamma is not on the device";
  if (b->precision != a->precision || b->nref != a->nref) return "another precision";
  if (memcmp(P.ng, Q.ng, sizeof(P.ng)) || memcmp(P.l, Q.l, sizeof(P.l)) || P.gtype != Q.gtype || P.gr != Q.gr)
    return "another grid";
  if (memcmp(P.cbcvel, Q.cbcvel, sizeof(P.cbcvel)) || memcmp(P.bcvel, Q.bcvel, sizeof(P.bcvel)) ||
      memcmp(P.cbcpre, Q.cbcpre, sizeof(P.cbcpre)) || memcmp(P.bcpre, Q.bcpre, sizeof(P.bcpre)) ||
      memcmp(P.cbcpsi, Q.cbcpsi, sizeof(P.cbcpsi)) || memcmp(P.bcpsi, Q.bcpsi, sizeof(P.bcpsi)) ||
      memcmp(P.cbcnor, Q.cbcnor, sizeof(P.cbcnor)) || memcmp(P.bcnor, Q.bcnor, sizeof(P.bcnor)) ||
      memcmp(b->cbccur, a->cbccur, sizeof(a->cbccur)) || memcmp(b->bccur, a->bccur, sizeof(a->bccur)))
    return "other boundary conditions";
  return nullptr;
}

const char *fz_case_batch_refusal(const fz_case *first, const fz_case *cs) { return batch_refusal(first, cs); }

extern "C" int fz_case_batch_destroy(fz_case_batch *b) {
  if (!b) return FZ_OK;
  if (!b->cs.empty()) (void)hipStreamSynchronize(b->cs[0]->ctx->st);
  if (b->dtab) (void)hipFree(b->dtab);
  if (b->dpar) (void)hipFree(b->dpar);
  if (b->dsolve) (void)hipFree(b->dsolve);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  delete b;
  return FZ_OK;
}

extern "C" int fz_case_batch_create(fz_case *const *cases, int n, fz_case_batch **out) {
  if (!cases || n < 1 || !out) return fz_set_error(FZ_ERR_ARG, "fz_case_batch_create: bad argument");
  if (n > 65535) return fz_set_error(FZ_ERR_ARG, "fz_case_batch_create: at most 65535 cases");
  for (int q = 0; q < n; q++) {
    if (!cases[q]) return fz_set_error(FZ_ERR_ARG, "fz_case_batch_create: null case %d", q);
    for (int p = 0; p < q; p++)
      if (cases[p] == cases[q]) return fz_set_error(FZ_ERR_ARG, "fz_case_batch_create: case %d given twice", q);
    const char *why = batch_refusal(cases[0], cases[q]);
    if (why) return fz_set_error(FZ_ERR_UNSUPPORTED, "fz_case_batch_create: case %d: %s", q, why);
  }
  fz_case_batch *b = new fz_case_batch();
  b->cs.assign(cases, cases + n);
  b->mixed = cases[0]->precision == FZ_PREC_MIXED;
  std::vector<fz_bcase> tab(n);
  std::vector<void *> sol(3 * (size_t)n);
  for (int q = 0; q < n; q++) {
    const fz_case *cs = cases[q];
    const fz_params &P = cs->prm;
    fz_bcase &e = tab[q];
    memset(&e, 0, sizeof(e));
    case_roles(cs, e.at[0]);
    memcpy(e.at[1], e.at[0], sizeof(e.at[0]));  // one step later: psi and p rotated
    std::swap(e.at[1][BR_PSI_OLD], e.at[1][BR_PSI_NEW]);
    std::swap(e.at[1][BR_P], e.at[1][BR_PO]);
    e.stc = cs->dstc;
    e.dsc = cs->dsc;
    e.gslots = cs->gslots;
    e.rho = P.rho12[1];
    e.drho = P.rho12[0] - P.rho12[1];
    e.mu = P.mu12[1];
    e.dmu = P.mu12[0] - P.mu12[1];
    e.surf_factor = P.sigma * 2.0 / (P.rho12[0] + P.rho12[1]);
    e.rho0 = cs->sc[S_RHO0];
    e.seps = cs->sc[S_SEPS];
    e.gam_factor = P.acdi_gam_factor;
    e.gam_min = P.acdi_gam_min;
    for (int d = 0; d < 3; d++) {
      e.bforce[d] = P.bforce[d];
      e.gacc[d] = P.gacc[d];
    }
    sol[q] = cs->dp;
    sol[n + q] = cs->ctx->pois->work;
    sol[2 * n + q] = cs->ctx->pois->piv;
  }
  int rc = FZ_OK;
  auto hip = [&](hipError_t e, const char *what) {
    if (rc == FZ_OK && e != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "fz_case_batch_create %s: %s", what, hipGetErrorString(e));
  };
  hip(hipMalloc((void **)&b->dtab, sizeof(fz_bcase) * n), "table");
  if (rc == FZ_OK) hip(hipMalloc((void **)&b->dpar, sizeof(int) * n), "state");
  if (rc == FZ_OK) hip(hipMalloc((void **)&b->dsolve, sizeof(void *) * 3 * n), "solve table");
  std::vector<int> par(n, 0);
  if (rc == FZ_OK) hip(hipMemcpy(b->dtab, tab.data(), sizeof(fz_bcase) * n, hipMemcpyHostToDevice), "upload");
  if (rc == FZ_OK) hip(hipMemcpy(b->dpar, par.data(), sizeof(int) * n, hipMemcpyHostToDevice), "upload");
  if (rc == FZ_OK) hip(hipMemcpy(b->dsolve, sol.data(), sizeof(void *) * 3 * n, hipMemcpyHostToDevice), "upload");
  b->ev.assign(n, nullptr);
  for (int q = 0; q < n && rc == FZ_OK; q++) hip(hipEventCreateWithFlags(&b->ev[q], hipEventDisableTiming), "event");
  if (rc != FZ_OK) {
    fz_case_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FZ_OK;
}

// one step of every case of the batch, enqueued on the first case's stream
static int batch_enqueue(fz_case_batch *b) {
  const int n = (int)b->cs.size();
  fz_case *c0s = b->cs[0];
  fz_ctx *c0 = c0s->ctx;
  const fz_params &P = c0s->prm;
  const fz_bref br = {b->dtab, b->dpar, 0};
  for (fz_case *cs : b->cs) {  // the host side of step_enqueue, per case
    StepArgs a;
    a.dt = cs->sc[S_DT];
    a.first = false;
    cs->sc[S_ISTEP] += 1.0;
    cs->sc[S_TIME] += a.dt;
    a.dt_r = a.dt / cs->sc[S_DTO];
    const double tm1 = 1.0 + 0.5 * a.dt_r, tm2 = -0.5 * a.dt_r;
    a.factor1 = tm1 * a.dt;
    a.factor2 = tm2 * a.dt;
    a.factor12 = a.factor1 + a.factor2;
    a.stc = cs->dstc;
    FZ_TRY(set_stc(cs, a));
    cs->sc[S_DTO] = a.dt;
  }
  // rk_2fl with acdi_transport_pf, the boundary updates, phi and the normals
  const FzBcItem fl[2] = {{1, P.cbcvel, P.bcvel, 0, {nullptr, nullptr, nullptr}},
                          {0, P.cbcpsi, P.bcpsi, 0, {nullptr, nullptr, nullptr}}};
  const int fl_roles[6] = {BR_FX, BR_FY, BR_FZ, BR_PSI_NEW, 0, 0};
  const FzBcItem ge[4] = {{0, c0s->cbccur, c0s->bccur, 0, {nullptr, nullptr, nullptr}},
                          {0, P.cbcnor + 0, P.bcnor + 0, 0, {nullptr, nullptr, nullptr}},
                          {0, P.cbcnor + 6, P.bcnor + 6, 0, {nullptr, nullptr, nullptr}},
                          {0, P.cbcnor + 12, P.bcnor + 12, 0, {nullptr, nullptr, nullptr}}};
  const int ge_roles[12] = {BR_KAPPA, 0, 0, BR_NX, 0, 0, BR_NY, 0, 0, BR_NZ, 0, 0};
  if (b->mixed) {
    FZ_TRY(fz_batch_acdi_step<float>(c0, br, n));
    FZ_TRY(fz_bounds_group_batch<float>(c0, 2, fl, fl_roles, br, n));
    FZ_TRY(fz_batch_phase_geometry<float>(c0, br, n));
    FZ_TRY(fz_bounds_group_batch<float>(c0, 4, ge, ge_roles, br, n));
    FZ_TRY((fz_batch_mom<float, float>(c0, br, n)));
  } else {
    FZ_TRY(fz_batch_acdi_step<double>(c0, br, n));
    FZ_TRY(fz_bounds_group_batch<double>(c0, 2, fl, fl_roles, br, n));
    FZ_TRY(fz_batch_phase_geometry<double>(c0, br, n));
    FZ_TRY(fz_bounds_group_batch<double>(c0, 4, ge, ge_roles, br, n));
    FZ_TRY((fz_batch_mom<double, double>(c0, br, n)));
  }
  // rk's result is in the spare buffers: bounduvw, fillps, the solve into dp
  // and boundp, then correc/updatep/gamma speeds into the first buffers
  const FzBcItem uv[1] = {{1, P.cbcvel, P.bcvel, 0, {nullptr, nullptr, nullptr}}};
  const int uv_roles[3] = {BR_UN, BR_VN, BR_WN};
  FZ_TRY(fz_bounds_group_batch<double>(c0, 1, uv, uv_roles, br, n));
  FZ_TRY(fz_batch_fillps(c0, br, n));
  FZ_TRY(fz_batch_solve(c0, b->dsolve, n));
  const FzBcItem pr[1] = {{0, P.cbcpre, P.bcpre, 0, {nullptr, nullptr, nullptr}}};
  const int pr_roles[3] = {BR_DP, 0, 0};
  FZ_TRY(fz_bounds_group_batch<double>(c0, 1, pr, pr_roles, br, n));
  FZ_TRY(fz_batch_correc_gamma(c0, br, n, true));
  const FzBcItem co[2] = {{1, P.cbcvel, P.bcvel, 1, {nullptr, nullptr, nullptr}},
                          {0, P.cbcpre, P.bcpre, 0, {nullptr, nullptr, nullptr}}};
  const int co_roles[6] = {BR_U, BR_V, BR_W, BR_PO, 0, 0};
  FZ_TRY(fz_bounds_group_batch<double>(c0, 2, co, co_roles, br, n));
  FZ_TRY(fz_batch_flip(c0, b->dpar, n));
  for (fz_case *cs : b->cs) {  // the rotations the kernels applied through the flipped state
    swap_ptr(cs->psio, cs->f[F_PSI]);
    swap_ptr(cs->psio32, cs->g[F_PSI]);
    swap_ptr(cs->f[F_P], cs->f[F_PO]);
  }
  return FZ_OK;
}

// the batch's stream is the first case's; every other case's context uses it
// while the batch steps (its own launches: dt scalars, checks), after the
// case's earlier work on its own stream
static int batch_enter(fz_case_batch *b) {
  const int n = (int)b->cs.size();
  hipStream_t st = b->cs[0]->ctx->st;
  b->own.assign(n, nullptr);
  for (int q = 0; q < n; q++) {
    b->own[q] = b->cs[q]->ctx->st;
    if (q > 0) {
      FZ_HIP(hipEventRecord(b->ev[q], b->own[q]));
      FZ_HIP(hipStreamWaitEvent(st, b->ev[q], 0));
    }
  }
  for (int q = 1; q < n; q++) b->cs[q]->ctx->st = st;
  return FZ_OK;
}
// each case's later work on its own stream after the batch's
static void batch_leave(fz_case_batch *b) {
  const int n = (int)b->cs.size();
  for (int q = 1; q < n; q++) b->cs[q]->ctx->st = b->own[q];
  if (hipEventRecord(b->ev[0], b->cs[0]->ctx->st) == hipSuccess)
    for (int q = 1; q < n; q++) (void)hipStreamWaitEvent(b->own[q], b->ev[0], 0);
}

// the checks enqueued by a step read back: the next dt and the stop conditions
static int batch_checks(fz_case_batch *b, int *killed, bool *stop) {
  bool pending = false;
  for (fz_case *cs : b->cs) pending = pending || cs->check_pending;
  if (!pending) return FZ_OK;
  FZ_HIP(hipStreamSynchronize(b->cs[0]->ctx->st));
  for (size_t q = 0; q < b->cs.size(); q++) {
    fz_case *cs = b->cs[q];
    if (!cs->check_pending) continue;
    int kill = 0;
    check_finish(cs, &kill);
    if (kill) {
      *stop = true;
      if (killed) killed[q] = 1;
    }
  }
  return FZ_OK;
}

// one batched step enqueued, with each case's checks when due
static int batch_step_enqueue(fz_case_batch *b, bool *pending) {
  FZ_TRY(batch_enqueue(b));
  *pending = false;
  for (fz_case *cs : b->cs) {
    const int istep = (int)cs->sc[S_ISTEP];
    if (cs->prm.icheck > 0 && istep % cs->prm.icheck == 0) {
      FZ_TRY(check_enqueue(cs));
      *pending = true;
    }
  }
  return FZ_OK;
}

extern "C" int fz_case_batch_step(fz_case_batch *b, int nsteps, int *killed) {
  if (!b || nsteps < 0) return fz_set_error(FZ_ERR_ARG, "fz_case_batch_step: bad argument");
  if (b->pending) return fz_set_error(FZ_ERR_STATE, "fz_case_batch_step: fz_case_batch_step_end has not been called");
  const int n = (int)b->cs.size();
  if (killed)
    for (int q = 0; q < n; q++) killed[q] = 0;
  for (fz_case *cs : b->cs)
    if (cs->check_pending) return fz_set_error(FZ_ERR_STATE, "fz_case_batch_step: a case has a step not ended");
  FZ_TRY(batch_enter(b));
  int rc = FZ_OK;
  bool stop = false;
  for (int s = 0; s < nsteps && !stop && rc == FZ_OK; s++) {
    bool pending = false;
    rc = batch_step_enqueue(b, &pending);
    if (rc == FZ_OK) rc = batch_checks(b, killed, &stop);
  }
  batch_leave(b);
  return rc;
}

extern "C" int fz_case_batch_step_begin(fz_case_batch *b, int *check_pending) {
  if (!b) return fz_set_error(FZ_ERR_ARG, "fz_case_batch_step_begin: null batch");
  if (b->pending) return fz_set_error(FZ_ERR_STATE, "fz_case_batch_step_begin: the previous step has not ended");
  for (fz_case *cs : b->cs)
    if (cs->check_pending) return fz_set_error(FZ_ERR_STATE, "fz_case_batch_step_begin: a case has a step not ended");
  FZ_TRY(batch_enter(b));
  bool pending = false;
  const int rc = batch_step_enqueue(b, &pending);
  batch_leave(b);
  b->pending = rc == FZ_OK;
  if (check_pending) *check_pending = pending ? 1 : 0;
  return rc;
}

extern "C" int fz_case_batch_step_end(fz_case_batch *b, int *killed) {
  if (!b) return fz_set_error(FZ_ERR_ARG, "fz_case_batch_step_end: null batch");
  if (killed)
    for (size_t q = 0; q < b->cs.size(); q++) killed[q] = 0;
  if (!b->pending) return FZ_OK;
  b->pending = false;
  bool stop = false;
  return batch_checks(b, killed, &stop);
}
