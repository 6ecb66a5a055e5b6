// Global reductions of the hot path: the ACDI gamma (src/acdi.f90:32-62),
// chkdiv (src/chkdiv.f90:15-49), the advective part of chkdt
// (src/chkdt.f90:48-67), bulk_mean_12_stag (src/utils.f90:68-136) and the
// chk_poisson residual (src/debug.f90:81-114). Two passes: per-block partials
// with wavefront shuffles, then one block over the partials; the order is
// fixed, so results do not depend on timing. Results stay on the device for
// the case driver; the reference-style entry points copy them to the host.
#include <cmath>
#include <mutex>

#include "fz_internal.h"

#define RED_THREADS 256
#define RED_BLOCKS 1024

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
// block reduce; result valid in thread 0
__device__ inline double block_max(double v) {
  __shared__ double sh[RED_THREADS / 64];
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = fmax(m, sh[q]);
    v = m;
  }
  __syncthreads();
  return v;
}
__device__ inline double block_sum(double v) {
  __shared__ double sh[RED_THREADS / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = sh[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) s += sh[q];
    v = s;
  }
  __syncthreads();
  return v;
}

// Each block reduces one contiguous range of the interior cells (in the
// x-fastest order), and the XCD remap gives each XCD a contiguous run of
// blocks: the y and z neighbours a cell reads lie in the ranges of the same
// XCD, so its L2 holds them (strided ranges sent every neighbour row to
// another XCD's L2, and the dt reduction fetched 3.7x its bytes).
#define FOR_CELLS(d)                                                                                     \
  const long ncell = (long)(d).n1 * (d).n2 * (d).n3;                                                     \
  const long per_ = (ncell + gridDim.x - 1) / gridDim.x;                                                 \
  const long b0_ = (long)fz_xcd_block(blockIdx.x, gridDim.x) * per_;                                     \
  const long b1_ = b0_ + per_ < ncell ? b0_ + per_ : ncell;                                              \
  for (long idx = b0_ + threadIdx.x; idx < b1_; idx += blockDim.x) {                                     \
    const int i = (int)(idx % (d).n1) + 1;                                                               \
    const long t_ = idx / (d).n1;                                                                        \
    const int j = (int)(t_ % (d).n2) + 1;                                                                \
    const int k = (int)(t_ / (d).n2) + 1;                                                                \
    const long q = i + (d).sy * j + (d).sz * k;

__global__ void k_red_gamma(fz_dims d, const double *__restrict__ u, const double *__restrict__ v,
                            const double *__restrict__ w, double *__restrict__ part) {
  double m = 0.0;
  FOR_CELLS(d)
  (void)k;
  const double uc = 0.5 * (u[q] + u[q - 1]);
  const double vc = 0.5 * (v[q] + v[q - d.sy]);
  const double wc = 0.5 * (w[q] + w[q - d.sz]);
  m = fmax(m, uc * uc + vc * vc + wc * wc);
}
m = block_max(m);
if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// gam = sqrt(velmax)*gam_factor, then max(gam, gam_min) (src/main.f90:526-527)
__global__ void k_fin_gamma(int nb, const double *__restrict__ part, double factor, double gmin,
                            double *__restrict__ out) {
  double m = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) m = fmax(m, part[q]);
  m = block_max(m);
  if (threadIdx.x == 0) out[0] = fmax(sqrt(m) * factor, gmin);
}

__global__ void k_red_div(fz_dims d, const double *__restrict__ dzfi, double dxi, double dyi,
                          const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
                          double *__restrict__ part) {
  double s = 0.0, m = 0.0;
  FOR_CELLS(d)
  const double div = (w[q] - w[q - d.sz]) * dzfi[k] + (v[q] - v[q - d.sy]) * dyi + (u[q] - u[q - 1]) * dxi;
  m = fmax(m, fabs(div));
  s = s + div;
}
s = block_sum(s);
m = block_max(m);
if (threadIdx.x == 0) {
  part[2 * blockIdx.x] = s;
  part[2 * blockIdx.x + 1] = m;
}
}

__global__ void k_fin_div(int nb, const double *__restrict__ part, double *__restrict__ out) {
  double s = 0.0, m = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) {
    s += part[2 * q];
    m = fmax(m, part[2 * q + 1]);
  }
  s = block_sum(s);
  m = block_max(m);
  if (threadIdx.x == 0) {
    out[0] = s;
    out[1] = m;
  }
}

__global__ void k_red_dt(fz_dims d, const double *__restrict__ dzci, const double *__restrict__ dzfi, double dxi,
                         double dyi, const double *__restrict__ u, const double *__restrict__ v,
                         const double *__restrict__ w, double *__restrict__ part) {
  double dti = 0.0;
  const long sy = d.sy, sz = d.sz;
  FOR_CELLS(d)
  const double ux = fabs(u[q]);
  const double vx = 0.25 * fabs(v[q] + v[q - sy] + v[q + 1] + v[q + 1 - sy]);
  const double wx = 0.25 * fabs(w[q] + w[q - sz] + w[q + 1] + w[q + 1 - sz]);
  const double dtix = ux * dxi + vx * dyi + wx * dzfi[k];
  const double uy = 0.25 * fabs(u[q] + u[q + sy] + u[q - 1 + sy] + u[q - 1]);
  const double vy = fabs(v[q]);
  const double wy = 0.25 * fabs(w[q] + w[q + sy] + w[q + sy - sz] + w[q - sz]);
  const double dtiy = uy * dxi + vy * dyi + wy * dzfi[k];
  const double uz = 0.25 * fabs(u[q] + u[q - 1] + u[q - 1 + sz] + u[q + sz]);
  const double vz = 0.25 * fabs(v[q] + v[q - sy] + v[q - sy + sz] + v[q + sz]);
  const double wz = fabs(w[q]);
  const double dtiz = uz * dxi + vz * dyi + wz * dzci[k];
  dti = fmax(dti, fmax(dtix, fmax(dtiy, dtiz)));
}
dti = block_max(dti);
if (threadIdx.x == 0) part[blockIdx.x] = dti;
}

__global__ void k_fin_max(int nb, const double *__restrict__ part, double *__restrict__ out) {
  double m = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) m = fmax(m, part[q]);
  m = block_max(m);
  if (threadIdx.x == 0) out[0] = m;
}

template <typename PT>  // FP64 sums also over the FP32 phase field of the mixed-precision step
__global__ void k_red_bulk(fz_dims d, long off, const double *__restrict__ gvr, double p1, double p2,
                           const PT *__restrict__ psi, double *__restrict__ part) {
  double s = 0.0;
  FOR_CELLS(d)
  const double pl = 0.5 * ((double)psi[q + off] + (double)psi[q]);
  s = s + (pl * p1 + (1.0 - pl) * p2) * gvr[k];
}
s = block_sum(s);
if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_fin_sum(int nb, const double *__restrict__ part, double *__restrict__ out) {
  double s = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) s += part[q];
  s = block_sum(s);
  if (threadIdx.x == 0) out[0] = s;
}

__global__ void k_red_poisson(fz_dims d, const double *__restrict__ dzci, const double *__restrict__ dzfi,
                              double dxi, double dyi, const double *__restrict__ fp, const double *__restrict__ fpp,
                              double *__restrict__ part) {
  double m = 0.0;
  const long sy = d.sy, sz = d.sz;
  FOR_CELLS(d)
  const double val = (fpp[q + 1] - 2.0 * fpp[q] + fpp[q - 1]) * (dxi * dxi) +
                     (fpp[q + sy] - 2.0 * fpp[q] + fpp[q - sy]) * (dyi * dyi) +
                     ((fpp[q + sz] - fpp[q]) * dzci[k] - (fpp[q] - fpp[q - sz]) * dzci[k - 1]) * dzfi[k];
  m = fmax(m, fabs(val - fp[q]));
}
m = block_max(m);
if (threadIdx.x == 0) part[blockIdx.x] = m;
}

static int red_blocks(const fz_ctx *c) {
  long nb = (c->ni + RED_THREADS - 1) / RED_THREADS;
  if (nb > RED_BLOCKS) nb = RED_BLOCKS;
  return (int)(nb < 1 ? 1 : nb);
}

// gamma into a device scalar (no host synchronisation)
int fz_acdi_gamma_dev(fz_ctx *c, double gam_factor, double gam_min, const double *u, const double *v,
                      const double *w, double *d_gam) {
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb));
  FZ_LAUNCH(c, "red_gamma", k_red_gamma, dim3(nb), dim3(RED_THREADS), 0, c->d, u, v, w, c->red);
  FZ_LAUNCHED();
  FZ_LAUNCH(c, "fin_gamma", k_fin_gamma, dim3(1), dim3(RED_THREADS), 0, nb, c->red, gam_factor, gam_min, d_gam);
  FZ_LAUNCHED();
  return FZ_OK;
}

static int fetch(fz_ctx *c, const double *d_src, int count, double *out) {
  FZ_HIP(hipMemcpyAsync(c->h_small, d_src, sizeof(double) * count, hipMemcpyDeviceToHost, c->st));
  FZ_HIP(hipStreamSynchronize(c->st));
  for (int q = 0; q < count; q++) out[q] = c->h_small[q];
  return FZ_OK;
}

extern "C" int fz_acdi_set_gamma(fz_ctx *c, double gam_factor, const double *u, const double *v, const double *w,
                                 double *gam) {
  if (!c || !u || !v || !w || !gam) return fz_set_error(FZ_ERR_ARG, "fz_acdi_set_gamma: null argument");
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  double *res = c->red + 2 * (c->red_cap - 4);
  FZ_TRY(fz_acdi_gamma_dev(c, gam_factor, 0.0, u, v, w, res));
  return fetch(c, res, 1, gam);
}

int fz_chkdiv_dev(fz_ctx *c, const double *u, const double *v, const double *w, double *d_out) {
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  FZ_LAUNCH(c, "red_div", k_red_div, dim3(nb), dim3(RED_THREADS), 0, c->d, c->dzfi, c->dli[0], c->dli[1], u, v, w,
                     c->red);
  FZ_LAUNCHED();
  FZ_LAUNCH(c, "fin_div", k_fin_div, dim3(1), dim3(RED_THREADS), 0, nb, c->red, d_out);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_chkdiv(fz_ctx *c, const double *u, const double *v, const double *w, double *divtot,
                         double *divmax) {
  if (!c || !u || !v || !w || !divtot || !divmax) return fz_set_error(FZ_ERR_ARG, "fz_chkdiv: null argument");
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  double *res = c->red + 2 * (c->red_cap - 4);
  FZ_TRY(fz_chkdiv_dev(c, u, v, w, res));
  double o[2];
  FZ_TRY(fetch(c, res, 2, o));
  *divtot = o[0];
  *divmax = o[1];
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// chkdiv without allocation. The per-block partials go to a file-scope device
// array, `scratch`, instead of the context's buffer, and the two results to a
// device buffer of the caller, so the call neither allocates nor synchronises.
// Each GPU holds its own copy of the array until the process ends; every call
// on one GPU shares it, so calls that overlap on one GPU race. Each call
// writes the partials it reads before reading them, so nothing needs clearing.
#define FZ_SCRATCH_BLOCKS RED_BLOCKS
#define FZ_SCRATCH_DEVICES 64
static __device__ __align__(256) double scratch[2 * FZ_SCRATCH_BLOCKS];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[FZ_SCRATCH_DEVICES];

// Called by fz_ctx_create with the context's device current: looks up, once
// per device, the device address of that device's `scratch` (in host code the
// array's name is no device address) and gives it to the context.
int fz_scratch_setup(fz_ctx *c) {
  int dev = 0;
  FZ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= FZ_SCRATCH_DEVICES)
    return fz_set_error(FZ_ERR_UNSUPPORTED, "device %d is beyond the %d devices of the scratch table", dev,
                        FZ_SCRATCH_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fz_set_error(FZ_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)p;
  }
  c->scratch = g_scratch_ptr[dev];
  return FZ_OK;
}

extern "C" int fz_chkdiv_noalloc(fz_ctx *c, const double *u, const double *v, const double *w, double *d_out) {
  if (!c || !u || !v || !w || !d_out) return fz_set_error(FZ_ERR_ARG, "fz_chkdiv_noalloc: null argument");
  if (!c->scratch) return fz_set_error(FZ_ERR_STATE, "fz_chkdiv_noalloc: the context has no scratch address");
  const int nb = red_blocks(c);  // the block count of fz_chkdiv, so the summation order is the same
  if (nb > FZ_SCRATCH_BLOCKS)
    return fz_set_error(FZ_ERR_ARG, "fz_chkdiv_noalloc: %d partial sums exceed the scratch maximum of %d", nb,
                        FZ_SCRATCH_BLOCKS);
  hipLaunchKernelGGL(k_red_div, dim3(nb), dim3(RED_THREADS), 0, c->st, c->d, c->dzfi, c->dli[0], c->dli[1], u, v, w,
                     c->scratch);
  FZ_LAUNCHED();
  hipLaunchKernelGGL(k_fin_div, dim3(1), dim3(RED_THREADS), 0, c->st, nb, c->scratch, d_out);
  FZ_LAUNCHED();
  return FZ_OK;
}

int fz_chkdt_adv_dev(fz_ctx *c, const double *u, const double *v, const double *w, double *d_out) {
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  FZ_LAUNCH(c, "red_dt", k_red_dt, dim3(nb), dim3(RED_THREADS), 0, c->d, c->dzci, c->dzfi, 1.0 / c->dl[0],
                     1.0 / c->dl[1], u, v, w, c->red);
  FZ_LAUNCHED();
  FZ_LAUNCH(c, "fin_max", k_fin_max, dim3(1), dim3(RED_THREADS), 0, nb, c->red, d_out);
  FZ_LAUNCHED();
  return FZ_OK;
}

// the scalar part of chkdt, src/chkdt.f90:71-93, from the advective maximum
double fz_chkdt_finish(const fz_ctx *c, double dti, int is_solve_ns, int is_track_interface, int vof,
                       const double mu12[2], const double rho12[2], double sigma, const double gacc[3], double gam,
                       double seps) {
  double dlmin = c->dl[0] < c->dl[1] ? c->dl[0] : c->dl[1];
  double m = 1.0 / c->h_dzfi[0];
  for (size_t k = 1; k < c->h_dzfi.size(); k++) m = fmin(m, 1.0 / c->h_dzfi[k]);
  dlmin = fmin(dlmin, m);
  double dtipsi;
  if (!vof) {
    dtipsi = gam * seps / (dlmin * dlmin);
    if (dtipsi == 0.0) dtipsi = 1.0;
  } else {
    dtipsi = dti;
  }
  double dtmax;
  if (!is_solve_ns) {
    dtmax = fmin(1.0 / dti, 1.0 / dtipsi);
  } else {
    const double dtiv = fmax(mu12[0] / rho12[0], mu12[1] / rho12[1]) * 2.0 * (3.0 / (dlmin * dlmin));
    const double dtik = sqrt(sigma / (fmin(rho12[0], rho12[1])) / (dlmin * dlmin * dlmin));
    const double dtig = fmax(fabs(gacc[0]), fmax(fabs(gacc[1]), fabs(gacc[2]))) / dlmin;
    dti = 0.5 * (dti + dtiv + sqrt((dti + dtiv) * (dti + dtiv) + 4.0 * (dtig * dtig + dtik * dtik)));
    if (dti == 0.0) dti = 1.0;
    dtmax = 1.0 / dti;
    if (is_track_interface) dtmax = fmin(dtmax, 1.0 / dtipsi);
  }
  return dtmax;
}

extern "C" int fz_chkdt(fz_ctx *c, int is_solve_ns, int is_track_interface, int vof, const double mu12[2],
                        const double rho12[2], double sigma, const double gacc[3], const double *u, const double *v,
                        const double *w, double gam, double seps, double *dtmax) {
  if (!c || !mu12 || !rho12 || !gacc || !u || !v || !w || !dtmax)
    return fz_set_error(FZ_ERR_ARG, "fz_chkdt: null argument");
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  double *res = c->red + 2 * (c->red_cap - 4);
  FZ_TRY(fz_chkdt_adv_dev(c, u, v, w, res));
  double dti;
  FZ_TRY(fetch(c, res, 1, &dti));
  *dtmax = fz_chkdt_finish(c, dti, is_solve_ns, is_track_interface, vof, mu12, rho12, sigma, gacc, gam, seps);
  return FZ_OK;
}

template <typename PT>
int fz_bulk_mean_dev(fz_ctx *c, int idir, const PT *psi, const double p12[2], const double *gvr, double *d_out) {
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  const long off = idir == 1 ? 1 : (idir == 2 ? c->d.sy : c->d.sz);
  FZ_LAUNCH(c, "red_bulk", k_red_bulk<PT>, dim3(nb), dim3(RED_THREADS), 0, c->d, off, gvr, p12[0], p12[1], psi,
            c->red);
  FZ_LAUNCHED();
  FZ_LAUNCH(c, "fin_sum", k_fin_sum, dim3(1), dim3(RED_THREADS), 0, nb, c->red, d_out);
  FZ_LAUNCHED();
  return FZ_OK;
}

template int fz_bulk_mean_dev<double>(fz_ctx *, int, const double *, const double *, const double *, double *);
template int fz_bulk_mean_dev<float>(fz_ctx *, int, const float *, const double *, const double *, double *);

extern "C" int fz_bulk_mean_12_stag(fz_ctx *c, int idir, const double *psi, const double p12[2], double *mean) {
  if (!c || !psi || !p12 || !mean || idir < 1 || idir > 3)
    return fz_set_error(FZ_ERR_ARG, "fz_bulk_mean_12_stag: bad argument");
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  // grid_vol_ratio_c for x and y faces, grid_vol_ratio_f for z faces (src/main.f90:268-269, :472-477)
  const int nz = c->d.n3 + 2;
  std::vector<double> g(nz);
  const double vol = c->l[0] * c->l[1] * c->l[2];
  for (int k = 0; k < nz; k++) g[k] = c->dl[0] * c->dl[1] * (idir < 3 ? c->h_dzc[k] : c->h_dzf[k]) / vol;
  double *dg = nullptr;
  FZ_HIP(hipMalloc((void **)&dg, sizeof(double) * nz));
  (void)hipMemcpyAsync(dg, g.data(), sizeof(double) * nz, hipMemcpyHostToDevice, c->st);
  double *res = c->red + 2 * (c->red_cap - 4);
  int rc = fz_bulk_mean_dev(c, idir, psi, p12, dg, res);
  if (rc == FZ_OK) rc = fetch(c, res, 1, mean);
  (void)hipFree(dg);
  return rc;
}

extern "C" int fz_chk_poisson(fz_ctx *c, const double *fp, const double *fpp, double *diffmax) {
  if (!c || !fp || !fpp || !diffmax) return fz_set_error(FZ_ERR_ARG, "fz_chk_poisson: null argument");
  const int nb = red_blocks(c);
  FZ_TRY(fz_reduce_ensure(c, nb + 8));
  FZ_LAUNCH(c, "red_poisson", k_red_poisson, dim3(nb), dim3(RED_THREADS), 0, c->d, c->dzci, c->dzfi, c->dli[0],
                     c->dli[1], fp, fpp, c->red);
  FZ_LAUNCHED();
  double *res = c->red + 2 * (c->red_cap - 4);
  FZ_LAUNCH(c, "fin_max", k_fin_max, dim3(1), dim3(RED_THREADS), 0, nb, c->red, res);
  FZ_LAUNCHED();
  return fetch(c, res, 1, diffmax);
}
