// Context, memory and grid set-up of libfizzy.
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>

#include "fz_internal.h"

static thread_local char g_err[512] = "";

int fz_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char *fz_last_error(void) { return g_err; }

// initgrid, src/initgrid.f90:14-78. The normalised face coordinate
// (k-0.)/(1.*n) is a default-real (single precision) quotient in the
// reference, so it is formed in float here as well.
static double gridpoint(int gtype, double alpha, double z0) {
  if (alpha == 0.0) return z0;
  switch (gtype) {
    case 2: return 1.0 * (1.0 + tanh((z0 - 1.0) * alpha) / tanh(alpha / 1.0));
    case 3: return 1.0 - 1.0 * (1.0 + tanh((1.0 - z0 - 1.0) * alpha) / tanh(alpha / 1.0));
    case 4:
      if (z0 <= 0.5) return 0.5 * (1.0 - 1.0 + tanh(2.0 * alpha * (z0 - 0.0)) / tanh(alpha));
      return 0.5 * (1.0 + 1.0 + tanh(2.0 * alpha * (z0 - 1.0)) / tanh(alpha));
    default: return 0.5 * (1.0 + tanh((z0 - 0.5) * alpha) / tanh(alpha / 2.0));
  }
}

extern "C" int fz_initgrid(int gtype, int n, double gr, double lz, double *dzc, double *dzf, double *zc,
                           double *zf) {
  if (n < 1 || !dzc || !dzf || !zc || !zf) return fz_set_error(FZ_ERR_ARG, "fz_initgrid: bad arguments");
  zf[0] = 0.0;
  for (int k = 1; k <= n; k++) {
    const double z0 = (double)((float)k / (float)n);
    zf[k] = gridpoint(gtype, gr, z0) * lz;
  }
  for (int k = 1; k <= n; k++) dzf[k] = zf[k] - zf[k - 1];
  dzf[0] = dzf[1];
  dzf[n + 1] = dzf[n];
  for (int k = 0; k <= n; k++) dzc[k] = 0.5 * (dzf[k] + dzf[k + 1]);
  dzc[n + 1] = dzc[n];
  zc[0] = -dzc[0] / 2.0;
  zf[0] = 0.0;
  for (int k = 1; k <= n + 1; k++) {
    zc[k] = zc[k - 1] + dzc[k - 1];
    zf[k] = zf[k - 1] + dzf[k];
  }
  return FZ_OK;
}

static bool valid_bc(char c) { return c == 'P' || c == 'D' || c == 'N'; }

extern "C" int fz_ctx_create(fz_ctx **out, const int n[3], const double l[3], const double *dzc, const double *dzf,
                             const double *zc, const double *zf, const char cbcpre[6], void *hip_stream) {
  if (!out || !n || !l || !dzc || !dzf || !zc || !zf || !cbcpre)
    return fz_set_error(FZ_ERR_ARG, "fz_ctx_create: null argument");
  for (int d = 0; d < 3; d++)
    if (n[d] < 1) return fz_set_error(FZ_ERR_ARG, "fz_ctx_create: n(%d)=%d must be >= 1", d + 1, n[d]);
  for (int q = 0; q < 6; q++)
    if (!valid_bc(cbcpre[q])) return fz_set_error(FZ_ERR_ARG, "fz_ctx_create: bad cbcpre code '%c'", cbcpre[q]);
  fz_ctx *c = new fz_ctx();
  c->d.n1 = n[0];
  c->d.n2 = n[1];
  c->d.n3 = n[2];
  c->d.sy = n[0] + 2;
  c->d.sz = (long)(n[0] + 2) * (n[1] + 2);
  c->d.tile_y = 8;
  if (const char *e = getenv("FZ_TILE_Y")) c->d.tile_y = atoi(e);  // block-order tuning
  if (const char *e = getenv("FZ_MOM_WPE")) c->mom_variant = atoi(e);
  if (const char *e = getenv("FZ_CORREC_KZ")) c->correc_kz = atoi(e);
  if (const char *e = getenv("FZ_CORREC_MY")) c->correc_my = atoi(e);
  if (c->correc_my != 8 && c->correc_my != 16) c->correc_my = 4;  // the tile heights the kernel is built for
  if (const char *e = getenv("FZ_P2_XDIV")) c->p2_xdiv = atoi(e);
  if (const char *e = getenv("FZ_ACDI_VARIANT")) c->acdi_variant = atoi(e);
  if (const char *e = getenv("FZ_YOUNGS_VARIANT")) c->youngs_variant = atoi(e);
  // the Thomas kernel without a pivot array pays off once the columns fill
  // the device (measured: slower at 256x256 columns, faster at 512x512)
  c->thomas_rc = (long)n[0] * n[1] >= 262144 ? 1 : 0;
  // below that, the FP64 solve of a non-periodic z reads pivots made once at
  // set-up (two more N-element arrays; 294 against 330 us at 256x256x512)
  c->thomas_pre = c->thomas_rc ? 0 : 1;
  if (const char *e = getenv("FZ_THOMAS_RC")) c->thomas_rc = atoi(e);
  if (const char *e = getenv("FZ_THOMAS_PRE")) c->thomas_pre = atoi(e);
  if (const char *e = getenv("FZ_THOMAS_ROWS32")) c->tk32 = atoi(e);
  if (const char *e = getenv("FZ_THOMAS_ROWS64")) c->tk64 = atoi(e);
  c->nh = c->d.sz * (n[2] + 2);
  c->ni = (long)n[0] * n[1] * n[2];
  for (int d = 0; d < 3; d++) {
    c->l[d] = l[d];
    // dl(:) = l(:)/(1.*ng(:)), src/param.f90:170-171
    c->dl[d] = l[d] / (double)(float)n[d];
    c->dli[d] = 1.0 / c->dl[d];
  }
  const int nz = n[2] + 2;
  c->h_dzc.assign(dzc, dzc + nz);
  c->h_dzf.assign(dzf, dzf + nz);
  c->h_zc.assign(zc, zc + nz);
  c->h_zf.assign(zf, zf + nz);
  c->h_dzci.resize(nz);
  c->h_dzfi.resize(nz);
  for (int k = 0; k < nz; k++) {  // dzci = dzc**(-1), src/main.f90:259-260
    c->h_dzci[k] = 1.0 / dzc[k];
    c->h_dzfi[k] = 1.0 / dzf[k];
  }
  // the Youngs kernel may leave out its per-lane range checks (given the
  // phase field's range flag) only where its grid factors lie in [2^-20, 2^20]
  bool fast = true;
  for (int d = 0; d < 2; d++) fast = fast && c->dli[d] >= 0x1p-20 && c->dli[d] <= 0x1p20;
  for (int k = 0; k < nz; k++)
    fast = fast && c->h_dzci[k] >= 0x1p-20 && c->h_dzci[k] <= 0x1p20 && c->h_dzfi[k] >= 0x1p-20 &&
           c->h_dzfi[k] <= 0x1p20;
  c->youngs_fast_grid = fast;
  memcpy(c->cbcpre, cbcpre, 6);
  for (int d = 0; d < 3; d++) c->per[d] = (cbcpre[2 * d] == 'P' && cbcpre[2 * d + 1] == 'P');
  hipError_t e;
  if (hip_stream) {
    c->st = (hipStream_t)hip_stream;
  } else {
    e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fz_set_error(FZ_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  double **dev[6] = {&c->dzc, &c->dzf, &c->zc, &c->zf, &c->dzci, &c->dzfi};
  const std::vector<double> *host[6] = {&c->h_dzc, &c->h_dzf, &c->h_zc, &c->h_zf, &c->h_dzci, &c->h_dzfi};
  for (int q = 0; q < 6; q++) {
    e = hipMalloc((void **)dev[q], sizeof(double) * nz);
    if (e == hipSuccess) e = hipMemcpy(*dev[q], host[q]->data(), sizeof(double) * nz, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      fz_ctx_destroy(c);
      return fz_set_error(FZ_ERR_HIP, "fz_ctx_create grid upload: %s", hipGetErrorString(e));
    }
  }
  e = hipHostMalloc((void **)&c->h_small, 64 * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) {
    fz_ctx_destroy(c);
    return fz_set_error(FZ_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  const int rs = fz_scratch_setup(c);
  if (rs != FZ_OK) {
    fz_ctx_destroy(c);
    return rs;
  }
  *out = c;
  return FZ_OK;
}

extern "C" int fz_ctx_destroy(fz_ctx *c) {
  if (!c) return FZ_OK;
  if (c->st) (void)hipStreamSynchronize(c->st);
  double *dev[6] = {c->dzc, c->dzf, c->zc, c->zf, c->dzci, c->dzfi};
  for (int q = 0; q < 6; q++)
    if (dev[q]) (void)hipFree(dev[q]);
  if (c->pois) fz_poisson_destroy(c->pois);
  fz_bcmaps_free(c);
  for (const fz_prof_rec &r : c->prof_recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
  if (c->red) (void)hipFree(c->red);
  if (c->ftab) (void)hipFree(c->ftab);
  if (c->h_small) (void)hipHostFree(c->h_small);
  if (c->own_stream && c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return FZ_OK;
}

extern "C" int fz_ctx_sync(fz_ctx *c) {
  if (!c) return fz_set_error(FZ_ERR_ARG, "fz_ctx_sync: null context");
  FZ_HIP(hipStreamSynchronize(c->st));
  return FZ_OK;
}

extern "C" int fz_malloc(fz_ctx *c, size_t bytes, void **dptr) {
  if (!c || !dptr) return fz_set_error(FZ_ERR_ARG, "fz_malloc: null argument");
  FZ_HIP(hipMalloc(dptr, bytes ? bytes : 1));
  return FZ_OK;
}

extern "C" int fz_free(fz_ctx *c, void *dptr) {
  if (!c) return fz_set_error(FZ_ERR_ARG, "fz_free: null context");
  if (dptr) {
    FZ_HIP(hipStreamSynchronize(c->st));
    FZ_HIP(hipFree(dptr));
  }
  return FZ_OK;
}

extern "C" int fz_upload(fz_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return fz_set_error(FZ_ERR_ARG, "fz_upload: null argument");
  FZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->st));
  FZ_HIP(hipStreamSynchronize(c->st));
  return FZ_OK;
}

extern "C" int fz_download(fz_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return fz_set_error(FZ_ERR_ARG, "fz_download: null argument");
  FZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->st));
  FZ_HIP(hipStreamSynchronize(c->st));
  return FZ_OK;
}

// plain streaming copy, 16 B per lane per iteration (also the HBM calibration kernel of the profiles)
__global__ void __launch_bounds__(256) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int fz_copy(fz_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return fz_set_error(FZ_ERR_ARG, "fz_copy: null argument");
  if (!bytes) return FZ_OK;
  const bool aligned = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0);
  const size_t nvec = aligned ? bytes / 16 : 0;
  if (nvec) {
    const long nb = (long)std::min<size_t>((nvec + 255) / 256, 8192);
    FZ_LAUNCH(c, "copy", k_copy, dim3(nb), dim3(256), 0, (const double2 *)src, (double2 *)dst, (long)nvec);
    FZ_LAUNCHED();
  }
  const size_t done = nvec * 16;
  if (done < bytes)
    FZ_HIP(hipMemcpyAsync((char *)dst + done, (const char *)src + done, bytes - done, hipMemcpyDeviceToDevice, c->st));
  return FZ_OK;
}

// ---- per-launch timing ----
static hipEvent_t prof_event(fz_ctx *c) {
  if (!c->prof_pool.empty()) {
    hipEvent_t e = c->prof_pool.back();
    c->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

fz_prof_scope::fz_prof_scope(fz_ctx *c_, const char *n) : c(c_), name(n) {
  if (!c || !c->prof_on) return;
  a = prof_event(c);
  if (a) (void)hipEventRecord(a, c->st);
}

fz_prof_scope::~fz_prof_scope() {
  if (!a) return;
  hipEvent_t b = prof_event(c);
  if (!b) return;
  (void)hipEventRecord(b, c->st);
  c->prof_recs.push_back(fz_prof_rec{name, a, b});
}

extern "C" int fz_prof_enable(fz_ctx *c, int on) {
  if (!c) return fz_set_error(FZ_ERR_ARG, "fz_prof_enable: null context");
  c->prof_on = on != 0;
  return FZ_OK;
}

// "name calls total_ms" lines, one per launch site, in first-seen order; clears the records
extern "C" int fz_prof_report(fz_ctx *c, char *buf, size_t cap) {
  if (!c || !buf || cap == 0) return fz_set_error(FZ_ERR_ARG, "fz_prof_report: bad argument");
  FZ_HIP(hipStreamSynchronize(c->st));
  std::vector<const char *> names;
  std::vector<double> tot;
  std::vector<long> cnt;
  for (const fz_prof_rec &r : c->prof_recs) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    size_t q = 0;
    while (q < names.size() && strcmp(names[q], r.name) != 0) q++;
    if (q == names.size()) {
      names.push_back(r.name);
      tot.push_back(0.0);
      cnt.push_back(0);
    }
    tot[q] += ms;
    cnt[q] += 1;
    c->prof_pool.push_back(r.a);
    c->prof_pool.push_back(r.b);
  }
  c->prof_recs.clear();
  size_t used = 0;
  buf[0] = '\0';
  for (size_t q = 0; q < names.size() && used + 1 < cap; q++) {
    const int w = snprintf(buf + used, cap - used, "%s %ld %.6f\n", names[q], cnt[q], tot[q]);
    if (w < 0) break;
    used += (size_t)w;
  }
  return FZ_OK;
}

int fz_reduce_ensure(fz_ctx *c, size_t nblocks) {
  if (nblocks <= c->red_cap) return FZ_OK;
  if (c->red) {
    FZ_HIP(hipStreamSynchronize(c->st));
    FZ_HIP(hipFree(c->red));
    c->red = nullptr;
  }
  FZ_HIP(hipMalloc((void **)&c->red, sizeof(double) * 2 * nblocks));
  c->red_cap = nblocks;
  return FZ_OK;
}

// acdi_set_epsilon, src/acdi.f90:16-30: dlmax over dl(1:2) and 1/dzfi(0:n3+1)
extern "C" int fz_acdi_set_epsilon(fz_ctx *c, double seps_factor, double *seps) {
  if (!c || !seps) return fz_set_error(FZ_ERR_ARG, "fz_acdi_set_epsilon: null argument");
  double dlmax = c->dl[0] > c->dl[1] ? c->dl[0] : c->dl[1];
  double m = 1.0 / c->h_dzfi[0];
  for (size_t k = 1; k < c->h_dzfi.size(); k++) {
    const double v = 1.0 / c->h_dzfi[k];
    if (v > m) m = v;
  }
  if (m > dlmax) dlmax = m;
  *seps = dlmax * seps_factor;
  return FZ_OK;
}
