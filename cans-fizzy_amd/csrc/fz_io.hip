// Field files of the reference, read and written from device fields:
//  - restart files (load_one/io_field, src/load.f90:19-168): the interior of
//    one field, FP64 in Fortran order, followed by the two reals
//    [time, istep]; a file whose size is not (ng1*ng2*ng3 + 2)*8 bytes is
//    refused, as load_one aborts on it (src/load.f90:41-51);
//  - 3-D output fields (out3d, src/output.f90:185-236): the interior
//    sampled every nskip points from index 1, FP64 in Fortran order, no
//    header;
//  - the grid files of src/main.f90:230-243: grid.bin (dzc, dzf, zc, zf over
//    k = 1..n3, stream), grid.out (one text line per k = 0..n3+1) and
//    geometry.out (ng, then l).
// Interior planes move between the halo'd device field and a packed device
// buffer with a gather/scatter kernel, then cross PCIe in one copy.
#include <cstdio>
#include <string>
#include <vector>

#include "fz_internal.h"

// packed[(i-1)/si + ni*((j-1)/sj + nj*(k-1)/sk)] <-> field(i,j,k), i = 1, 1+si, ...
__global__ void k_pack(fz_dims d, int si, int sj, int sk, int ni, int nj, int nk, int to_packed,
                       double *__restrict__ field, double *__restrict__ packed) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)ni * nj * nk;
  if (t >= n) return;
  const int a = (int)(t % ni);
  const long rest = t / ni;
  const int b = (int)(rest % nj);
  const int c = (int)(rest / nj);
  const long q = (1 + (long)a * si) + d.sy * (1 + (long)b * sj) + d.sz * (1 + (long)c * sk);
  if (to_packed) packed[t] = field[q];
  else field[q] = packed[t];
}

static int pack_interior(fz_ctx *c, const int s[3], int to_packed, double *field, std::vector<double> &host) {
  const fz_dims &d = c->d;
  const int ni = (d.n1 - 1) / s[0] + 1, nj = (d.n2 - 1) / s[1] + 1, nk = (d.n3 - 1) / s[2] + 1;
  const long n = (long)ni * nj * nk;
  double *buf = nullptr;
  FZ_HIP(hipMalloc((void **)&buf, sizeof(double) * n));
  int rc = FZ_OK;
  hipError_t e = hipSuccess;
  if (!to_packed) e = hipMemcpyAsync(buf, host.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->st);
  if (e == hipSuccess) {
    const long nb = (n + 255) / 256;
    FZ_LAUNCH(c, "pack", k_pack, dim3((unsigned)nb), dim3(256), 0, d, s[0], s[1], s[2], ni, nj, nk, to_packed, field,
              buf);
    e = hipGetLastError();
  }
  if (e == hipSuccess && to_packed) {
    host.resize(n);
    e = hipMemcpyAsync(host.data(), buf, sizeof(double) * n, hipMemcpyDeviceToHost, c->st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "field transfer: %s", hipGetErrorString(e));
  (void)hipFree(buf);
  return rc;
}

extern "C" int fz_load_one(fz_ctx *c, char io, const char *filename, double *p, double *time, int *istep) {
  if (!c || !filename || !p || !time || !istep) return fz_set_error(FZ_ERR_ARG, "fz_load_one: null argument");
  const int one[3] = {1, 1, 1};
  const long n = c->ni;
  if (io == 'w') {
    std::vector<double> h;
    FZ_TRY(pack_interior(c, one, 1, p, h));
    FILE *fp = fopen(filename, "wb");
    if (!fp) return fz_set_error(FZ_ERR_ARG, "fz_load_one: cannot open %s for writing", filename);
    const double info[2] = {*time, (double)*istep};  // fldinfo = [time, 1._rp*istep]
    const bool ok = fwrite(h.data(), sizeof(double), n, fp) == (size_t)n && fwrite(info, sizeof(double), 2, fp) == 2;
    if (fclose(fp) != 0 || !ok) return fz_set_error(FZ_ERR_ARG, "fz_load_one: write to %s failed", filename);
    return FZ_OK;
  }
  if (io != 'r') return fz_set_error(FZ_ERR_ARG, "fz_load_one: io must be 'r' or 'w'");
  FILE *fp = fopen(filename, "rb");
  if (!fp) return fz_set_error(FZ_ERR_ARG, "fz_load_one: cannot open %s", filename);
  fseek(fp, 0, SEEK_END);
  const long size = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  const long good = (n + 2) * (long)sizeof(double);
  if (size != good) {
    fclose(fp);
    return fz_set_error(FZ_ERR_ARG, "checkpoint file with incorrect size: %s | expected size: %ld | actual size: %ld",
                        filename, good, size);
  }
  std::vector<double> h(n);
  double info[2];
  const bool ok = fread(h.data(), sizeof(double), n, fp) == (size_t)n && fread(info, sizeof(double), 2, fp) == 2;
  fclose(fp);
  if (!ok) return fz_set_error(FZ_ERR_ARG, "fz_load_one: read of %s failed", filename);
  FZ_TRY(pack_interior(c, one, 0, p, h));
  *time = info[0];
  *istep = (int)lround(info[1]);  // istep = nint(fldinfo(2))
  return FZ_OK;
}

extern "C" int fz_out3d(fz_ctx *c, const char *filename, const int nskip[3], const double *p) {
  if (!c || !filename || !nskip || !p) return fz_set_error(FZ_ERR_ARG, "fz_out3d: null argument");
  for (int q = 0; q < 3; q++)
    if (nskip[q] < 1) return fz_set_error(FZ_ERR_ARG, "fz_out3d: nskip must be >= 1");
  std::vector<double> h;
  FZ_TRY(pack_interior(c, nskip, 1, (double *)p, h));
  FILE *fp = fopen(filename, "wb");
  if (!fp) return fz_set_error(FZ_ERR_ARG, "fz_out3d: cannot open %s", filename);
  const bool ok = fwrite(h.data(), sizeof(double), h.size(), fp) == h.size();
  if (fclose(fp) != 0 || !ok) return fz_set_error(FZ_ERR_ARG, "fz_out3d: write to %s failed", filename);
  return FZ_OK;
}

extern "C" int fz_write_grid(fz_ctx *c, const char *datadir) {
  if (!c || !datadir) return fz_set_error(FZ_ERR_ARG, "fz_write_grid: null argument");
  const std::string dir(datadir);
  const int n3 = c->d.n3;
  FILE *fp = fopen((dir + "grid.bin").c_str(), "wb");
  if (!fp) return fz_set_error(FZ_ERR_ARG, "fz_write_grid: cannot open %sgrid.bin", datadir);
  const std::vector<double> *cols[4] = {&c->h_dzc, &c->h_dzf, &c->h_zc, &c->h_zf};
  bool ok = true;
  for (int q = 0; q < 4; q++) ok = ok && fwrite(cols[q]->data() + 1, sizeof(double), n3, fp) == (size_t)n3;
  if (fclose(fp) != 0 || !ok) return fz_set_error(FZ_ERR_ARG, "fz_write_grid: grid.bin write failed");
  fp = fopen((dir + "grid.out").c_str(), "w");
  if (!fp) return fz_set_error(FZ_ERR_ARG, "fz_write_grid: cannot open %sgrid.out", datadir);
  for (int k = 0; k <= n3 + 1; k++)
    fprintf(fp, " %.17g %.17g %.17g %.17g %.17g\n", 0.0, c->h_zf[k], c->h_zc[k], c->h_dzf[k], c->h_dzc[k]);
  fclose(fp);
  fp = fopen((dir + "geometry.out").c_str(), "w");
  if (!fp) return fz_set_error(FZ_ERR_ARG, "fz_write_grid: cannot open %sgeometry.out", datadir);
  fprintf(fp, " %d %d %d\n %.17g %.17g %.17g\n", c->d.n1, c->d.n2, c->d.n3, c->l[0], c->l[1], c->l[2]);
  fclose(fp);
  return FZ_OK;
}
