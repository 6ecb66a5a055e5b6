// Boundary conditions: boundp / bounduvw / set_bc of src/bound.f90:15-293.
//
// With the whole domain on one device the reference's single-rank sequence
// is kept exactly, because edge and corner ghost cells take the value of the
// last operation that writes them: first the periodic halos of y and z
// (updthalo with the periods of cbcpre, src/initmpi.f90:48-49 and
// src/bound.f90:398-475), then set_bc along x (always a boundary of the
// x-pencil), then y and z where they are not periodic. Each direction is one
// launch over its full halo-inclusive planes; the sub-operations of one
// direction run in order in the thread that owns a plane position.
#include <cstring>

#include "fz_internal.h"

namespace {
struct SubOp {
  int mode;  // 1 copy src->dst, 2 dst = factor + sgn*src, 3 dst = factor
  int dst, src;
  double factor, sgn;
};
struct FieldOps {
  int nops;
  SubOp op[4];
};
struct DirOps {
  FieldOps f[3];
};
}  // namespace

__global__ void k_plane_ops(fz_dims d, int dir, double *__restrict__ f0, double *__restrict__ f1,
                            double *__restrict__ f2, DirOps ops) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  long base, st;
  if (dir == 0) {
    if (a > d.n2 + 1) return;
    base = d.sy * a + d.sz * b;
    st = 1;
  } else if (dir == 1) {
    if (a > d.n1 + 1) return;
    base = a + d.sz * b;
    st = d.sy;
  } else {
    if (a > d.n1 + 1) return;
    base = a + d.sy * b;
    st = d.sz;
  }
  double *fs[3] = {f0, f1, f2};
  for (int fi = 0; fi < 3; fi++) {
    double *p = fs[fi];
    if (!p) continue;
    const FieldOps &fo = ops.f[fi];
    for (int s = 0; s < fo.nops; s++) {
      const SubOp &o = fo.op[s];
      double val;
      if (o.mode == 1) val = p[base + st * o.src];
      else if (o.mode == 2) val = o.factor + o.sgn * p[base + st * o.src];
      else val = o.factor;
      p[base + st * o.dst] = val;
    }
  }
}

static void push(FieldOps &fo, int mode, int dst, int src, double factor, double sgn) {
  fo.op[fo.nops++] = SubOp{mode, dst, src, factor, sgn};
}

// set_bc, src/bound.f90:126-293 with nh = 1
static void set_bc_ops(FieldOps &fo, char ctype, int ib, int nn, bool centered, double rvalue, double dr) {
  double factor = rvalue, sgn = 0.0;
  if (ctype == 'D' && centered) {
    factor = 2.0 * factor;
    sgn = -1.0;
  }
  if (ctype == 'N') {
    factor = (ib == 0) ? -dr * factor : dr * factor;
    sgn = 1.0;
  }
  if (ctype == 'P') {
    push(fo, 1, 0, nn, 0.0, 0.0);
    push(fo, 1, nn + 1, 1, 0.0, 0.0);
    return;
  }
  if (centered) {
    if (ib == 0) push(fo, 2, 0, 1, factor, sgn);
    else push(fo, 2, nn + 1, nn, factor, sgn);
  } else if (ctype == 'D') {
    if (ib == 0) push(fo, 3, 0, 0, factor, 0.0);
    else {
      push(fo, 1, nn + 1, nn - 1, 0.0, 0.0);
      push(fo, 3, nn, nn, factor, 0.0);
    }
  } else {
    if (ib == 0) push(fo, 2, 0, 1, factor, 1.0);
    else {
      push(fo, 1, nn + 1, nn, 0.0, 0.0);
      push(fo, 2, nn, nn - 1, factor, 1.0);
    }
  }
}

static int launch_dir(fz_ctx *c, int dir, double *f0, double *f1, double *f2, const DirOps &ops) {
  const fz_dims &d = c->d;
  int na, nb;
  if (dir == 0) { na = d.n2 + 2; nb = d.n3 + 2; }
  else if (dir == 1) { na = d.n1 + 2; nb = d.n3 + 2; }
  else { na = d.n1 + 2; nb = d.n2 + 2; }
  FZ_LAUNCH(c, "plane_ops", k_plane_ops, dim3((na + 63) / 64, nb), dim3(64), 0, d, dir, f0, f1, f2, ops);
  FZ_LAUNCHED();
  return FZ_OK;
}

static int periodic_halos(fz_ctx *c, double *f0, double *f1, double *f2) {
  const int nn[3] = {c->d.n1, c->d.n2, c->d.n3};
  for (int dir = 1; dir <= 2; dir++) {
    if (!c->per[dir]) continue;
    DirOps ops;
    memset(&ops, 0, sizeof(ops));
    for (int fi = 0; fi < 3; fi++) {
      push(ops.f[fi], 1, nn[dir] + 1, 1, 0.0, 0.0);
      push(ops.f[fi], 1, 0, nn[dir], 0.0, 0.0);
    }
    FZ_TRY(launch_dir(c, dir, f0, f1, f2, ops));
  }
  return FZ_OK;
}

static bool valid_codes(const char *cbc, int n) {
  for (int q = 0; q < n; q++)
    if (cbc[q] != 'P' && cbc[q] != 'D' && cbc[q] != 'N') return false;
  return true;
}

extern "C" int fz_boundp(fz_ctx *c, const char cbc[6], const double bc[6], double *p) {
  if (!c || !cbc || !bc || !p) return fz_set_error(FZ_ERR_ARG, "fz_boundp: null argument");
  if (!valid_codes(cbc, 6)) return fz_set_error(FZ_ERR_ARG, "fz_boundp: bad boundary code");
  FZ_TRY(periodic_halos(c, p, nullptr, nullptr));
  const int nn[3] = {c->d.n1, c->d.n2, c->d.n3};
  const double dr[3][2] = {{c->dl[0], c->dl[0]}, {c->dl[1], c->dl[1]}, {c->h_dzc[0], c->h_dzc[c->d.n3]}};
  for (int dir = 0; dir < 3; dir++) {
    if (dir > 0 && c->per[dir]) continue;
    DirOps ops;
    memset(&ops, 0, sizeof(ops));
    for (int ib = 0; ib < 2; ib++)
      set_bc_ops(ops.f[0], cbc[2 * dir + ib], ib, nn[dir], true, bc[2 * dir + ib], dr[dir][ib]);
    FZ_TRY(launch_dir(c, dir, p, nullptr, nullptr, ops));
  }
  return FZ_OK;
}

extern "C" int fz_bounduvw(fz_ctx *c, const char cbc[18], const double bc[18], int is_correc, double *u, double *v,
                           double *w) {
  if (!c || !cbc || !bc || !u || !v || !w) return fz_set_error(FZ_ERR_ARG, "fz_bounduvw: null argument");
  if (!valid_codes(cbc, 18)) return fz_set_error(FZ_ERR_ARG, "fz_bounduvw: bad boundary code");
  FZ_TRY(periodic_halos(c, u, v, w));
  const int nn[3] = {c->d.n1, c->d.n2, c->d.n3};
  const double *dzc = c->h_dzc.data(), *dzf = c->h_dzf.data();
#define CB(b, dd, cc) cbc[(b) + 2 * ((dd) + 3 * (cc))]
#define BV(b, dd, cc) bc[(b) + 2 * ((dd) + 3 * (cc))]
  for (int dir = 0; dir < 3; dir++) {
    if (dir > 0 && c->per[dir]) continue;
    // the face-normal component is skipped on the correction call unless
    // periodic, src/bound.f90:46,57,68
    const bool impose = (!is_correc) || (CB(0, dir, dir) == 'P' && CB(1, dir, dir) == 'P');
    DirOps ops;
    memset(&ops, 0, sizeof(ops));
    for (int ib = 0; ib < 2; ib++) {
      for (int comp = 0; comp < 3; comp++) {
        const bool normal = (comp == dir);
        if (normal && !impose) continue;
        double dr;
        if (dir < 2) dr = c->dl[dir];
        else dr = normal ? (ib == 0 ? dzf[0] : dzf[nn[2]]) : (ib == 0 ? dzc[0] : dzc[nn[2]]);
        set_bc_ops(ops.f[comp], CB(ib, dir, comp), ib, nn[dir], !normal, BV(ib, dir, comp), dr);
      }
    }
    FZ_TRY(launch_dir(c, dir, u, v, w, ops));
  }
#undef CB
#undef BV
  return FZ_OK;
}
