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
#include <algorithm>
#include <cstdlib>
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

// T = double: the reference's FP64 fields; T = float: the FP32 phase fields
// of the mixed-precision step (the operations then round in float)
template <typename T>
__global__ void k_plane_ops(fz_dims d, int dir, T *__restrict__ f0, T *__restrict__ f1, T *__restrict__ f2,
                            DirOps ops) {
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
  T *fs[3] = {f0, f1, f2};
  for (int fi = 0; fi < 3; fi++) {
    T *p = fs[fi];
    if (!p) continue;
    const FieldOps &fo = ops.f[fi];
    for (int s = 0; s < fo.nops; s++) {
      const SubOp &o = fo.op[s];
      T val;
      if (o.mode == 1) val = p[base + st * o.src];
      else if (o.mode == 2) val = (T)o.factor + (T)o.sgn * p[base + st * o.src];
      else val = (T)o.factor;
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

template <typename T>
static int launch_dir(fz_ctx *c, int dir, T *f0, T *f1, T *f2, const DirOps &ops, bool periodic_copy) {
  const fz_dims &d = c->d;
  int na, nb;
  if (dir == 0) { na = d.n2 + 2; nb = d.n3 + 2; }
  else if (dir == 1) { na = d.n1 + 2; nb = d.n3 + 2; }
  else { na = d.n1 + 2; nb = d.n2 + 2; }
  static const char *names[2][3] = {{"bc_x", "bc_y", "bc_z"}, {"bc_per_x", "bc_per_y", "bc_per_z"}};
  FZ_LAUNCH(c, names[periodic_copy ? 1 : 0][dir], k_plane_ops<T>, dim3((na + 63) / 64, nb), dim3(64), 0, d, dir, f0,
            f1, f2, ops);
  FZ_LAUNCHED();
  return FZ_OK;
}

// the ordered phases of one boundp / bounduvw call
struct Phase {
  int dir;
  bool periodic_copy;
  DirOps ops;
};

static void periodic_phases(fz_ctx *c, int nfield, std::vector<Phase> &ph) {
  const int nn[3] = {c->d.n1, c->d.n2, c->d.n3};
  for (int dir = 1; dir <= 2; dir++) {
    if (!c->per[dir]) continue;
    Phase P;
    memset(&P, 0, sizeof(P));
    P.dir = dir;
    P.periodic_copy = true;
    for (int fi = 0; fi < nfield; fi++) {
      push(P.ops.f[fi], 1, nn[dir] + 1, 1, 0.0, 0.0);
      push(P.ops.f[fi], 1, 0, nn[dir], 0.0, 0.0);
    }
    ph.push_back(P);
  }
}

static void boundp_phases(fz_ctx *c, const char cbc[6], const double bc[6], std::vector<Phase> &ph) {
  periodic_phases(c, 1, ph);
  const int nn[3] = {c->d.n1, c->d.n2, c->d.n3};
  const double dr[3][2] = {{c->dl[0], c->dl[0]}, {c->dl[1], c->dl[1]}, {c->h_dzc[0], c->h_dzc[c->d.n3]}};
  for (int dir = 0; dir < 3; dir++) {
    if (dir > 0 && c->per[dir]) continue;
    Phase P;
    memset(&P, 0, sizeof(P));
    P.dir = dir;
    for (int ib = 0; ib < 2; ib++)
      set_bc_ops(P.ops.f[0], cbc[2 * dir + ib], ib, nn[dir], true, bc[2 * dir + ib], dr[dir][ib]);
    ph.push_back(P);
  }
}

static void bounduvw_phases(fz_ctx *c, const char cbc[18], const double bc[18], int is_correc,
                            std::vector<Phase> &ph) {
  periodic_phases(c, 3, ph);
  const int nn[3] = {c->d.n1, c->d.n2, c->d.n3};
  const double *dzc = c->h_dzc.data(), *dzf = c->h_dzf.data();
#define CB(b, dd, cc) cbc[(b) + 2 * ((dd) + 3 * (cc))]
#define BV(b, dd, cc) bc[(b) + 2 * ((dd) + 3 * (cc))]
  for (int dir = 0; dir < 3; dir++) {
    if (dir > 0 && c->per[dir]) continue;
    // the face-normal component is skipped on the correction call unless
    // periodic, src/bound.f90:46,57,68
    const bool impose = (!is_correc) || (CB(0, dir, dir) == 'P' && CB(1, dir, dir) == 'P');
    Phase P;
    memset(&P, 0, sizeof(P));
    P.dir = dir;
    for (int ib = 0; ib < 2; ib++) {
      for (int comp = 0; comp < 3; comp++) {
        const bool normal = (comp == dir);
        if (normal && !impose) continue;
        double dr;
        if (dir < 2) dr = c->dl[dir];
        else dr = normal ? (ib == 0 ? dzf[0] : dzf[nn[2]]) : (ib == 0 ? dzc[0] : dzc[nn[2]]);
        set_bc_ops(P.ops.f[comp], CB(ib, dir, comp), ib, nn[dir], !normal, BV(ib, dir, comp), dr);
      }
    }
    ph.push_back(P);
  }
#undef CB
#undef BV
}

template <typename T>
static int run_phases(fz_ctx *c, const std::vector<Phase> &ph, T *f0, T *f1, T *f2) {
  for (const Phase &P : ph) FZ_TRY(launch_dir(c, P.dir, f0, f1, f2, P.ops, P.periodic_copy));
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// The phases composed into one gather per call. Every operation writes one
// cell from one other cell (copy or affine) or from a constant, so the final
// value of every written cell is a chain of at most a few such operations
// applied to one original value. Replaying the phases symbolically on the
// host gives, per written cell, its source cell and its chain; one launch
// then evaluates the chains with the same arithmetic as the planes kernel
// (so the result is bit for bit that of the ordered phases). When some
// source cell is itself written by the call the gather would race, and the
// ordered phases are used instead.
// ---------------------------------------------------------------------------
#define BC_MAXOPS 6
struct BcChain {
  int nops;
  int mode[BC_MAXOPS];
  double factor[BC_MAXOPS], sgn[BC_MAXOPS];
};
struct fz_bcmap {
  bool ok = false;  // false: use the ordered phases
  int n = 0;
  int *dst = nullptr, *src = nullptr;
  unsigned char *cid = nullptr;
  BcChain *chains = nullptr;
};
struct BcMapDev {
  const int *dst, *src;
  const unsigned char *cid;
  const BcChain *chains;
  int n;
};

// up to FZ_GATHER_MAX (map, field) pairs, one per blockIdx.y: the boundary
// updates of several fields in one launch
#define FZ_GATHER_MAX 4
template <typename T>
struct GatherSet {
  BcMapDev m[FZ_GATHER_MAX];
  T *f[FZ_GATHER_MAX];
};

template <typename T>
__global__ void k_bc_gather(GatherSet<T> g) {
  const unsigned y = blockIdx.y;
  const BcMapDev m = y == 0 ? g.m[0] : (y == 1 ? g.m[1] : (y == 2 ? g.m[2] : g.m[3]));
  T *f = y == 0 ? g.f[0] : (y == 1 ? g.f[1] : (y == 2 ? g.f[2] : g.f[3]));
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m.n; t += gridDim.x * blockDim.x) {
    const BcChain &ch = m.chains[m.cid[t]];
    const int sidx = m.src[t];
    T v = sidx >= 0 ? f[sidx] : (T)0;
    for (int o = 0; o < ch.nops; o++) v = (ch.mode[o] == 2) ? (T)ch.factor[o] + (T)ch.sgn[o] * v : (T)ch.factor[o];
    f[m.dst[t]] = v;
  }
}

// the same for a batch of cases (fz_case_batch): blockIdx.z is the case,
// blockIdx.y the (map, role) pair; the maps are those of the first case,
// which every case of the batch shares (same grid and conditions)
template <typename T>
__global__ void k_bc_gather_b(GatherSet<T> g, fz_bref br, int r0, int r1, int r2, int r3) {
  const unsigned y = blockIdx.y, cs = blockIdx.z;
  const BcMapDev m = y == 0 ? g.m[0] : (y == 1 ? g.m[1] : (y == 2 ? g.m[2] : g.m[3]));
  T *f = (T *)fz_brole(br, cs, y == 0 ? r0 : (y == 1 ? r1 : (y == 2 ? r2 : r3)));
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m.n; t += gridDim.x * blockDim.x) {
    const BcChain &ch = m.chains[m.cid[t]];
    const int sidx = m.src[t];
    T v = sidx >= 0 ? f[sidx] : (T)0;
    for (int o = 0; o < ch.nops; o++) v = (ch.mode[o] == 2) ? (T)ch.factor[o] + (T)ch.sgn[o] * v : (T)ch.factor[o];
    f[m.dst[t]] = v;
  }
}

namespace {
// The cells a boundary operation can write: those with i, j or k in the
// layers {0, n, n+1} of their direction (a staggered wall value sits on
// layer n), numbered densely: the three z layers, then the three y rows of
// the other layers, then the three x cells of the other rows.
struct ShellIndex {
  long sy, sz;
  int n1, n2, n3;
  long size;
  explicit ShellIndex(const fz_dims &d) : sy(d.sy), sz(d.sz), n1(d.n1), n2(d.n2), n3(d.n3) {
    size = 3 * sz + 3 * sy * (n3 - 1) + 3 * (long)(n2 - 1) * (n3 - 1);
  }
  static int layer(long x, int n) { return x == 0 ? 0 : (x == n ? 1 : (x == n + 1 ? 2 : -1)); }
  long pos3(long i, long j, long k) const {  // -1 for a cell no operation can write
    const int lz = layer(k, n3);
    if (lz >= 0) return lz * sz + j * sy + i;
    const long kk = k - 1;  // 0 .. n3-2
    long base = 3 * sz;
    const int ly = layer(j, n2);
    if (ly >= 0) return base + (kk * 3 + ly) * sy + i;
    base += 3 * sy * (n3 - 1);
    const int lx = layer(i, n1);
    if (lx >= 0) return base + (kk * (n2 - 1) + (j - 1)) * 3 + lx;
    return -1;
  }
};

// The distinct chains met while composing, each with its successors under
// every operation (memo: chain x operation -> chain, -1 not yet made)
struct ChainTable {
  std::vector<BcChain> ch;
  std::vector<int> memo;
  int nop;
  explicit ChainTable(int nop_) : nop(nop_) { add(BcChain()); }
  int add(const BcChain &c) {
    for (size_t q = 0; q < ch.size(); q++)
      if (memcmp(&ch[q], &c, sizeof(c)) == 0) return (int)q;
    ch.push_back(c);
    memo.resize(ch.size() * nop, -1);
    return (int)ch.size() - 1;
  }
  // chain a followed by operation o (number g); -2 when too long
  int next(int a, const SubOp &o, int g) {
    int &m = memo[(size_t)a * nop + g];
    if (m != -1) return m;
    BcChain c;
    memset(&c, 0, sizeof(c));
    if (o.mode == 3) {
      c.nops = 1;
      c.mode[0] = 3;
      c.factor[0] = o.factor;
      c.sgn[0] = o.sgn;
    } else {
      c = ch[a];
      if (o.mode == 2) {
        if (c.nops == BC_MAXOPS) return memo[(size_t)a * nop + g] = -2;
        c.mode[c.nops] = 2;
        c.factor[c.nops] = o.factor;
        c.sgn[c.nops] = o.sgn;
        c.nops++;
      }
    }
    const int id = add(c);
    memo[(size_t)a * nop + g] = id;  // add() may have moved the memo
    return id;
  }
};
}  // namespace

static void plane_geometry(const fz_dims &d, int dir, int &na, int &nb) {
  if (dir == 0) { na = d.n2 + 2; nb = d.n3 + 2; }
  else if (dir == 1) { na = d.n1 + 2; nb = d.n3 + 2; }
  else { na = d.n1 + 2; nb = d.n2 + 2; }
}

// Each written cell holds (source cell or -1, chain): the phases are applied
// in order, an operation's destination taking its source's pair extended by
// the operation (a source not written so far is its own original value).
// xdone: the kernel that produced the field also wrote, for every interior
// row (j in 1..n2, k in 1..n3), the x halo cells 0 and n1+1 with the values
// of their periodic images n1 and 1; the cells whose composed value is
// exactly that copy are left out of the map.
static fz_bcmap *compose(fz_ctx *c, const std::vector<Phase> &ph, int fi, bool xdone) {
  fz_bcmap *m = new fz_bcmap();
  const fz_dims &d = c->d;
  if (c->nh >= 2147483647L) return m;
  const ShellIndex sh(d);
  std::vector<long> wsrc(sh.size), wsp(sh.size);  // source cell and its shell position (-1: none)
  std::vector<int> wch(sh.size, -1);              // chain, -1: not written
  int nop = 0;
  for (const Phase &P : ph) nop += P.ops.f[fi].nops;
  ChainTable tab(nop > 0 ? nop : 1);
  int g0 = 0;
  size_t nw = 0;  // cells written
  for (const Phase &P : ph) {
    const FieldOps &fo = P.ops.f[fi];
    if (fo.nops == 0) continue;
    int na, nb;
    plane_geometry(d, P.dir, na, nb);
    for (int b = 0; b < nb; b++)
      for (int a = 0; a < na; a++) {
        long base, st;
        if (P.dir == 0) { base = d.sy * a + d.sz * b; st = 1; }
        else if (P.dir == 1) { base = a + d.sz * b; st = d.sy; }
        else { base = a + d.sy * b; st = d.sz; }
        // the cell at offset x along the phase's direction, in shell numbering
        auto spos = [&](long x) {
          return P.dir == 0 ? sh.pos3(x, a, b) : (P.dir == 1 ? sh.pos3(a, x, b) : sh.pos3(a, b, x));
        };
        for (int s = 0; s < fo.nops; s++) {
          const SubOp &o = fo.op[s];
          const long src = base + st * o.src;
          long esrc = -1, esp = -1;
          int ech = 0;
          if (o.mode != 3) {
            const long ps = spos(o.src);
            if (ps >= 0 && wch[ps] >= 0) { esrc = wsrc[ps]; esp = wsp[ps]; ech = wch[ps]; }
            else { esrc = src; esp = ps; }
          }
          ech = tab.next(ech, o, g0 + s);
          if (ech == -2) return m;  // longer than a chain holds: the ordered phases
          const long pd = spos(o.dst);
          if (pd < 0) return m;  // not a boundary layer: the ordered phases
          nw += wch[pd] < 0;
          wsrc[pd] = esrc;
          wsp[pd] = esp;
          wch[pd] = ech;
        }
      }
    g0 += fo.nops;
  }
  // the written cells in increasing linear order
  std::vector<std::pair<long, long>> recs;  // (cell, shell position)
  recs.reserve(nw);
  auto visit = [&](long i, long j, long k) {
    const long p = sh.pos3(i, j, k);
    if (wch[p] >= 0) recs.push_back({k * d.sz + j * d.sy + i, p});
  };
  for (long k = 0; k <= d.n3 + 1; k++)
    for (long j = 0; j <= d.n2 + 1; j++) {
      if (ShellIndex::layer(k, d.n3) >= 0 || ShellIndex::layer(j, d.n2) >= 0) {
        for (long i = 0; i <= d.n1 + 1; i++) visit(i, j, k);
      } else {
        visit(0, j, k);
        visit(d.n1, j, k);
        visit(d.n1 + 1, j, k);
      }
    }
  for (const auto &rp : recs) {
    const long pe = wsp[rp.second];
    if (pe >= 0 && wch[pe] >= 0) return m;  // a written cell is also read: not a gather
  }
  if (xdone) {
    std::vector<std::pair<long, long>> keep;
    keep.reserve(recs.size());
    for (const auto &rp : recs) {
      const long cell = rp.first, k = cell / d.sz, j = (cell - k * d.sz) / d.sy, i = cell - k * d.sz - j * d.sy;
      const bool xrow = (i == 0 || i == d.n1 + 1) && j >= 1 && j <= d.n2 && k >= 1 && k <= d.n3;
      const long image = cell + (i == 0 ? (long)d.n1 : -(long)d.n1);
      if (xrow && tab.ch[wch[rp.second]].nops == 0 && wsrc[rp.second] == image) continue;
      keep.push_back(rp);
    }
    recs.swap(keep);
  }
  // chain numbers in order of first use over the sorted cells
  std::vector<int> renum(tab.ch.size(), -1);
  std::vector<BcChain> chains;
  std::vector<int> hdst(recs.size()), hsrc(recs.size());
  std::vector<unsigned char> hcid(recs.size());
  for (size_t r = 0; r < recs.size(); r++) {
    const long p = recs[r].second;
    int &id = renum[wch[p]];
    if (id < 0) {
      id = (int)chains.size();
      if (id > 255) return m;
      chains.push_back(tab.ch[wch[p]]);
    }
    hdst[r] = (int)recs[r].first;
    hsrc[r] = (int)wsrc[p];
    hcid[r] = (unsigned char)id;
  }
  const int n = (int)recs.size();
  bool good = hipMalloc((void **)&m->dst, sizeof(int) * (n ? n : 1)) == hipSuccess &&
              hipMalloc((void **)&m->src, sizeof(int) * (n ? n : 1)) == hipSuccess &&
              hipMalloc((void **)&m->cid, (n ? n : 1)) == hipSuccess &&
              hipMalloc((void **)&m->chains, sizeof(BcChain) * (chains.empty() ? 1 : chains.size())) == hipSuccess;
  if (good && n) {
    good = hipMemcpy(m->dst, hdst.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(m->src, hsrc.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(m->cid, hcid.data(), n, hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(m->chains, chains.data(), sizeof(BcChain) * chains.size(), hipMemcpyHostToDevice) == hipSuccess;
  }
  m->n = n;
  m->ok = good;
  return m;
}

static void bcmap_free(fz_bcmap *m) {
  if (!m) return;
  if (m->dst) (void)hipFree(m->dst);
  if (m->src) (void)hipFree(m->src);
  if (m->cid) (void)hipFree(m->cid);
  if (m->chains) (void)hipFree(m->chains);
  delete m;
}

void fz_bcmaps_free(fz_ctx *c) {
  for (auto &kv : c->bcmaps) bcmap_free(kv.second);
  c->bcmaps.clear();
}

static fz_bcmap *get_map(fz_ctx *c, const std::string &key, const std::vector<Phase> &ph, int fi,
                         bool xdone = false) {
  auto it = c->bcmaps.find(key);
  if (it != c->bcmaps.end()) return it->second;
  fz_bcmap *m = compose(c, ph, fi, xdone);
  c->bcmaps[key] = m;
  return m;
}

static std::string signature(const char *kind, const char *cbc, int ncbc, const double *bc, int is_correc, int fi,
                             bool xdone = false) {
  std::string k(kind);
  k.append(cbc, ncbc);
  k.append((const char *)bc, sizeof(double) * ncbc);
  k.push_back((char)('0' + is_correc));
  k.push_back((char)('0' + fi));
  k.push_back(xdone ? 'x' : '-');
  return k;
}

static BcMapDev devmap(const fz_bcmap *m) {
  BcMapDev d;
  d.dst = m ? m->dst : nullptr;
  d.src = m ? m->src : nullptr;
  d.cid = m ? m->cid : nullptr;
  d.chains = m ? m->chains : nullptr;
  d.n = m ? m->n : 0;
  return d;
}

// nf (map, field) pairs, FZ_GATHER_MAX per launch
template <typename T>
static int launch_gather(fz_ctx *c, int nf, const fz_bcmap *const *maps, T *const *fields) {
  for (int q0 = 0; q0 < nf; q0 += FZ_GATHER_MAX) {
    const int k = std::min(FZ_GATHER_MAX, nf - q0);
    GatherSet<T> g;
    int nmax = 0;
    for (int q = 0; q < FZ_GATHER_MAX; q++) {
      g.m[q] = devmap(q < k ? maps[q0 + q] : nullptr);
      g.f[q] = q < k ? fields[q0 + q] : nullptr;
      if (q < k) nmax = std::max(nmax, maps[q0 + q]->n);
    }
    if (nmax == 0) continue;
    const int nb = std::min((nmax + 255) / 256, 4096);
    FZ_LAUNCH(c, "bc_gather", k_bc_gather<T>, dim3(nb, k), dim3(256), 0, g);
    FZ_LAUNCHED();
  }
  return FZ_OK;
}

static bool valid_codes(const char *cbc, int n) {
  for (int q = 0; q < n; q++)
    if (cbc[q] != 'P' && cbc[q] != 'D' && cbc[q] != 'N') return false;
  return true;
}

static bool g_gather_off = getenv("FZ_BC_PHASES") != nullptr;  // the ordered phases, for comparison

template <typename T>
static int boundp_t(fz_ctx *c, const char cbc[6], const double bc[6], T *p) {
  std::vector<Phase> ph;
  boundp_phases(c, cbc, bc, ph);
  if (!g_gather_off) {
    fz_bcmap *m = get_map(c, signature("p", cbc, 6, bc, 0, 0), ph, 0);
    if (m->ok) return launch_gather<T>(c, 1, &m, &p);
  }
  return run_phases<T>(c, ph, p, nullptr, nullptr);
}

template <typename T>
static int bounduvw_t(fz_ctx *c, const char cbc[18], const double bc[18], int is_correc, T *u, T *v, T *w) {
  std::vector<Phase> ph;
  bounduvw_phases(c, cbc, bc, is_correc ? 1 : 0, ph);
  if (!g_gather_off) {
    fz_bcmap *m[3];
    bool ok = true;
    for (int fi = 0; fi < 3; fi++) {
      m[fi] = get_map(c, signature("uvw", cbc, 18, bc, is_correc ? 1 : 0, fi), ph, fi);
      ok = ok && m[fi]->ok;
    }
    T *const f[3] = {u, v, w};
    if (ok) return launch_gather<T>(c, 3, m, f);
  }
  return run_phases<T>(c, ph, u, v, w);
}

extern "C" int fz_boundp(fz_ctx *c, const char cbc[6], const double bc[6], double *p) {
  if (!c || !cbc || !bc || !p) return fz_set_error(FZ_ERR_ARG, "fz_boundp: null argument");
  if (!valid_codes(cbc, 6)) return fz_set_error(FZ_ERR_ARG, "fz_boundp: bad boundary code");
  return boundp_t<double>(c, cbc, bc, p);
}

extern "C" int fz_bounduvw(fz_ctx *c, const char cbc[18], const double bc[18], int is_correc, double *u, double *v,
                           double *w) {
  if (!c || !cbc || !bc || !u || !v || !w) return fz_set_error(FZ_ERR_ARG, "fz_bounduvw: null argument");
  if (!valid_codes(cbc, 18)) return fz_set_error(FZ_ERR_ARG, "fz_bounduvw: bad boundary code");
  return bounduvw_t<double>(c, cbc, bc, is_correc, u, v, w);
}

// Several boundary updates in as few launches as the gather allows: items
// of boundp (one field) or bounduvw (three fields) on distinct fields, each
// result as its own call's. Items whose composed map is unavailable run as
// their own calls.
template <typename T>
int fz_bounds_group(fz_ctx *c, int n, const FzBcItem *items) {
  std::vector<fz_bcmap *> maps;
  std::vector<T *> fields;
  for (int q = 0; q < n; q++) {
    const FzBcItem &it = items[q];
    const int ncbc = it.kind == 0 ? 6 : 18;
    if (!valid_codes(it.cbc, ncbc)) return fz_set_error(FZ_ERR_ARG, "boundary update: bad boundary code");
    std::vector<Phase> ph;
    if (it.kind == 0) boundp_phases(c, it.cbc, it.bc, ph);
    else bounduvw_phases(c, it.cbc, it.bc, it.is_correc ? 1 : 0, ph);
    const int nf = it.kind == 0 ? 1 : 3;
    bool ok = !g_gather_off;
    fz_bcmap *m[3] = {nullptr, nullptr, nullptr};
    for (int fi = 0; fi < nf && ok; fi++) {
      m[fi] = get_map(c,
                      signature(it.kind == 0 ? "p" : "uvw", it.cbc, ncbc, it.bc, it.is_correc ? 1 : 0, fi,
                                it.xdone != 0),
                      ph, fi, it.xdone != 0);
      ok = ok && m[fi]->ok;
    }
    if (!ok) {
      FZ_TRY(run_phases<T>(c, ph, (T *)it.f[0], (T *)it.f[1], (T *)it.f[2]));
      continue;
    }
    for (int fi = 0; fi < nf; fi++) {
      maps.push_back(m[fi]);
      fields.push_back((T *)it.f[fi]);
    }
  }
  if (maps.empty()) return FZ_OK;
  return launch_gather<T>(c, (int)maps.size(), maps.data(), fields.data());
}
template int fz_bounds_group<double>(fz_ctx *, int, const FzBcItem *);
template int fz_bounds_group<float>(fz_ctx *, int, const FzBcItem *);

// fz_bounds_group for a batch of ncase same-shaped cases: the items' fields
// are given as roles (roles[3 q + fi] for field fi of item q), each case's
// buffer of the role taken from the batch table. FZ_ERR_UNSUPPORTED when an
// item has no composed map (the ordered phases are not batched).
template <typename T>
int fz_bounds_group_batch(fz_ctx *c, int n, const FzBcItem *items, const int *roles, fz_bref br, int ncase) {
  std::vector<fz_bcmap *> maps;
  std::vector<int> rl;
  for (int q = 0; q < n; q++) {
    const FzBcItem &it = items[q];
    const int ncbc = it.kind == 0 ? 6 : 18;
    if (!valid_codes(it.cbc, ncbc)) return fz_set_error(FZ_ERR_ARG, "boundary update: bad boundary code");
    std::vector<Phase> ph;
    if (it.kind == 0) boundp_phases(c, it.cbc, it.bc, ph);
    else bounduvw_phases(c, it.cbc, it.bc, it.is_correc ? 1 : 0, ph);
    const int nf = it.kind == 0 ? 1 : 3;
    for (int fi = 0; fi < nf; fi++) {
      fz_bcmap *m = g_gather_off ? nullptr
                                 : get_map(c, signature(it.kind == 0 ? "p" : "uvw", it.cbc, ncbc, it.bc,
                                                        it.is_correc ? 1 : 0, fi), ph, fi);
      if (!m || !m->ok) return fz_set_error(FZ_ERR_UNSUPPORTED, "batched boundary update without a composed map");
      maps.push_back(m);
      rl.push_back(roles[3 * q + fi]);
    }
  }
  for (size_t q0 = 0; q0 < maps.size(); q0 += FZ_GATHER_MAX) {
    const int k = (int)std::min((size_t)FZ_GATHER_MAX, maps.size() - q0);
    GatherSet<T> g;
    int nmax = 0, r[FZ_GATHER_MAX] = {0, 0, 0, 0};
    for (int q = 0; q < FZ_GATHER_MAX; q++) {
      g.m[q] = devmap(q < k ? maps[q0 + q] : nullptr);
      g.f[q] = nullptr;
      if (q < k) {
        nmax = std::max(nmax, maps[q0 + q]->n);
        r[q] = rl[q0 + q];
      }
    }
    if (nmax == 0) continue;
    const int nb = std::min((nmax + 255) / 256, 4096);
    FZ_LAUNCH(c, "bc_gather", k_bc_gather_b<T>, dim3(nb, k, ncase), dim3(256), 0, g, br, r[0], r[1], r[2], r[3]);
    FZ_LAUNCHED();
  }
  return FZ_OK;
}
template int fz_bounds_group_batch<double>(fz_ctx *, int, const FzBcItem *, const int *, fz_bref, int);
template int fz_bounds_group_batch<float>(fz_ctx *, int, const FzBcItem *, const int *, fz_bref, int);

// FP32 fields of the mixed-precision step (same sequence and maps)
int fz_boundp_f32(fz_ctx *c, const char cbc[6], const double bc[6], float *p) {
  if (!valid_codes(cbc, 6)) return fz_set_error(FZ_ERR_ARG, "fz_boundp: bad boundary code");
  return boundp_t<float>(c, cbc, bc, p);
}

int fz_bounduvw_f32(fz_ctx *c, const char cbc[18], const double bc[18], int is_correc, float *u, float *v,
                    float *w) {
  if (!valid_codes(cbc, 18)) return fz_set_error(FZ_ERR_ARG, "fz_bounduvw: bad boundary code");
  return bounduvw_t<float>(c, cbc, bc, is_correc, u, v, w);
}
