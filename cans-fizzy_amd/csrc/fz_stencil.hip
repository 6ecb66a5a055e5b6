// Cell stencils of the per-timestep hot path: pressure RHS and projection,
// ACDI and THINC/QQ phase-field transport, Youngs normals and curvature,
// momentum advection/diffusion and the pressure/surface-tension/gravity
// split, AB2 updates, forcing and the initial phase field.
//
// One thread per cell, x fastest across a wavefront so every field load is a
// coalesced 512-byte row segment; neighbours in y and z come from L1/L2.
// Expressions keep the reference's operation order (and the library is built
// without FMA contraction) so that, transcendental functions aside, results
// round exactly as the reference's own CPU build does.
#include <cmath>
#include <cstdlib>

#include "fz_internal.h"
#include "fz_math.h"

// block coordinates in the XCD-aware pencil order
#define FZ_BLOCK_XYZ(d, bx, by, bz) \
  unsigned bx, by, bz;              \
  fz_stencil_block((d).tile_y, bx, by, bz);

// the same in a kernel that may run batched (fz_bref br): the block's case in cs
#define FZ_BLOCK_XYZB(d, br, cs, bx, by, bz) \
  unsigned cs, bx, by, bz;                   \
  fz_stencil_block_b((d).tile_y, BT ? (br).gz : gridDim.z, cs, bx, by, bz);

#define CELL_IJK(d)                                   \
  FZ_BLOCK_XYZ(d, bx_, by_, bz_);                     \
  const int i = bx_ * blockDim.x + threadIdx.x + 1;   \
  const int j = by_ * blockDim.y + threadIdx.y + 1;   \
  const int k = bz_ + 1;                              \
  if (i > (d).n1 || j > (d).n2) return;               \
  const long q = i + (d).sy * j + (d).sz * k;         \
  const long r = (i - 1) + (long)(d).n1 * ((j - 1) + (long)(d).n2 * (k - 1));

#define CELL_IJKB(d, br, cs)                          \
  FZ_BLOCK_XYZB(d, br, cs, bx_, by_, bz_);            \
  const int i = bx_ * blockDim.x + threadIdx.x + 1;   \
  const int j = by_ * blockDim.y + threadIdx.y + 1;   \
  const int k = bz_ + 1;                              \
  if (i > (d).n1 || j > (d).n2) return;               \
  const long q = i + (d).sy * j + (d).sz * k;         \
  const long r = (i - 1) + (long)(d).n1 * ((j - 1) + (long)(d).n2 * (k - 1));

#define HALO_IJK(d)                                   \
  FZ_BLOCK_XYZ(d, bx_, by_, bz_);                     \
  const int i = bx_ * blockDim.x + threadIdx.x;       \
  const int j = by_ * blockDim.y + threadIdx.y;       \
  const int k = bz_;                                  \
  if (i > (d).n1 + 1 || j > (d).n2 + 1) return;       \
  const long q = i + (d).sy * j + (d).sz * k;

// ---------------------------------------------------------------------------
// fillps, src/fillps.f90:33-50
__global__ void k_fillps(fz_dims d, const double *__restrict__ dzfi, double dti, double dtidxi, double dtidyi,
                         double rho0, const double *__restrict__ u, const double *__restrict__ v,
                         const double *__restrict__ w, double *__restrict__ p) {
  CELL_IJK(d);
  (void)r;
  p[q] = rho0 * ((w[q] - w[q - d.sz]) * dti * dzfi[k] + (v[q] - v[q - d.sy]) * dtidyi + (u[q] - u[q - 1]) * dtidxi);
}

// correc, src/correc.f90:34-66 (constant coefficients: rhop = rho0); the
// three loops cover different halo-inclusive ranges and are fused here.
__global__ void k_correc(fz_dims d, const double *__restrict__ dzci, double factori, double factorj, double dt,
                         double rhop, const double *__restrict__ p, double *__restrict__ u, double *__restrict__ v,
                         double *__restrict__ w) {
  HALO_IJK(d);
  if (i <= d.n1) u[q] = u[q] - factori * (p[q + 1] - p[q]) / rhop;
  if (j <= d.n2) v[q] = v[q] - factorj * (p[q + d.sy] - p[q]) / rhop;
  if (k <= d.n3) w[q] = w[q] - dt * dzci[k] * (p[q + d.sz] - p[q]) / rhop;
}

// The end of the projection in one pass over the halo-inclusive grid:
// correc (src/correc.f90:34-66) into the spare velocity buffers un/vn/wn,
// updatep (src/updatep.f90:21-23) into the buffer of the previous pressure,
// and the cell-centred speeds of acdi_set_gamma (src/acdi.f90:32-62) over
// the interior. Faces that correc leaves alone (u at i = n1+1, v at
// j = n2+1, w at k = n3+1) are carried over. An interior thread recomputes
// the corrected velocity of its lower faces with the same expression as
// the thread that owns them, so the speeds are those of the corrected
// field bit for bit. Each block folds the maximum squared speed of its
// cells into one of GAMMA_SLOTS slots with a 64-bit atomic max on the bit
// pattern (which orders non-negative doubles like their values; NaN speeds
// are skipped, as fmax does in the reduction kernel); spreading the blocks
// over the slots keeps same-address atomics from serialising, and
// k_gamma_fin reduces and clears the slots.
__global__ void __launch_bounds__(256) k_correc_step(fz_dims d, const double *__restrict__ dzci, double factori_v,
                                                     double factorj_v, double dt_v, double dxi, double dyi,
                                                     const double *__restrict__ stc, double rhop,
                                                     const double *__restrict__ dp, const double *__restrict__ p,
                                                     const double *__restrict__ u, const double *__restrict__ v,
                                                     const double *__restrict__ w, double *__restrict__ un,
                                                     double *__restrict__ vn, double *__restrict__ wn,
                                                     double *__restrict__ p_out,
                                                     unsigned long long *__restrict__ slots) {
  // dt from the case's device scalars when given (the products are the
  // host's, operand for operand)
  const double dt = stc ? stc[FZ_STC_DT] : dt_v;
  const double factori = stc ? dt * dxi : factori_v, factorj = stc ? dt * dyi : factorj_v;
  FZ_BLOCK_XYZ(d, bx, by, bz);
  const int i = bx * blockDim.x + threadIdx.x, j = by * blockDim.y + threadIdx.y, k = bz;
  double m = 0.0;
  if (i <= d.n1 + 1 && j <= d.n2 + 1) {
    const long sy = d.sy, sz = d.sz;
    const long q = i + sy * j + sz * k;
    const double uq = (i <= d.n1) ? u[q] - factori * (dp[q + 1] - dp[q]) / rhop : u[q];
    const double vq = (j <= d.n2) ? v[q] - factorj * (dp[q + sy] - dp[q]) / rhop : v[q];
    const double wq = (k <= d.n3) ? w[q] - dt * dzci[k] * (dp[q + sz] - dp[q]) / rhop : w[q];
    un[q] = uq;
    vn[q] = vq;
    wn[q] = wq;
    p_out[q] = dp[q] + p[q];
    if (slots && i >= 1 && i <= d.n1 && j >= 1 && j <= d.n2 && k >= 1 && k <= d.n3) {
      const double um = u[q - 1] - factori * (dp[q] - dp[q - 1]) / rhop;
      const double vm = v[q - sy] - factorj * (dp[q] - dp[q - sy]) / rhop;
      const double wm = w[q - sz] - dt * dzci[k - 1] * (dp[q] - dp[q - sz]) / rhop;
      const double uc = 0.5 * (uq + um), vc = 0.5 * (vq + vm), wc = 0.5 * (wq + wm);
      m = fmax(m, uc * uc + vc * vc + wc * wc);
    }
  }
  if (!slots) return;  // uniform across the grid
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  __shared__ double sh[16];
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  const int nw = (int)((blockDim.x * blockDim.y + 63) >> 6);
  if ((t & 63) == 0) sh[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    double b = sh[0];
    for (int q = 1; q < nw; q++) b = fmax(b, sh[q]);
    if (b > 0.0)
      atomicMax(slots + ((blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) & (FZ_GAMMA_SLOTS - 1)),
                (unsigned long long)__double_as_longlong(b));
  }
}

// k_correc_step with each thread marching up KZ cells in z: dp at the cell
// above becomes dp at the cell of the next level, and the cell's new w
// becomes the next cell's lower face velocity for the speed maximum, so the
// planes above and below are read once instead of three times. Same
// expressions, same values.
template <int KZ>
__global__ void __launch_bounds__(256) k_correc_step_mz(fz_dims d, const double *__restrict__ dzci,
                                                        double factori_v, double factorj_v, double dt_v, double dxi,
                                                        double dyi, const double *__restrict__ stc, double rhop,
                                                        const double *__restrict__ dp, const double *__restrict__ p,
                                                        const double *__restrict__ u, const double *__restrict__ v,
                                                        const double *__restrict__ w, double *__restrict__ un,
                                                        double *__restrict__ vn, double *__restrict__ wn,
                                                        double *__restrict__ p_out,
                                                        unsigned long long *__restrict__ slots) {
  const double dt = stc ? stc[FZ_STC_DT] : dt_v;
  const double factori = stc ? dt * dxi : factori_v, factorj = stc ? dt * dyi : factorj_v;
  FZ_BLOCK_XYZ(d, bx, by, bz);
  const int i = bx * blockDim.x + threadIdx.x, j = by * blockDim.y + threadIdx.y;
  const int k0 = bz * KZ, kend = min(k0 + KZ - 1, d.n3 + 1);
  double m = 0.0;
  if (i <= d.n1 + 1 && j <= d.n2 + 1) {
    const long sy = d.sy, sz = d.sz;
    const bool in_ij = i >= 1 && i <= d.n1 && j >= 1 && j <= d.n2;
    long q = i + sy * j + sz * (long)k0;
    double dpc = dp[q];
    double wbelow = 0.0;  // the new w of the cell below (the lower face of this one)
    if (slots && in_ij && k0 >= 1 && k0 <= d.n3)
      wbelow = w[q - sz] - dt * dzci[k0 - 1] * (dpc - dp[q - sz]) / rhop;
    for (int k = k0; k <= kend; k++, q += sz) {
      const double dpz = k <= d.n3 ? dp[q + sz] : 0.0;
      const double uq = (i <= d.n1) ? u[q] - factori * (dp[q + 1] - dpc) / rhop : u[q];
      const double vq = (j <= d.n2) ? v[q] - factorj * (dp[q + sy] - dpc) / rhop : v[q];
      const double wq = (k <= d.n3) ? w[q] - dt * dzci[k] * (dpz - dpc) / rhop : w[q];
      un[q] = uq;
      vn[q] = vq;
      wn[q] = wq;
      p_out[q] = dpc + p[q];
      if (slots && in_ij && k >= 1 && k <= d.n3) {
        const double um = u[q - 1] - factori * (dpc - dp[q - 1]) / rhop;
        const double vm = v[q - sy] - factorj * (dpc - dp[q - sy]) / rhop;
        const double uc = 0.5 * (uq + um), vc = 0.5 * (vq + vm), wc = 0.5 * (wq + wbelow);
        m = fmax(m, uc * uc + vc * vc + wc * wc);
      }
      wbelow = wq;
      dpc = dpz;
    }
  }
  if (!slots) return;  // uniform across the grid
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  __shared__ double sh[16];
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  const int nw = (int)((blockDim.x * blockDim.y + 63) >> 6);
  if ((t & 63) == 0) sh[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    double b = sh[0];
    for (int q = 1; q < nw; q++) b = fmax(b, sh[q]);
    if (b > 0.0)
      atomicMax(slots + ((blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) & (FZ_GAMMA_SLOTS - 1)),
                (unsigned long long)__double_as_longlong(b));
  }
}

// gam = max(sqrt(velmax)*gam_factor, gam_min) (src/acdi.f90:57-60, src/main.f90:526-527)
// from the slots of k_correc_step, which are cleared for the next step
// (clear: a word set to zero, the range flag of the next step's phase field)
template <bool BT = false>
__global__ void __launch_bounds__(FZ_GAMMA_SLOTS) k_gamma_fin(unsigned long long *__restrict__ slots, double factor,
                                                              double gmin, double *__restrict__ out, fz_bref br,
                                                              unsigned *__restrict__ clear) {
  if constexpr (BT) {  // batched: one block per case
    const fz_bcase &B = br.bt[blockIdx.x];
    slots = B.gslots;
    factor = B.gam_factor;
    gmin = B.gam_min;
    out = B.dsc;
  }
  double m = __longlong_as_double((long long)slots[threadIdx.x]);
  slots[threadIdx.x] = 0ull;
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  __shared__ double sh[FZ_GAMMA_SLOTS / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < FZ_GAMMA_SLOTS / 64; q++) m = fmax(m, sh[q]);
    out[0] = fmax(sqrt(m) * factor, gmin);
    if (clear) clear[0] = 0u;
  }
}


int fz_launch_gamma_fin(fz_ctx *c, unsigned long long *slots, double factor, double gmin, double *out,
                        unsigned *clear) {
  FZ_LAUNCH(c, "gamma_fin", k_gamma_fin<false>, dim3(1), dim3(FZ_GAMMA_SLOTS), 0, slots, factor, gmin, out, kNoBRef,
            clear);
  FZ_LAUNCHED();
  return FZ_OK;
}

// updatep, src/updatep.f90:21-23, whole array
__global__ void k_updatep(long nh, const double *__restrict__ pp, double *__restrict__ p) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nh; q += (long)gridDim.x * blockDim.x)
    p[q] = p[q] + pp[q];
}

// updt_rhs_b, src/bound.f90:357-395 (one launch per face)
__global__ void k_add_face(fz_dims d, int dir, int plane, const double *__restrict__ rhs, double *__restrict__ p) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;  // first in-plane index (0-based interior)
  const int b = blockIdx.y;
  int i, j, k, na;
  if (dir == 0) { na = d.n2; i = plane; j = a + 1; k = b + 1; }
  else if (dir == 1) { na = d.n1; i = a + 1; j = plane; k = b + 1; }
  else { na = d.n1; i = a + 1; j = b + 1; k = plane; }
  if (a >= na) return;
  p[i + d.sy * j + d.sz * k] += rhs[a + (long)na * b];
}

// Stencil access to a halo'd field around one cell, (a,b,c) the neighbour
// offsets in x, y, z. PtrAt indexes through a 64-bit address per load;
// BufAt addresses through a buffer resource: one 32-bit byte offset per
// cell in a vector register (the (-1,-1,-1) neighbour), the y/z strides as
// scalar offsets and the x offset as an immediate, which spares the vector
// registers and instructions of a 64-bit address per field and plane. The
// values read are the same; a read outside the field returns 0 instead of
// faulting (the kernels never make one).
template <typename T>
struct PtrAt {
  const T *__restrict__ p;
  long q, sy, sz;
  __device__ __forceinline__ T operator()(int a, int b, int c) const { return p[q + a + b * sy + c * sz]; }
};
template <typename T>
struct BufAt {
  __amdgpu_buffer_rsrc_t r;
  unsigned v0, sy, sz;  // bytes
  __device__ __forceinline__ T operator()(int a, int b, int c) const {
    const unsigned so = (unsigned)(b + 1) * sy + (unsigned)(c + 1) * sz;
    const unsigned vo = v0 + (unsigned)((a + 1) * (int)sizeof(T));
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
  }
};
template <typename T>
__device__ __forceinline__ PtrAt<T> ptr_at(const T *p, const fz_dims &d, long q) {
  return PtrAt<T>{p, q, d.sy, d.sz};
}
// (fields below 4 GiB: the launchers check fits_buffers on the host)
template <typename T>
__device__ __forceinline__ BufAt<T> buf_at(const T *p, const fz_dims &d, long q) {
  const unsigned sy = (unsigned)d.sy, sz = (unsigned)d.sz;
  const unsigned bytes = (unsigned)((unsigned long long)d.sz * (unsigned)(d.n3 + 2) * sizeof(T));
  BufAt<T> b;
  b.r = __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
  b.v0 = ((unsigned)q - 1u - sy - sz) * (unsigned)sizeof(T);
  b.sy = sy * (unsigned)sizeof(T);
  b.sz = sz * (unsigned)sizeof(T);
  return b;
}

// buffer loads at element e of a halo'd field (byte offsets in 32 bits, a
// uniform extra byte offset in soffset; past the field's end they read 0)
template <typename T>
struct BufFld {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ T operator()(unsigned e, unsigned so_bytes) const {
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, e * 8u, so_bytes, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, e * 4u, so_bytes, 0));
  }
};
template <typename T>
__device__ __forceinline__ BufFld<T> buf_fld(const T *p, const fz_dims &d) {
  const unsigned bytes = (unsigned)((unsigned long long)d.sz * (unsigned)(d.n3 + 2) * sizeof(T));
  return BufFld<T>{__builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000)};
}

// whether a halo'd FP64 field is small enough for 32-bit buffer offsets
static int fits_buffers(const fz_ctx *c, const char *who) {
  if ((unsigned long long)c->nh * sizeof(double) < (1ull << 32)) return FZ_OK;
  return fz_set_error(FZ_ERR_UNSUPPORTED, "%s: halo'd fields above 4 GiB are not supported", who);
}

// A kernel that writes the interior cells of a field that is periodic in x
// can write the x halo of the cell's row as well (xper): the halo cells 0
// and n1+1 take the values of the cells n1 and 1, as the periodic boundary
// update would, and that update then leaves them alone. Interior rows only
// (j in 1..n2, k in 1..n3). The halo cell shares its cache line with the
// row's first or last cells, which the same kernel writes, so the line is
// written once instead of read and written again by the boundary update.
// The stores go through buffer resources with 32-bit element offsets (the
// launchers check fits_buffers), which costs a lane one offset instead of a
// 64-bit address per field.
struct XHalo {
  bool lo, hi;  // the cell is n1 (its image is 0) / 1 (its image is n1+1)
  unsigned n1;
};
__device__ __forceinline__ XHalo xhalo_of(const fz_dims &d, int i) {
  return XHalo{i == d.n1, i == 1, (unsigned)d.n1};
}
template <typename T>
struct BufOut {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ void put(unsigned e, T v) const {
    if constexpr (sizeof(T) == 8)
      __builtin_amdgcn_raw_buffer_store_b64(
          __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0u, 0)), v), r, e * 8u, 0u, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b32(
          __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0u, 0u, 0)), v), r, e * 4u, 0u, 0);
  }
};
template <typename T>
__device__ __forceinline__ BufOut<T> buf_out(T *p, const fz_dims &d) {
  const unsigned bytes = (unsigned)((unsigned long long)d.sz * (unsigned)(d.n3 + 2) * sizeof(T));
  return BufOut<T>{__builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000)};
}
// the halo images of cell e (its halo'd index) for up to four fields
template <typename T>
__device__ __forceinline__ void xhalo_put(const XHalo &x, unsigned e, const BufOut<T> &f0, T v0) {
  if (x.lo) f0.put(e - x.n1, v0);
  if (x.hi) f0.put(e + x.n1, v0);
}

// the same after the fact, for kernel variants that do not write the halo:
// one thread per interior row, up to four fields (blockIdx.y)
template <typename T>
__global__ void k_xhalo_copy(fz_dims d, T *__restrict__ f0, T *__restrict__ f1, T *__restrict__ f2,
                             T *__restrict__ f3) {
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  T *f = blockIdx.y == 0 ? f0 : (blockIdx.y == 1 ? f1 : (blockIdx.y == 2 ? f2 : f3));
  if (row >= (long)d.n2 * d.n3 || !f) return;
  const long q0 = d.sy * (row % d.n2 + 1) + d.sz * (row / d.n2 + 1);
  const T a = f[q0 + d.n1], b = f[q0 + 1];
  f[q0] = a;
  f[q0 + d.n1 + 1] = b;
}
template <typename T>
static int xhalo_copy(fz_ctx *c, T *f0, T *f1, T *f2, T *f3) {
  const long rows = (long)c->d.n2 * c->d.n3;
  FZ_LAUNCH(c, "xhalo_copy", k_xhalo_copy<T>, dim3((unsigned)((rows + 255) / 256), 4), dim3(256), 0, c->d, f0, f1,
            f2, f3);
  FZ_LAUNCHED();
  return FZ_OK;
}

// k_correc_step_mz with the correction dp and the velocities u, v staged in
// LDS: a block of 64 x 4 cells (of the halo-inclusive range) marches up 16
// planes and keeps, per field, the planes k and k+1 of its tile with a
// one-cell x/y frame, so each value is read from HBM once instead of through
// five (dp) or two (u, v) neighbour loads, whose address processing bounds
// the one-cell-per-load kernel. Same expressions, same values.
constexpr int COR_MX = 64, COR_PX = COR_MX + 2;
template <int KZ, int MY, bool BT = false>
__global__ void __launch_bounds__(COR_MX *MY) k_correc_lds(fz_dims d, const double *__restrict__ dzci,
                                                              double factori_v, double factorj_v, double dt_v,
                                                              double dxi, double dyi, const double *__restrict__ stc,
                                                              double rhop, const double *__restrict__ dp,
                                                              const double *__restrict__ p,
                                                              const double *__restrict__ u,
                                                              const double *__restrict__ v,
                                                              const double *__restrict__ w, double *__restrict__ un,
                                                              double *__restrict__ vn, double *__restrict__ wn,
                                                              double *__restrict__ p_out,
                                                              unsigned long long *__restrict__ slots, fz_bref br,
                                                              int xper, unsigned rlo) {
  constexpr int NT = COR_MX * MY, PL = COR_PX * (MY + 2), NPT = (PL + NT - 1) / NT;
  __shared__ double s_dp[2 * PL], s_u[2 * PL], s_v[2 * PL];
  FZ_BLOCK_XYZB(d, br, cs, bx, by, bz);
  if constexpr (BT) {  // this block's case of a batch: the momentum step's velocity in the spare buffers
    const fz_bcase &B = br.bt[cs];
    stc = B.stc;
    rhop = B.rho0;
    slots = B.gslots;
    dp = (const double *)fz_brole(br, cs, BR_DP);
    p = (const double *)fz_brole(br, cs, BR_P);
    u = (const double *)fz_brole(br, cs, BR_UN);
    v = (const double *)fz_brole(br, cs, BR_VN);
    w = (const double *)fz_brole(br, cs, BR_WN);
    un = (double *)fz_brole(br, cs, BR_U);
    vn = (double *)fz_brole(br, cs, BR_V);
    wn = (double *)fz_brole(br, cs, BR_W);
    p_out = (double *)fz_brole(br, cs, BR_PO);
  }
  const double dt = stc ? stc[FZ_STC_DT] : dt_v;
  const double factori = stc ? dt * dxi : factori_v, factorj = stc ? dt * dyi : factorj_v;
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + COR_MX * ty;
  // the range: halo to halo, or in the directions of bit d of rlo (whose
  // halo cells the boundary conditions that follow overwrite) 1 to n
  const int xlo = rlo & 1u, ylo = (rlo >> 1) & 1u, zlo = (rlo >> 2) & 1u;
  const int i0 = xlo + (int)bx * COR_MX, j0 = ylo + (int)by * MY;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = zlo + (int)bz * KZ, kend = min(k0 + KZ - 1, d.n3 + 1 - zlo), ktop = min(kend + 1, d.n3 + 1);
  const bool act = i <= d.n1 + 1 - xlo && j <= d.n2 + 1 - ylo;
  const bool in_ij = i >= 1 && i <= d.n1 && j >= 1 && j <= d.n2;
  const int c = (ty + 1) * COR_PX + (tx + 1);
  const BufFld<double> GDP = buf_fld(dp, d), GU = buf_fld(u, d), GV = buf_fld(v, d);
  const XHalo xh = xhalo_of(d, i);
  const BufOut<double> OUN = buf_out(un, d), OVN = buf_out(vn, d), OWN = buf_out(wn, d), OPO = buf_out(p_out, d);
  // halo'd index of plane position pos; frame positions outside the array
  // (no cell uses them: the tiles of the last block column and row hang
  // past the halo-inclusive range) are not loaded
  unsigned gpos[NPT];
  bool inr[NPT];
#pragma unroll
  for (int r = 0; r < NPT; r++) {
    const int pos = min(t + r * NT, PL - 1), py = pos / COR_PX, px = pos - py * COR_PX;
    const int gi = i0 - 1 + px, gj = j0 - 1 + py;
    inr[r] = gi >= 0 && gi <= d.n1 + 1 && gj >= 0 && gj <= d.n2 + 1;
    gpos[r] = (unsigned)gi + (unsigned)d.sy * (unsigned)gj;
  }
  double qd[NPT], qu[NPT], qv[NPT];
  auto load_plane = [&](int kp) {
    const unsigned zo = (unsigned)d.sz * (unsigned)kp;
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      qd[r] = inr[r] ? GDP(gpos[r] + zo, 0) : 0.0;
      qu[r] = inr[r] ? GU(gpos[r] + zo, 0) : 0.0;
      qv[r] = inr[r] ? GV(gpos[r] + zo, 0) : 0.0;
    }
  };
  auto store_plane = [&](int kp) {
    const int so = (kp & 1) * PL;
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      const int pos = t + r * NT;
      if (pos < PL) {
        s_dp[so + pos] = qd[r];
        s_u[so + pos] = qu[r];
        s_v[so + pos] = qv[r];
      }
    }
  };
  load_plane(k0);
  store_plane(k0);
  if (k0 + 1 <= ktop) {
    load_plane(k0 + 1);
    store_plane(k0 + 1);
  }
  double wbelow = 0.0;  // the new w of the cell below (the lower face of this one)
  if (slots && in_ij && k0 >= 1 && k0 <= d.n3) {
    const long q = i + d.sy * j + d.sz * (long)k0;
    wbelow = w[q - d.sz] - dt * dzci[k0 - 1] * (dp[q] - dp[q - d.sz]) / rhop;
  }
  __syncthreads();
  double m = 0.0;
  for (int k = k0; k <= kend; k++) {
    if (k + 2 <= ktop) load_plane(k + 2);  // in flight during this level
    if (act) {
      const long q = i + d.sy * j + d.sz * (long)k;
      const int sa = (k & 1) * PL, sb = ((k + 1) & 1) * PL;
      const double dpc = s_dp[sa + c];
      const double dpz = k <= d.n3 ? s_dp[sb + c] : 0.0;
      const double uq = (i <= d.n1) ? s_u[sa + c] - factori * (s_dp[sa + c + 1] - dpc) / rhop : s_u[sa + c];
      const double vq =
          (j <= d.n2) ? s_v[sa + c] - factorj * (s_dp[sa + c + COR_PX] - dpc) / rhop : s_v[sa + c];
      const double wq = (k <= d.n3) ? w[q] - dt * dzci[k] * (dpz - dpc) / rhop : w[q];
      const double pq = dpc + p[q];
      // (xper: in an interior row the halo cells take the values of their
      // periodic images, which the cells n1 and 1 write)
      const bool xrow = xper && j >= 1 && j <= d.n2 && k >= 1 && k <= d.n3;
      if (!xrow || (i >= 1 && i <= d.n1)) {
        un[q] = uq;
        vn[q] = vq;
        wn[q] = wq;
        p_out[q] = pq;
      }
      if (xrow) {
        xhalo_put(xh, (unsigned)q, OUN, uq);
        xhalo_put(xh, (unsigned)q, OVN, vq);
        xhalo_put(xh, (unsigned)q, OWN, wq);
        xhalo_put(xh, (unsigned)q, OPO, pq);
      }
      if (slots && in_ij && k >= 1 && k <= d.n3) {
        const double um = s_u[sa + c - 1] - factori * (dpc - s_dp[sa + c - 1]) / rhop;
        const double vm = s_v[sa + c - COR_PX] - factorj * (dpc - s_dp[sa + c - COR_PX]) / rhop;
        const double uc = 0.5 * (uq + um), vc = 0.5 * (vq + vm), wc = 0.5 * (wq + wbelow);
        m = fmax(m, uc * uc + vc * vc + wc * wc);
      }
      wbelow = wq;
    }
    __syncthreads();  // the slot of plane k is free
    if (k + 2 <= ktop) store_plane(k + 2);
    __syncthreads();
  }
  if (!slots) return;  // uniform across the grid
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  __shared__ double sh[NT / 64];
  if ((t & 63) == 0) sh[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    double b = sh[0];
    for (int q = 1; q < NT / 64; q++) b = fmax(b, sh[q]);
    if (b > 0.0)
      atomicMax(slots + ((blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)) & (FZ_GAMMA_SLOTS - 1)),
                (unsigned long long)__double_as_longlong(b));
  }
}

int fz_launch_correc_step(fz_ctx *c, double rho0, double dt, const double *stc, const double *dp, const double *p,
                          const double *u, const double *v, const double *w, double *un, double *vn, double *wn,
                          double *p_out, unsigned long long *slots, int xper, unsigned rlo) {
  if (c->correc_kz == 1) {  // dp, u, v staged in LDS
    FZ_TRY(fits_buffers(c, "correc_step"));
    constexpr int KZ = 16;
    const int my = c->correc_my;
    // (rlo: the directions whose range is 1 to n, see k_correc_lds)
    const int ex = 2 - 2 * (int)(rlo & 1u), ey = 2 - 2 * (int)((rlo >> 1) & 1u), ez = 2 - 2 * (int)((rlo >> 2) & 1u);
    const dim3 grid((c->d.n1 + ex + COR_MX - 1) / COR_MX, (c->d.n2 + ey + my - 1) / my, (c->d.n3 + ez + KZ - 1) / KZ);
#define FZ_CORREC_LDS(MY)                                                                                          \
  FZ_LAUNCH(c, "correc_step", (k_correc_lds<KZ, MY>), grid, dim3(COR_MX, MY), 0, c->d, c->dzci, dt * c->dli[0],     \
            dt * c->dli[1], dt, c->dli[0], c->dli[1], stc, rho0, dp, p, u, v, w, un, vn, wn, p_out, slots, kNoBRef, \
            xper, rlo)
    if (my == 16) FZ_CORREC_LDS(16);
    else if (my == 8) FZ_CORREC_LDS(8);
    else FZ_CORREC_LDS(4);
#undef FZ_CORREC_LDS
  } else if (c->correc_kz > 0) {
    dim3 grid = fz_grid_halo(c->d);
    const int kz = c->correc_kz == 8 ? 8 : 16;
    grid.z = (c->d.n3 + 2 + kz - 1) / kz;
    FZ_LAUNCH(c, "correc_step", kz == 8 ? k_correc_step_mz<8> : k_correc_step_mz<16>, grid, fz_block3(), 0, c->d,
              c->dzci, dt * c->dli[0], dt * c->dli[1], dt, c->dli[0], c->dli[1], stc, rho0, dp, p, u, v, w, un, vn,
              wn, p_out, slots);
  } else {
    FZ_LAUNCH(c, "correc_step", k_correc_step, fz_grid_halo(c->d), fz_block3(), 0, c->d, c->dzci, dt * c->dli[0],
              dt * c->dli[1], dt, c->dli[0], c->dli[1], stc, rho0, dp, p, u, v, w, un, vn, wn, p_out, slots);
  }
  FZ_LAUNCHED();
  if (xper && c->correc_kz != 1) return xhalo_copy<double>(c, un, vn, wn, p_out);
  return FZ_OK;
}


// transcendentals by phase-field type: the shared FP64 routines (identical
// on the device and in the oracle), or the device library's FP32 ones
// The FP32 instances (the phase field of the mixed-precision step, compared
// with the FP64 step within stated per-field bounds) use the hardware's
// approximate operations: v_rcp_f32, v_sqrt_f32, and exp/log through
// v_exp_f32/v_log_f32 (base 2), each within about 1 ulp, instead of the
// correctly rounded FP32 sequences; the FP64 instances are plain '/' and
// sqrt (correctly rounded, as the oracle's).
__device__ inline double fz_div(double a, double b) { return a / b; }
__device__ inline float fz_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline double fz_sqrt(double x) { return sqrt(x); }
__device__ inline float fz_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double fz_tanh(double x) { return fzm::tanh(x); }
// tanh(x) = (1 - e) / (1 + e), e = exp(-2|x|) (no overflow for any x)
__device__ inline float fz_tanh(float x) {
  const float e = __expf(-2.0f * fabsf(x));
  return copysignf((1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e), x);
}
__device__ inline double fz_log(double x) { return fzm::log(x); }
__device__ inline float fz_log(float x) { return __logf(x); }

// ---------------------------------------------------------------------------
// acdi_transport_pf at one cell, src/acdi.f90:102-224 (SDF-normal sharpening):
// the right-hand side R and the three face fluxes of the cell's + faces
// T is the type of the phase fields and of the arithmetic: double, or float
// for the mixed-precision step (the FP64 velocities and grid are rounded to T)
template <typename T, class AD, class AP>
__device__ __forceinline__ void acdi_cell(int k, const double *__restrict__ dzci, const double *__restrict__ dzfi,
                                          double dxi_d, double dyi_d, double gam_d, double seps_d, const AD &u,
                                          const AD &v, const AD &w, const AP &normx, const AP &normy,
                                          const AP &normz, const AP &phi, const AP &psi, T &R, T &fx, T &fy,
                                          T &fz) {
  const T dxi = (T)dxi_d, dyi = (T)dyi_d, gam = (T)gam_d, seps = (T)seps_d;
  const T dzfi_c = (T)dzfi[k], dzci_c = (T)dzci[k], dzci_m = (T)dzci[k - 1];
  const T psi_ccm = psi(0, 0, -1), psi_cmc = psi(0, -1, 0), psi_mcc = psi(-1, 0, 0), psi_ccc = psi(0, 0, 0),
          psi_pcc = psi(1, 0, 0), psi_cpc = psi(0, 1, 0), psi_ccp = psi(0, 0, 1);
  const T phi_ccm = phi(0, 0, -1), phi_cmc = phi(0, -1, 0), phi_mcc = phi(-1, 0, 0), phi_ccc = phi(0, 0, 0),
          phi_pcc = phi(1, 0, 0), phi_cpc = phi(0, 1, 0), phi_ccp = phi(0, 0, 1);
  const T u_mcc = (T)u(-1, 0, 0), u_ccc = (T)u(0, 0, 0), v_cmc = (T)v(0, -1, 0), v_ccc = (T)v(0, 0, 0),
          w_ccm = (T)w(0, 0, -1), w_ccc = (T)w(0, 0, 0);
  const T nxc = normx(0, 0, 0), nyc = normy(0, 0, 0), nzc = normz(0, 0, 0);
  const T nx_xm = (T)0.5 * (normx(-1, 0, 0) + nxc), nx_xp = (T)0.5 * (normx(1, 0, 0) + nxc);
  const T nx_ym = (T)0.5 * (normx(0, -1, 0) + nxc), nx_yp = (T)0.5 * (normx(0, 1, 0) + nxc);
  const T nx_zm = (T)0.5 * (normx(0, 0, -1) + nxc), nx_zp = (T)0.5 * (normx(0, 0, 1) + nxc);
  const T ny_xm = (T)0.5 * (normy(-1, 0, 0) + nyc), ny_xp = (T)0.5 * (normy(1, 0, 0) + nyc);
  const T ny_ym = (T)0.5 * (normy(0, -1, 0) + nyc), ny_yp = (T)0.5 * (normy(0, 1, 0) + nyc);
  const T ny_zm = (T)0.5 * (normy(0, 0, -1) + nyc), ny_zp = (T)0.5 * (normy(0, 0, 1) + nyc);
  const T nz_xm = (T)0.5 * (normz(-1, 0, 0) + nzc), nz_xp = (T)0.5 * (normz(1, 0, 0) + nzc);
  const T nz_ym = (T)0.5 * (normz(0, -1, 0) + nzc), nz_yp = (T)0.5 * (normz(0, 1, 0) + nzc);
  const T nz_zm = (T)0.5 * (normz(0, 0, -1) + nzc), nz_zp = (T)0.5 * (normz(0, 0, 1) + nzc);
  const T upsiim = (T)0.5 * (psi_mcc + psi_ccc) * u_mcc, upsiip = (T)0.5 * (psi_pcc + psi_ccc) * u_ccc;
  const T vpsijm = (T)0.5 * (psi_cmc + psi_ccc) * v_cmc, vpsijp = (T)0.5 * (psi_cpc + psi_ccc) * v_ccc;
  const T wpsikm = (T)0.5 * (psi_ccm + psi_ccc) * w_ccm, wpsikp = (T)0.5 * (psi_ccp + psi_ccc) * w_ccc;
  const T adv = dxi * (upsiip - upsiim) + dyi * (vpsijp - vpsijm) + dzfi_c * (wpsikp - wpsikm);
  const T gs = gam * seps;
  const T dpsidxm = gs * (psi_ccc - psi_mcc) * dxi, dpsidxp = gs * (psi_pcc - psi_ccc) * dxi;
  const T dpsidym = gs * (psi_ccc - psi_cmc) * dyi, dpsidyp = gs * (psi_cpc - psi_ccc) * dyi;
  const T dpsidzm = gs * (psi_ccc - psi_ccm) * dzci_m, dpsidzp = gs * (psi_ccp - psi_ccc) * dzci_c;
  const T rn_01 = fz_div(nx_xm, fz_sqrt(nx_xm * nx_xm + ny_xm * ny_xm + nz_xm * nz_xm) + (T)FZ_EPS);
  const T rn_11 = fz_div(nx_xp, fz_sqrt(nx_xp * nx_xp + ny_xp * ny_xp + nz_xp * nz_xp) + (T)FZ_EPS);
  const T rn_02 = fz_div(ny_ym, fz_sqrt(nx_ym * nx_ym + ny_ym * ny_ym + nz_ym * nz_ym) + (T)FZ_EPS);
  const T rn_12 = fz_div(ny_yp, fz_sqrt(nx_yp * nx_yp + ny_yp * ny_yp + nz_yp * nz_yp) + (T)FZ_EPS);
  const T rn_03 = fz_div(nz_zm, fz_sqrt(nx_zm * nx_zm + ny_zm * ny_zm + nz_zm * nz_zm) + (T)FZ_EPS);
  const T rn_13 = fz_div(nz_zp, fz_sqrt(nx_zp * nx_zp + ny_zp * ny_zp + nz_zp * nz_zp) + (T)FZ_EPS);
  const T g4 = (T)0.25 * gam;
  T t;
  t = fz_tanh(fz_div((T)0.25 * (phi_ccc + phi_mcc), seps)); const T sharpxm = g4 * (((T)1.0 - t * t) * rn_01);
  t = fz_tanh(fz_div((T)0.25 * (phi_pcc + phi_ccc), seps)); const T sharpxp = g4 * (((T)1.0 - t * t) * rn_11);
  t = fz_tanh(fz_div((T)0.25 * (phi_ccc + phi_cmc), seps)); const T sharpym = g4 * (((T)1.0 - t * t) * rn_02);
  t = fz_tanh(fz_div((T)0.25 * (phi_cpc + phi_ccc), seps)); const T sharpyp = g4 * (((T)1.0 - t * t) * rn_12);
  t = fz_tanh(fz_div((T)0.25 * (phi_ccc + phi_ccm), seps)); const T sharpzm = g4 * (((T)1.0 - t * t) * rn_03);
  t = fz_tanh(fz_div((T)0.25 * (phi_ccp + phi_ccc), seps)); const T sharpzp = g4 * (((T)1.0 - t * t) * rn_13);
  const T rglrxm = dpsidxm - sharpxm, rglrxp = dpsidxp - sharpxp;
  const T rglrym = dpsidym - sharpym, rglryp = dpsidyp - sharpyp;
  const T rglrzm = dpsidzm - sharpzm, rglrzp = dpsidzp - sharpzp;
  const T rglr = (rglrxp - rglrxm) * dxi + (rglryp - rglrym) * dyi + (rglrzp - rglrzm) * dzfi_c;
  R = -adv + rglr;
  fx = upsiip - rglrxp;
  fy = vpsijp - rglryp;
  fz = wpsikp - rglrzp;
}

__global__ void __launch_bounds__(256) k_acdi_transport(fz_dims d, const double *__restrict__ dzci,
                                                        const double *__restrict__ dzfi, double dxi, double dyi,
                                                        const double *__restrict__ d_gam, double gam_v, double seps,
                                                        const double *__restrict__ u, const double *__restrict__ v,
                                                        const double *__restrict__ w, const double *__restrict__ normx,
                                                        const double *__restrict__ normy,
                                                        const double *__restrict__ normz,
                                                        const double *__restrict__ phi, const double *__restrict__ psi,
                                                        double *__restrict__ dpsidt, double *__restrict__ flux_x,
                                                        double *__restrict__ flux_y, double *__restrict__ flux_z) {
  CELL_IJK(d);
  const double gam = d_gam ? *d_gam : gam_v;
  double R, fx, fy, fz;
  acdi_cell<double>(k, dzci, dzfi, dxi, dyi, gam, seps, ptr_at(u, d, q), ptr_at(v, d, q), ptr_at(w, d, q),
                    ptr_at(normx, d, q), ptr_at(normy, d, q), ptr_at(normz, d, q), ptr_at(phi, d, q),
                    ptr_at(psi, d, q), R, fx, fy, fz);
  dpsidt[r] = R;
  flux_x[q] = fx;
  flux_y[q] = fy;
  flux_z[q] = fz;
}

// rk_2fl's phase-field part in one pass (src/rk.f90:226-252): acdi_transport_pf,
// the AB2 update psi = psio + f1*R + f2*R_old and the clip to [0,1] (the
// reference also clips the halos, which the boundary update that follows
// overwrites from the interior). The history dpsio is read and replaced by
// R in place; on the first step the history is R itself.
template <typename T>
__global__ void __launch_bounds__(256) k_acdi_step(fz_dims d, const double *__restrict__ dzci,
                                                   const double *__restrict__ dzfi, double dxi, double dyi,
                                                   const double *__restrict__ d_gam, double gam_v, double seps,
                                                   double factor1_v, double factor2_v,
                                                   const double *__restrict__ stc, int first,
                                                   const double *__restrict__ u, const double *__restrict__ v,
                                                   const double *__restrict__ w, const T *__restrict__ normx,
                                                   const T *__restrict__ normy, const T *__restrict__ normz,
                                                   const T *__restrict__ phi, const T *__restrict__ psio,
                                                   T *__restrict__ dpsio, T *__restrict__ psi_out,
                                                   T *__restrict__ flux_x, T *__restrict__ flux_y,
                                                   T *__restrict__ flux_z) {
  CELL_IJK(d);
  const double gam = d_gam ? *d_gam : gam_v;
  const double factor1 = stc ? stc[FZ_STC_F1] : factor1_v, factor2 = stc ? stc[FZ_STC_F2] : factor2_v;
  T R, fx, fy, fz;
  const BufAt<T> PSIO = buf_at(psio, d, q);
  acdi_cell<T>(k, dzci, dzfi, dxi, dyi, gam, seps, buf_at(u, d, q), buf_at(v, d, q), buf_at(w, d, q),
               buf_at(normx, d, q), buf_at(normy, d, q), buf_at(normz, d, q), buf_at(phi, d, q), PSIO, R, fx, fy, fz);
  const T Ro = first ? R : dpsio[r];
  dpsio[r] = R;
  flux_x[q] = fx;
  flux_y[q] = fy;
  flux_z[q] = fz;
  const T pn = PSIO(0, 0, 0) + (T)factor1 * R + (T)factor2 * Ro;
  psi_out[q] = fmin(fmax((T)0.0, pn), (T)1.0);
}

// k_acdi_step with its eight fields staged in LDS (the address unit, not the
// arithmetic, was the busier of the two in the one-cell-per-thread kernel:
// TA_TA_BUSY 84 % against 72 % for the VALU at 256x256x512). A block of
// 64 x 4 cells marches up 16 planes and keeps, per field, the planes k and
// k+1 of its tile with a one-cell x/y frame; the stencil reads below the
// cell are the cell's own values of the level before, kept in registers,
// and the reads above it come from plane k+1. The cell arithmetic is
// acdi_cell's through the accessor LdsZ, so the results are the same bit for
// bit.
constexpr int ACD_MX = 64, ACD_MY = 4, ACD_PX = ACD_MX + 2, ACD_PL = ACD_PX * (ACD_MY + 2);

// a field's planes k, k+1 in LDS seen from one cell, and its value below
// the cell (acdi_cell reads only (0,0,-1) below and (0,0,+1) above)
template <typename T>
struct LdsZ {
  const T *base;
  int c, s1, s2;
  T below;
  __device__ __forceinline__ T operator()(int a, int b, int cc) const {
    if (cc < 0) return below;
    return base[(cc == 0 ? s1 : s2) + c + a + b * ACD_PX];
  }
};

template <typename T, int KZ, bool BT = false>  // BT: a batched launch (fz_bref)
__global__ void __launch_bounds__(ACD_MX *ACD_MY) k_acdi_lds(fz_dims d, const double *__restrict__ dzci,
                                                            const double *__restrict__ dzfi, double dxi, double dyi,
                                                            const double *__restrict__ d_gam, double gam_v,
                                                            double seps, double factor1_v, double factor2_v,
                                                            const double *__restrict__ stc, int first,
                                                            const double *__restrict__ u, const double *__restrict__ v,
                                                            const double *__restrict__ w, const T *__restrict__ normx,
                                                            const T *__restrict__ normy, const T *__restrict__ normz,
                                                            const T *__restrict__ phi, const T *__restrict__ psio,
                                                            T *__restrict__ dpsio, T *__restrict__ psi_out,
                                                            T *__restrict__ flux_x, T *__restrict__ flux_y,
                                                            T *__restrict__ flux_z, fz_bref br) {
  constexpr int NT = ACD_MX * ACD_MY, PL = ACD_PL, NPT = (PL + NT - 1) / NT;
  __shared__ double s_u[2 * PL], s_v[2 * PL], s_w[2 * PL];
  __shared__ T s_nx[2 * PL], s_ny[2 * PL], s_nz[2 * PL], s_phi[2 * PL], s_psi[2 * PL];
  FZ_BLOCK_XYZB(d, br, cs, bx, by, bz);
  if constexpr (BT) {  // this block's case of a batch
    d_gam = br.bt[cs].dsc;
    seps = br.bt[cs].seps;
    stc = br.bt[cs].stc;
    first = 0;
    u = (const double *)fz_brole(br, cs, BR_U);
    v = (const double *)fz_brole(br, cs, BR_V);
    w = (const double *)fz_brole(br, cs, BR_W);
    normx = (const T *)fz_brole(br, cs, BR_NX);
    normy = (const T *)fz_brole(br, cs, BR_NY);
    normz = (const T *)fz_brole(br, cs, BR_NZ);
    phi = (const T *)fz_brole(br, cs, BR_PHI);
    psio = (const T *)fz_brole(br, cs, BR_PSI_OLD);
    dpsio = (T *)fz_brole(br, cs, BR_DPSIO);
    psi_out = (T *)fz_brole(br, cs, BR_PSI_NEW);
    flux_x = (T *)fz_brole(br, cs, BR_FX);
    flux_y = (T *)fz_brole(br, cs, BR_FY);
    flux_z = (T *)fz_brole(br, cs, BR_FZ);
  }
  const double gam = d_gam ? *d_gam : gam_v;
  const double factor1 = stc ? stc[FZ_STC_F1] : factor1_v, factor2 = stc ? stc[FZ_STC_F2] : factor2_v;
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + ACD_MX * ty;
  const int i0 = (int)bx * ACD_MX + 1, j0 = (int)by * ACD_MY + 1, k0 = (int)bz * KZ + 1;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = i <= d.n1 && j <= d.n2;
  const int kend = min(k0 + KZ - 1, d.n3);
  const int c = (ty + 1) * ACD_PX + (tx + 1);
  const BufFld<double> GU = buf_fld(u, d), GV = buf_fld(v, d), GW = buf_fld(w, d);
  const BufFld<T> GNX = buf_fld(normx, d), GNY = buf_fld(normy, d), GNZ = buf_fld(normz, d), GPH = buf_fld(phi, d),
                  GPS = buf_fld(psio, d);
  unsigned gpos[NPT];
#pragma unroll
  for (int r = 0; r < NPT; r++) {
    const int pos = min(t + r * NT, PL - 1), py = pos / ACD_PX, px = pos - py * ACD_PX;
    gpos[r] = (unsigned)(i0 - 1 + px) + (unsigned)d.sy * (unsigned)(j0 - 1 + py);
  }
  double qu[NPT], qv[NPT], qw[NPT];
  T qx[NPT], qy[NPT], qz[NPT], qf[NPT], qs[NPT];
  auto load_plane = [&](int kp) {
    const unsigned zo = (unsigned)d.sz * (unsigned)kp;
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      const unsigned e = gpos[r] + zo;
      qu[r] = GU(e, 0);
      qv[r] = GV(e, 0);
      qw[r] = GW(e, 0);
      qx[r] = GNX(e, 0);
      qy[r] = GNY(e, 0);
      qz[r] = GNZ(e, 0);
      qf[r] = GPH(e, 0);
      qs[r] = GPS(e, 0);
    }
  };
  auto store_plane = [&](int kp) {
    const int so = (kp & 1) * PL;
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      const int pos = t + r * NT;
      if (pos < PL) {
        s_u[so + pos] = qu[r];
        s_v[so + pos] = qv[r];
        s_w[so + pos] = qw[r];
        s_nx[so + pos] = qx[r];
        s_ny[so + pos] = qy[r];
        s_nz[so + pos] = qz[r];
        s_phi[so + pos] = qf[r];
        s_psi[so + pos] = qs[r];
      }
    }
  };
  // the values below the first level, then the planes k0 and k0+1
  double wb = 0.0;
  T xb = (T)0, yb = (T)0, zb = (T)0, fb = (T)0, sb_ = (T)0;
  if (act) {
    const unsigned e = (unsigned)i + (unsigned)d.sy * (unsigned)j + (unsigned)d.sz * (unsigned)(k0 - 1);
    wb = GW(e, 0);
    xb = GNX(e, 0);
    yb = GNY(e, 0);
    zb = GNZ(e, 0);
    fb = GPH(e, 0);
    sb_ = GPS(e, 0);
  }
  load_plane(k0);
  store_plane(k0);
  load_plane(k0 + 1);
  store_plane(k0 + 1);
  __syncthreads();
  for (int k = k0; k <= kend; k++) {
    if (k < kend) load_plane(k + 2);  // in flight during this level
    const int s1 = (k & 1) * PL, s2 = ((k + 1) & 1) * PL;
    if (act) {
      const long q = i + d.sy * j + d.sz * (long)k;
      const long r = (i - 1) + (long)d.n1 * ((j - 1) + (long)d.n2 * (k - 1));
      const LdsZ<double> U{s_u, c, s1, s2, 0.0}, V{s_v, c, s1, s2, 0.0}, W{s_w, c, s1, s2, wb};
      const LdsZ<T> NX{s_nx, c, s1, s2, xb}, NY{s_ny, c, s1, s2, yb}, NZ{s_nz, c, s1, s2, zb},
          PH{s_phi, c, s1, s2, fb}, PSIO{s_psi, c, s1, s2, sb_};
      T R, fx, fy, fz;
      acdi_cell<T>(k, dzci, dzfi, dxi, dyi, gam, seps, U, V, W, NX, NY, NZ, PH, PSIO, R, fx, fy, fz);
      const T Ro = first ? R : dpsio[r];
      dpsio[r] = R;
      flux_x[q] = fx;
      flux_y[q] = fy;
      flux_z[q] = fz;
      const T pn = PSIO(0, 0, 0) + (T)factor1 * R + (T)factor2 * Ro;
      psi_out[q] = fmin(fmax((T)0.0, pn), (T)1.0);
      // this level's values: below the next one
      wb = s_w[s1 + c];
      xb = s_nx[s1 + c];
      yb = s_ny[s1 + c];
      zb = s_nz[s1 + c];
      fb = s_phi[s1 + c];
      sb_ = s_psi[s1 + c];
    }
    __syncthreads();  // the slot of plane k is free
    if (k < kend) store_plane(k + 2);
    __syncthreads();
  }
}

// One face of acdi_transport_pf (src/acdi.f90:102-224) between the cells a
// (below) and b (above) along direction DIR (0 x, 1 y, 2 z): the advective
// flux 0.5 (psi_b + psi_a) vel and the regularisation flux
// gs (psi_b - psi_a) dfac - g4 (1 - t^2) rn, with the face normal
// (n_a + n_b) / 2 and t = tanh(0.25 (phi_b + phi_a) / seps). acdi_cell
// evaluates these expressions at each face from both of its cells, and the
// two evaluations differ only in the operand order of sums of two terms, so
// a face evaluated once gives both cells' values bit for bit.
template <int DIR, typename T>
__device__ __forceinline__ void acdi_face(T psa, T psb, T pha, T phb, T nxa, T nxb, T nya, T nyb, T nza, T nzb, T vel,
                                          T dfac, T gs, T g4, T seps, T &upsi, T &rglr) {
  upsi = (T)0.5 * (psb + psa) * vel;
  const T dpsi = gs * (psb - psa) * dfac;
  const T nxf = (T)0.5 * (nxb + nxa), nyf = (T)0.5 * (nyb + nya), nzf = (T)0.5 * (nzb + nza);
  const T nf = DIR == 0 ? nxf : (DIR == 1 ? nyf : nzf);
  const T rn = fz_div(nf, fz_sqrt(nxf * nxf + nyf * nyf + nzf * nzf) + (T)FZ_EPS);
  const T t = fz_tanh(fz_div((T)0.25 * (phb + pha), seps));
  const T sharp = (T)g4 * (((T)1.0 - t * t) * rn);
  rglr = dpsi - sharp;
}

// The ACDI step with each face evaluated once (acdi_cell evaluates all six
// faces of a cell, so every face twice). The block of 64 x 4 cells marches
// up 16 planes with the normals, phi and the old psi of the planes k, k+1 of
// its tile and a one-cell frame in LDS; at each level every cell evaluates
// its three upper faces (x, y from LDS, z with the plane above), keeps the
// z face in registers for the level above, and puts the x and y faces in
// LDS; the faces below the tile's first row and column are evaluated by the
// first and second waves. After a barrier each cell assembles its
// right-hand side from its six faces in acdi_cell's order, so the results
// are those of k_acdi_lds bit for bit. The velocities are read at the cell
// only (and below the tile's edges), straight from memory.
template <typename T, int KZ, bool XPER = false, int MY = ACD_MY>
__global__ void __launch_bounds__(ACD_MX *MY) k_acdi_sf(fz_dims d, const double *__restrict__ dzci,
                                                           const double *__restrict__ dzfi, double dxi_d,
                                                           double dyi_d, const double *__restrict__ d_gam,
                                                           double gam_v, double seps_d, double factor1_v,
                                                           double factor2_v, const double *__restrict__ stc,
                                                           int first, const double *__restrict__ u,
                                                           const double *__restrict__ v, const double *__restrict__ w,
                                                           const T *__restrict__ normx, const T *__restrict__ normy,
                                                           const T *__restrict__ normz, const T *__restrict__ phi,
                                                           const T *__restrict__ psio, T *__restrict__ dpsio,
                                                           T *__restrict__ psi_out, T *__restrict__ flux_x,
                                                           T *__restrict__ flux_y, T *__restrict__ flux_z, int xper) {
  constexpr int NT = ACD_MX * MY, PL = ACD_PX * (MY + 2), NPT = (PL + NT - 1) / NT;
  __shared__ T s_nx[2 * PL], s_ny[2 * PL], s_nz[2 * PL], s_phi[2 * PL], s_psi[2 * PL];
  // the faces of a level: x faces of the tile's rows (index 0 the face below
  // the first column, tx+1 the face above cell tx), y faces of its columns
  // (row 0 the faces below the first row, ty+1 those above row ty);
  // advective and regularisation parts
  __shared__ T f_xu[MY][ACD_MX + 1], f_xr[MY][ACD_MX + 1];
  __shared__ T f_yu[MY + 1][ACD_MX], f_yr[MY + 1][ACD_MX];
  unsigned cs_, bx, by, bz;
  fz_stencil_block_b(d.tile_y, gridDim.z, cs_, bx, by, bz);
  const double gam_dd = d_gam ? *d_gam : gam_v;
  const double factor1 = stc ? stc[FZ_STC_F1] : factor1_v, factor2 = stc ? stc[FZ_STC_F2] : factor2_v;
  const T dxi = (T)dxi_d, dyi = (T)dyi_d, gam = (T)gam_dd, seps = (T)seps_d;
  const T gs = gam * seps, g4 = (T)0.25 * gam;
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + ACD_MX * ty;
  const int i0 = (int)bx * ACD_MX + 1, j0 = (int)by * MY + 1, k0 = (int)bz * KZ + 1;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = i <= d.n1 && j <= d.n2;
  const int kend = min(k0 + KZ - 1, d.n3);
  const int c = (ty + 1) * ACD_PX + (tx + 1);
  const BufFld<double> GU = buf_fld(u, d), GV = buf_fld(v, d), GW = buf_fld(w, d);
  const BufFld<T> GNX = buf_fld(normx, d), GNY = buf_fld(normy, d), GNZ = buf_fld(normz, d), GPH = buf_fld(phi, d),
                  GPS = buf_fld(psio, d);
  unsigned gpos[NPT];
#pragma unroll
  for (int r = 0; r < NPT; r++) {
    const int pos = min(t + r * NT, PL - 1), py = pos / ACD_PX, px = pos - py * ACD_PX;
    gpos[r] = (unsigned)(i0 - 1 + px) + (unsigned)d.sy * (unsigned)(j0 - 1 + py);
  }
  T qx[NPT], qy[NPT], qz[NPT], qf[NPT], qs[NPT];
  auto load_plane = [&](int kp) {
    const unsigned zo = (unsigned)d.sz * (unsigned)kp;
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      const unsigned e = gpos[r] + zo;
      qx[r] = GNX(e, 0);
      qy[r] = GNY(e, 0);
      qz[r] = GNZ(e, 0);
      qf[r] = GPH(e, 0);
      qs[r] = GPS(e, 0);
    }
  };
  auto store_plane = [&](int kp) {
    const int so = (kp & 1) * PL;
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      const int pos = t + r * NT;
      if (pos < PL) {
        s_nx[so + pos] = qx[r];
        s_ny[so + pos] = qy[r];
        s_nz[so + pos] = qz[r];
        s_phi[so + pos] = qf[r];
        s_psi[so + pos] = qs[r];
      }
    }
  };
  load_plane(k0);
  store_plane(k0);
  load_plane(k0 + 1);
  store_plane(k0 + 1);
  // the face below the first level, from the cell's values of the plane below
  T zu = (T)0, zr = (T)0;
  if (act) {
    const unsigned e = (unsigned)i + (unsigned)d.sy * (unsigned)j + (unsigned)d.sz * (unsigned)(k0 - 1);
    const T wb = (T)GW(e, 0), xb = GNX(e, 0), yb = GNY(e, 0), zb = GNZ(e, 0), fb = GPH(e, 0), sb = GPS(e, 0);
    const unsigned ec = e + (unsigned)d.sz;
    acdi_face<2, T>(sb, GPS(ec, 0), fb, GPH(ec, 0), xb, GNX(ec, 0), yb, GNY(ec, 0), zb, GNZ(ec, 0), wb,
                    (T)dzci[k0 - 1], gs, g4, seps, zu, zr);
  }
  __syncthreads();
  for (int k = k0; k <= kend; k++) {
    if (k < kend) load_plane(k + 2);  // in flight during this level
    const int s1 = (k & 1) * PL, s2 = ((k + 1) & 1) * PL;
    const unsigned ek = (unsigned)d.sz * (unsigned)k;
    T xu = (T)0, xr = (T)0, yu = (T)0, yr = (T)0, wu = (T)0, wr = (T)0, pso = (T)0;
    if (act) {  // the cell's upper faces
      const unsigned e = (unsigned)i + (unsigned)d.sy * (unsigned)j + ek;
      const int a = s1 + c;
      pso = s_psi[a];
      acdi_face<0, T>(s_psi[a], s_psi[a + 1], s_phi[a], s_phi[a + 1], s_nx[a], s_nx[a + 1], s_ny[a], s_ny[a + 1],
                      s_nz[a], s_nz[a + 1], (T)GU(e, 0), dxi, gs, g4, seps, xu, xr);
      acdi_face<1, T>(s_psi[a], s_psi[a + ACD_PX], s_phi[a], s_phi[a + ACD_PX], s_nx[a], s_nx[a + ACD_PX], s_ny[a],
                      s_ny[a + ACD_PX], s_nz[a], s_nz[a + ACD_PX], (T)GV(e, 0), dyi, gs, g4, seps, yu, yr);
      const int b = s2 + c;
      acdi_face<2, T>(s_psi[a], s_psi[b], s_phi[a], s_phi[b], s_nx[a], s_nx[b], s_ny[a], s_ny[b], s_nz[a], s_nz[b],
                      (T)GW(e, 0), (T)dzci[k], gs, g4, seps, wu, wr);
      f_xu[ty][tx + 1] = xu;
      f_xr[ty][tx + 1] = xr;
      f_yu[ty + 1][tx] = yu;
      f_yr[ty + 1][tx] = yr;
    }
    if (ty == 0 && i <= d.n1) {  // the y face below the tile's first row, column tx
      const int a = s1 + c - ACD_PX;
      const unsigned e = (unsigned)i + (unsigned)d.sy * (unsigned)(j0 - 1) + ek;
      T fu, fr;
      acdi_face<1, T>(s_psi[a], s_psi[a + ACD_PX], s_phi[a], s_phi[a + ACD_PX], s_nx[a], s_nx[a + ACD_PX], s_ny[a],
                      s_ny[a + ACD_PX], s_nz[a], s_nz[a + ACD_PX], (T)GV(e, 0), dyi, gs, g4, seps, fu, fr);
      f_yu[0][tx] = fu;
      f_yr[0][tx] = fr;
    }
    if (ty == 1 && tx < MY && j0 + tx <= d.n2) {  // the x face left of the tile's first column, row tx
      const int a = s1 + (tx + 1) * ACD_PX;
      const unsigned e = (unsigned)(i0 - 1) + (unsigned)d.sy * (unsigned)(j0 + tx) + ek;
      T fu, fr;
      acdi_face<0, T>(s_psi[a], s_psi[a + 1], s_phi[a], s_phi[a + 1], s_nx[a], s_nx[a + 1], s_ny[a], s_ny[a + 1],
                      s_nz[a], s_nz[a + 1], (T)GU(e, 0), dxi, gs, g4, seps, fu, fr);
      f_xu[tx][0] = fu;
      f_xr[tx][0] = fr;
    }
    __syncthreads();  // the faces of the level are in LDS; the ring's slot of plane k is free
    if (act) {
      const long q = i + d.sy * j + d.sz * (long)k;
      const long r = (i - 1) + (long)d.n1 * ((j - 1) + (long)d.n2 * (k - 1));
      const T dzfi_c = (T)dzfi[k];
      const T upsiim = f_xu[ty][tx], rglrxm = f_xr[ty][tx], vpsijm = f_yu[ty][tx], rglrym = f_yr[ty][tx];
      const T adv = dxi * (xu - upsiim) + dyi * (yu - vpsijm) + dzfi_c * (wu - zu);
      const T rglr = (xr - rglrxm) * dxi + (yr - rglrym) * dyi + (wr - zr) * dzfi_c;
      const T R = -adv + rglr;
      const T Ro = first ? R : dpsio[r];
      dpsio[r] = R;
      const T fx = xu - xr, fy = yu - yr, fz = wu - wr;
      flux_x[q] = fx;
      flux_y[q] = fy;
      flux_z[q] = fz;
      const T pn = pso + (T)factor1 * R + (T)factor2 * Ro;
      const T ps = fmin(fmax((T)0.0, pn), (T)1.0);
      psi_out[q] = ps;
      if (XPER && xper && (i == 1 || i == d.n1)) {  // the periodic x halo of the row
        const XHalo xh = xhalo_of(d, i);
        xhalo_put(xh, (unsigned)q, buf_out(flux_x, d), fx);
        xhalo_put(xh, (unsigned)q, buf_out(flux_y, d), fy);
        xhalo_put(xh, (unsigned)q, buf_out(flux_z, d), fz);
        xhalo_put(xh, (unsigned)q, buf_out(psi_out, d), ps);
      }
      zu = wu;  // this level's upper z face: the face below the next level
      zr = wr;
    }
    if (k < kend) store_plane(k + 2);
    __syncthreads();  // the plane k+2 is in place and the face arrays are free
  }
}

template <typename T>
int fz_launch_acdi_step(fz_ctx *c, const double *d_gam, double gam, double seps, double factor1, double factor2,
                        const double *stc, int first, const double *u, const double *v, const double *w,
                        const T *normx, const T *normy, const T *normz, const T *phi, const T *psio, T *dpsio,
                        T *psi_out, T *flux_x, T *flux_y, T *flux_z, int xper) {
  FZ_TRY(fits_buffers(c, "acdi_step"));
  // 0: each face evaluated once (3), at 256x256x512 FP64 928 against
  // 1000-1040 us for the LDS kernel of six faces per cell (1), FP32 547
  // against 557 us; the LDS kernel measured FP32 619 against 663 us and
  // FP64 1047 against 1120 us for the plain one (2)
  const int variant = c->acdi_variant ? c->acdi_variant : 3;
  if (variant == 3) {  // each face evaluated once
    constexpr int KZ = 16;
    // (the x-halo stores compiled only into the instance that makes them; a
    // run-time test there too: measured 550 against 565 us in FP32 without
    // it. FP32 tiles of 64 x 8: 550 -> 532 us; FP64 ones, 70 KB of LDS, 925
    // against 923 us, keep 64 x 4)
    constexpr int MY = sizeof(T) == 4 ? 8 : ACD_MY;
    const dim3 grid((c->d.n1 + ACD_MX - 1) / ACD_MX, (c->d.n2 + MY - 1) / MY, (c->d.n3 + KZ - 1) / KZ);
    auto kern = xper ? k_acdi_sf<T, KZ, true, MY> : k_acdi_sf<T, KZ, false, MY>;
    FZ_LAUNCH(c, sizeof(T) == 4 ? "acdi_step_f32" : "acdi_step", kern, grid, dim3(ACD_MX, MY), 0, c->d, c->dzci,
              c->dzfi, c->dli[0], c->dli[1], d_gam, gam, seps, factor1, factor2, stc, first, u, v, w, normx, normy,
              normz, phi, psio, dpsio, psi_out, flux_x, flux_y, flux_z, xper);
  } else if (variant == 1) {  // fields staged in LDS
    constexpr int KZ = 16;
    const dim3 grid((c->d.n1 + ACD_MX - 1) / ACD_MX, (c->d.n2 + ACD_MY - 1) / ACD_MY, (c->d.n3 + KZ - 1) / KZ);
    FZ_LAUNCH(c, sizeof(T) == 4 ? "acdi_step_f32" : "acdi_step", (k_acdi_lds<T, KZ>), grid, dim3(ACD_MX, ACD_MY), 0,
              c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1], d_gam, gam, seps, factor1, factor2, stc, first, u, v, w,
              normx, normy, normz, phi, psio, dpsio, psi_out, flux_x, flux_y, flux_z, kNoBRef);
  } else {
    FZ_LAUNCH(c, sizeof(T) == 4 ? "acdi_step_f32" : "acdi_step", k_acdi_step<T>, fz_grid_interior(c->d),
              fz_block3(), 0, c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1], d_gam, gam, seps, factor1, factor2, stc,
              first, u, v, w, normx, normy, normz, phi, psio, dpsio, psi_out, flux_x, flux_y, flux_z);
  }
  FZ_LAUNCHED();
  // (the other variants leave the x halo to a copy)
  if (xper && variant != 3) return xhalo_copy<T>(c, flux_x, flux_y, flux_z, psi_out);
  return FZ_OK;
}
template int fz_launch_acdi_step<double>(fz_ctx *, const double *, double, double, double, double, const double *,
                                         int, const double *, const double *, const double *, const double *,
                                         const double *, const double *, const double *, const double *, double *,
                                         double *, double *, double *, double *, int);
template int fz_launch_acdi_step<float>(fz_ctx *, const double *, double, double, double, double, const double *,
                                        int, const double *, const double *, const double *, const float *,
                                        const float *, const float *, const float *, const float *, float *, float *,
                                        float *, float *, float *, int);

// acdi_cmpt_phi, src/acdi.f90:239-247 (halo-inclusive)
// (rflag: set when some phi is neither zero nor of a magnitude in
// [2^-100, 2^20], the range k_youngs needs to skip its per-lane checks)
template <typename T, bool BT = false>
__global__ void k_acdi_phi(long nh, double seps_d, const T *__restrict__ psi, T *__restrict__ phi, fz_bref br,
                           unsigned *__restrict__ rflag) {
  if constexpr (BT) {  // batched: the case is blockIdx.y
    const unsigned cs = blockIdx.y;
    seps_d = br.bt[cs].seps;
    psi = (const T *)fz_brole(br, cs, BR_PSI_NEW);
    phi = (T *)fz_brole(br, cs, BR_PHI);
  }
  const T seps = (T)seps_d;
  bool bad = false;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nh; q += (long)gridDim.x * blockDim.x) {
    const T f = seps * fz_log(fz_div(psi[q] + (T)FZ_EPS, (T)1.0 - psi[q] + (T)FZ_EPS));
    phi[q] = f;
    if (rflag) bad = bad || !(f == (T)0 || (fabs(f) >= (T)0x1p-100 && fabs(f) <= (T)0x1p20));
  }
  if (rflag && __ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(rflag, 1u);
}

// epsilon(1._rp) of the phase-field type (the reference's single-precision
// build uses epsilon(1._sp) where its double build uses epsilon(1._dp))
template <typename T>
__device__ __forceinline__ T fz_eps_of() {
  return sizeof(T) == 4 ? (T)1.1920928955078125e-07 : (T)FZ_EPS;
}
__device__ inline double fz_exp(double x) { return fzm::exp(x); }
__device__ inline float fz_exp(float x) { return __expf(x); }

// vof_thinc_cmpt_phi, src/vof_thinc_qq.f90:148-156 (T = float: the mixed step)
template <typename T>
__global__ void k_vof_phi(long nh, double beta_d, const T *__restrict__ psi, T *__restrict__ phi) {
  const T beta = (T)beta_d, eps = fz_eps_of<T>();
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nh; q += (long)gridDim.x * blockDim.x) {
    const T pa = fmin(fmax(psi[q], eps), (T)1.0 - eps);
    phi[q] = (T)0.5 / beta * fz_log(fz_div(pa, (T)1.0 - pa));
  }
}

// vof_thinc_transport_psi fluxes, src/vof_thinc_qq.f90:66-121 (faces 0..n)
template <typename T>
__global__ void k_vof_flux(fz_dims d, double beta_d, const double *__restrict__ u, const double *__restrict__ v,
                           const double *__restrict__ w, const T *__restrict__ normx, const T *__restrict__ normy,
                           const T *__restrict__ normz, const T *__restrict__ phi, T *__restrict__ flux_x,
                           T *__restrict__ flux_y, T *__restrict__ flux_z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > d.n1 || j > d.n2) return;
  const T beta = (T)beta_d;
  const long q = i + d.sy * j + d.sz * k;
  const T vel_x = (T)u[q], vel_y = (T)v[q], vel_z = (T)w[q];
  long qx, qy, qz;
  T xx, yy, zz;
  if (vel_x > (T)0.0) { qx = q; xx = (T)0.5; } else { qx = q + 1; xx = (T)-0.5; }
  if (vel_y > (T)0.0) { qy = q; yy = (T)0.5; } else { qy = q + d.sy; yy = (T)-0.5; }
  if (vel_z > (T)0.0) { qz = q; zz = (T)0.5; } else { qz = q + d.sz; zz = (T)-0.5; }
  const T lfx = fz_div((T)1.0, (T)1.0 + fz_exp((T)-2.0 * beta * (normx[qx] * xx + phi[qx])));
  const T lfy = fz_div((T)1.0, (T)1.0 + fz_exp((T)-2.0 * beta * (normy[qy] * yy + phi[qy])));
  const T lfz = fz_div((T)1.0, (T)1.0 + fz_exp((T)-2.0 * beta * (normz[qz] * zz + phi[qz])));
  flux_x[q] = vel_x * lfx;
  flux_y[q] = vel_y * lfy;
  flux_z[q] = vel_z * lfz;
}

// THINC/QQ face flux at face index f of the x (dir 0), y (1) or z (2)
// direction, the expression of k_vof_flux
template <typename T>
__device__ __forceinline__ T vof_face_flux(const fz_dims &d, int dir, long f, T beta, const double *__restrict__ vel,
                                           const T *__restrict__ nrm, const T *__restrict__ phi) {
  const T vv = (T)vel[f];
  const long up = dir == 0 ? 1 : (dir == 1 ? d.sy : d.sz);
  long qq;
  T hh;
  if (vv > (T)0.0) { qq = f; hh = (T)0.5; } else { qq = f + up; hh = (T)-0.5; }
  const T lf = fz_div((T)1.0, (T)1.0 + fz_exp((T)-2.0 * beta * (nrm[qq] * hh + phi[qq])));
  return vv * lf;
}

// rk_2fl's THINC/QQ phase-field part in one pass (src/rk.f90:226-252 with
// vof_thinc_transport_psi, src/vof_thinc_qq.f90:39-132): the three face
// fluxes of the cell's + faces (stored) and - faces (recomputed with the
// same expressions as the neighbour that stores them), the divergence, the
// AB2 update psi = psio + f1*R + f2*R_old and the clip to [0,1]. The
// history dpsio is read and replaced by R in place (on the first step the
// history is R itself). Fluxes of the faces at index 0 are left to the
// boundary update that follows, which sets every halo face.
template <typename T>
__global__ void __launch_bounds__(256) k_vof_step(fz_dims d, const double *__restrict__ dzfi, double dxi_d,
                                                  double dyi_d, double beta_d, double factor1_v, double factor2_v,
                                                  const double *__restrict__ stc, int first,
                                                  const double *__restrict__ u,
                                                  const double *__restrict__ v, const double *__restrict__ w,
                                                  const T *__restrict__ normx, const T *__restrict__ normy,
                                                  const T *__restrict__ normz, const T *__restrict__ phi,
                                                  const T *__restrict__ psio, T *__restrict__ dpsio,
                                                  T *__restrict__ psi_out, T *__restrict__ flux_x,
                                                  T *__restrict__ flux_y, T *__restrict__ flux_z) {
  CELL_IJK(d);
  const double factor1 = stc ? stc[FZ_STC_F1] : factor1_v, factor2 = stc ? stc[FZ_STC_F2] : factor2_v;
  const T beta = (T)beta_d, dxi = (T)dxi_d, dyi = (T)dyi_d;
  const T fxp = vof_face_flux<T>(d, 0, q, beta, u, normx, phi), fxm = vof_face_flux<T>(d, 0, q - 1, beta, u, normx, phi);
  const T fyp = vof_face_flux<T>(d, 1, q, beta, v, normy, phi),
          fym = vof_face_flux<T>(d, 1, q - d.sy, beta, v, normy, phi);
  const T fzp = vof_face_flux<T>(d, 2, q, beta, w, normz, phi),
          fzm = vof_face_flux<T>(d, 2, q - d.sz, beta, w, normz, phi);
  flux_x[q] = fxp;
  flux_y[q] = fyp;
  flux_z[q] = fzp;
  const T R = -(dxi * (fxp - fxm) + dyi * (fyp - fym) + (T)dzfi[k] * (fzp - fzm));
  const T Ro = first ? R : dpsio[r];
  dpsio[r] = R;
  const T pn = psio[q] + (T)factor1 * R + (T)factor2 * Ro;
  psi_out[q] = fmin(fmax((T)0.0, pn), (T)1.0);
}

template <typename T>
int fz_launch_vof_step(fz_ctx *c, double beta, double factor1, double factor2, const double *stc, int first,
                       const double *u, const double *v, const double *w, const T *normx, const T *normy,
                       const T *normz, const T *phi, const T *psio, T *dpsio, T *psi_out, T *flux_x, T *flux_y,
                       T *flux_z) {
  FZ_LAUNCH(c, sizeof(T) == 4 ? "vof_step_f32" : "vof_step", k_vof_step<T>, fz_grid_interior(c->d), fz_block3(), 0,
            c->d, c->dzfi, c->dli[0], c->dli[1], beta, factor1, factor2, stc, first, u, v, w, normx, normy, normz,
            phi, psio, dpsio, psi_out, flux_x, flux_y, flux_z);
  FZ_LAUNCHED();
  return FZ_OK;
}
template int fz_launch_vof_step<double>(fz_ctx *, double, double, double, const double *, int, const double *,
                                        const double *, const double *, const double *, const double *,
                                        const double *, const double *, const double *, double *, double *, double *,
                                        double *, double *);
template int fz_launch_vof_step<float>(fz_ctx *, double, double, double, const double *, int, const double *,
                                       const double *, const double *, const float *, const float *, const float *,
                                       const float *, const float *, float *, float *, float *, float *, float *);

// flux divergence, src/vof_thinc_qq.f90:123-132
template <typename T>
__global__ void k_vof_div(fz_dims d, const double *__restrict__ dzfi, double dxi_d, double dyi_d,
                          const T *__restrict__ flux_x, const T *__restrict__ flux_y, const T *__restrict__ flux_z,
                          T *__restrict__ dpsidt) {
  CELL_IJK(d);
  const T dxi = (T)dxi_d, dyi = (T)dyi_d;
  dpsidt[r] = -(dxi * (flux_x[q] - flux_x[q - 1]) + dyi * (flux_y[q] - flux_y[q - d.sy]) +
                (T)dzfi[k] * (flux_z[q] - flux_z[q - d.sz]));
}

// clip_field, src/two_fluid.f90:54-57
template <typename T>
__global__ void k_clip(long nh, double rmin, double rmax, T *__restrict__ p) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nh; q += (long)gridDim.x * blockDim.x)
    p[q] = fmin(fmax((T)rmin, p[q]), (T)rmax);
}

// 0.25 * x for a normal x of magnitude at least 2^-1020, as a decrement of the
// exponent field (integer arithmetic: a uniform x stays on the scalar unit)
__device__ inline double quarter_normal(double x) {
  return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) - (2ull << 52));
}

// sqrt(s) by the device's correctly rounded FP64 sequence without its range
// scaling: the same result for s = 0 and s in [2^-767, 2^1023]
__device__ inline double sqrt_inrange(double s) {
  const double y0 = __builtin_amdgcn_rsq(s);
  double g = s * y0, h = y0 * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double e = __builtin_fma(-g, g, s);
  g = __builtin_fma(e, h, g);
  e = __builtin_fma(-g, g, s);
  g = __builtin_fma(e, h, g);
  return s == 0.0 ? s : g;
}

// (x, y, z) / (sqrt(x*x + y*y + z*z) + eps), src/two_fluid.f90:140-162.
// FAST (FP64 only): the square root as sqrt_inrange and the three divisions
// with one reciprocal of the shared denominator (fzm::div_inrange's sequence
// per numerator, then the device's division fix-up, which gives the signed
// zero of a zero numerator). Bit for bit the plain sqrt and '/' when the sum
// is at most 2^600 and every non-zero component is at least 2^-300 in
// magnitude (the denominator then lies in [eps, 2^301]); ok is cleared for a
// lane outside that range, whose caller must take the plain path.
// (NOCHECK: the caller knows the operands are in the fast sequences' range
// and that no component is -0; the fix-up then changes nothing and is left out:
// for finite non-zero operands it only restores the quotient's sign, and a +0
// numerator gives +0 through the sequence by itself)
template <bool FAST, typename T, bool NOCHECK = false>
__device__ __forceinline__ void unit3(T &x, T &y, T &z, bool &ok) {
  if constexpr (FAST && sizeof(T) == 8) {
    const double s = x * x + y * y + z * z;
    if constexpr (!NOCHECK) {
      const int e = min(min(__builtin_amdgcn_frexp_exp(x), __builtin_amdgcn_frexp_exp(y)),
                        __builtin_amdgcn_frexp_exp(z));
      ok = ok && s <= 0x1p600 && e >= -299;
    }
    const double d = sqrt_inrange(s) + (double)FZ_EPS;
    double w = __builtin_amdgcn_rcp(d);
    double t = __builtin_fma(-d, w, 1.0);
    w = __builtin_fma(w, t, w);
    t = __builtin_fma(-d, w, 1.0);
    w = __builtin_fma(w, t, w);
    const double qx = x * w, qy = y * w, qz = z * w;
    if constexpr (NOCHECK) {
      x = __builtin_fma(__builtin_fma(-d, qx, x), w, qx);
      y = __builtin_fma(__builtin_fma(-d, qy, y), w, qy);
      z = __builtin_fma(__builtin_fma(-d, qz, z), w, qz);
    } else {
      x = __builtin_amdgcn_div_fixup(__builtin_fma(__builtin_fma(-d, qx, x), w, qx), d, x);
      y = __builtin_amdgcn_div_fixup(__builtin_fma(__builtin_fma(-d, qy, y), w, qy), d, y);
      z = __builtin_amdgcn_div_fixup(__builtin_fma(__builtin_fma(-d, qz, z), w, qz), d, z);
    }
  } else {
    const T nrm = fz_sqrt(x * x + y * y + z * z) + (T)FZ_EPS;
    x = fz_div(x, nrm);
    y = fz_div(y, nrm);
    z = fz_div(z, nrm);
  }
}

// cmpt_norm_curv_youngs, src/two_fluid.f90:81-174, at one cell: the 27-point
// phase field through the accessor PS (T = float: the mixed-precision step,
// FP32 fields and arithmetic). FAST: the normalisations by unit3's fast
// sequences; ok is cleared where they do not apply.
template <typename T, bool FAST = false, bool NOCHECK = false, class A>
__device__ __forceinline__ void youngs_cell(const A &PS, double dxi_d, double dyi_d, double dzci_k, double dzci_km,
                                            double dzfi_k, T &onx, T &ony, T &onz, T &okap, bool &ok) {
  const T dxi = (T)dxi_d, dyi = (T)dyi_d;
#define P_(a, b, c) PS(a, b, c)
  const T mmm = P_(-1, -1, -1), mcm = P_(-1, 0, -1), mpm = P_(-1, 1, -1), cmm = P_(0, -1, -1), ccm = P_(0, 0, -1),
               cpm = P_(0, 1, -1), pmm = P_(1, -1, -1), pcm = P_(1, 0, -1), ppm = P_(1, 1, -1);
  const T mmc = P_(-1, -1, 0), mcc = P_(-1, 0, 0), mpc = P_(-1, 1, 0), cmc = P_(0, -1, 0), ccc = P_(0, 0, 0),
               cpc = P_(0, 1, 0), pmc = P_(1, -1, 0), pcc = P_(1, 0, 0), ppc = P_(1, 1, 0);
  const T mmp = P_(-1, -1, 1), mcp = P_(-1, 0, 1), mpp = P_(-1, 1, 1), cmp = P_(0, -1, 1), ccp = P_(0, 0, 1),
               cpp = P_(0, 1, 1), pmp = P_(1, -1, 1), pcp = P_(1, 0, 1), ppp = P_(1, 1, 1);
#undef P_
  // NOCHECK (FP64): every difference of sums below is zero or at least
  // 2^-152 in magnitude, so 0.25 * a is exact and 0.25 * a * f equals
  // a * (0.25 * f) bit for bit (0.25 * f is exact for the grid factors in
  // [2^-20, 2^20] the launcher requires): one product per component, with
  // the quartered factors made on the scalar unit
  constexpr bool FOLD = NOCHECK && sizeof(T) == 8;
  T fx = dxi, fy = dyi, zc0 = (T)dzci_k, zcm = (T)dzci_km, fzf = (T)dzfi_k;
  if constexpr (FOLD) {
    fx = quarter_normal(dxi_d);
    fy = quarter_normal(dyi_d);
    zc0 = quarter_normal(dzci_k);
    zcm = quarter_normal(dzci_km);
    fzf = quarter_normal(dzfi_k);
  }
  auto S = [](T a, T f) { return FOLD ? a * f : (T)0.25 * a * f; };
  T mx[8], my[8], mz[8];
  if constexpr (sizeof(T) == 4 && FAST) {
    // FP32 (the mixed-precision step, whose phase-field kernels are held to
    // error bounds, not to the reference's bits): the same vertex gradients
    // from sums over the faces' quads, each shared by the corners it serves.
    // A gradient component is the difference of two 4-value quads of one
    // pair of planes: pairs of values along z make the quads in x and y
    // planes, pairs along x the quads in z planes.
    T zp[3][3];          // z pairs at column (i, j) of one k pair
    T xl[3][3], xu[3][3];  // x pairs (i-1, i) and (i, i+1) at (j, k)
    T qx[3][2][2], qy[3][2][2], qz[3][2][2];
    for (int h = 0; h < 2; h++) {  // h: the k pair (k-1, k) or (k, k+1)
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) zp[a][b] = PS(a - 1, b - 1, h - 1) + PS(a - 1, b - 1, h);
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int q = 0; q < 2; q++) {
          qx[a][q][h] = zp[a][q] + zp[a][q + 1];  // x plane a, j pair q
          qy[a][q][h] = zp[q][a] + zp[q + 1][a];  // y plane a, i pair q
        }
    }
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        xl[b][c] = PS(-1, b - 1, c - 1) + PS(0, b - 1, c - 1);
        xu[b][c] = PS(0, b - 1, c - 1) + PS(1, b - 1, c - 1);
      }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int q = 0; q < 2; q++) {
        qz[c][0][q] = xl[q][c] + xl[q + 1][c];  // z plane c, i pair (i-1, i), j pair q
        qz[c][1][q] = xu[q][c] + xu[q + 1][c];
      }
    const T fx4 = (T)0.25 * dxi, fy4 = (T)0.25 * dyi, fz0 = (T)0.25 * zc0, fzm = (T)0.25 * zcm;
#pragma unroll
    for (int m = 0; m < 8; m++) {  // corner m: (+,+,+) (+,-,+) (+,+,-) (+,-,-), then the same with -x
      const int sx = m < 4, sy = (m & 1) == 0, sz = (m & 2) == 0;
      mx[m] = (sx ? qx[2][sy][sz] - qx[1][sy][sz] : qx[1][sy][sz] - qx[0][sy][sz]) * fx4;
      my[m] = (sy ? qy[2][sx][sz] - qy[1][sx][sz] : qy[1][sx][sz] - qy[0][sx][sz]) * fy4;
      mz[m] = sz ? (qz[2][sx][sy] - qz[1][sx][sy]) * fz0 : (qz[1][sx][sy] - qz[0][sx][sy]) * fzm;
    }
  } else {
  mx[0] = S((pcc + ppc + pcp + ppp) - (ccc + cpc + ccp + cpp), fx);
  mx[1] = S((pcc + pmc + pcp + pmp) - (ccc + cmc + ccp + cmp), fx);
  mx[2] = S((pcc + ppc + pcm + ppm) - (ccc + cpc + ccm + cpm), fx);
  mx[3] = S((pcc + pmc + pcm + pmm) - (ccc + cmc + ccm + cmm), fx);
  mx[4] = S((ccc + cpc + ccp + cpp) - (mcc + mpc + mcp + mpp), fx);
  mx[5] = S((ccc + cmc + ccp + cmp) - (mcc + mmc + mcp + mmp), fx);
  mx[6] = S((ccc + cpc + ccm + cpm) - (mcc + mpc + mcm + mpm), fx);
  mx[7] = S((ccc + cmc + ccm + cmm) - (mcc + mmc + mcm + mmm), fx);
  my[0] = S((cpc + ppc + cpp + ppp) - (ccc + pcc + ccp + pcp), fy);
  my[1] = S((ccc + pcc + ccp + pcp) - (cmc + pmc + cmp + pmp), fy);
  my[2] = S((cpc + ppc + cpm + ppm) - (ccc + pcc + ccm + pcm), fy);
  my[3] = S((ccc + pcc + ccm + pcm) - (cmc + pmc + cmm + pmm), fy);
  my[4] = S((cpc + mpc + cpp + mpp) - (ccc + mcc + ccp + mcp), fy);
  my[5] = S((ccc + mcc + ccp + mcp) - (cmc + mmc + cmp + mmp), fy);
  my[6] = S((cpc + mpc + cpm + mpm) - (ccc + mcc + ccm + mcm), fy);
  my[7] = S((ccc + mcc + ccm + mcm) - (cmc + mmc + cmm + mmm), fy);
  mz[0] = S((ccp + pcp + cpp + ppp) - (ccc + pcc + cpc + ppc), zc0);
  mz[1] = S((ccp + pcp + cmp + pmp) - (ccc + pcc + cmc + pmc), zc0);
  mz[2] = S((ccc + pcc + cpc + ppc) - (ccm + pcm + cpm + ppm), zcm);
  mz[3] = S((ccc + pcc + cmc + pmc) - (ccm + pcm + cmm + pmm), zcm);
  mz[4] = S((ccp + mcp + cpp + mpp) - (ccc + mcc + cpc + mpc), zc0);
  mz[5] = S((ccp + mcp + cmp + mmp) - (ccc + mcc + cmc + mmc), zc0);
  mz[6] = S((ccc + mcc + cpc + mpc) - (ccm + mcm + cpm + mpm), zcm);
  mz[7] = S((ccc + mcc + cmc + mmc) - (ccm + mcm + cmm + mmm), zcm);
  }
#pragma unroll
  for (int m = 0; m < 8; m++) unit3<FAST, T, NOCHECK>(mx[m], my[m], mz[m], ok);
  T nx = (T).125 * (mx[0] + mx[1] + mx[2] + mx[3] + mx[4] + mx[5] + mx[6] + mx[7]);
  T ny = (T).125 * (my[0] + my[1] + my[2] + my[3] + my[4] + my[5] + my[6] + my[7]);
  T nz = (T).125 * (mz[0] + mz[1] + mz[2] + mz[3] + mz[4] + mz[5] + mz[6] + mz[7]);
  unit3<FAST, T, NOCHECK>(nx, ny, nz, ok);
  onx = nx;
  ony = ny;
  onz = nz;
  // (the differences here are zero or at least 2^-268 in magnitude on the NOCHECK path)
  okap = -(S((mx[0] + mx[1] + mx[2] + mx[3]) - (mx[4] + mx[5] + mx[6] + mx[7]), fx) +
           S((my[0] + my[2] + my[4] + my[6]) - (my[1] + my[3] + my[5] + my[7]), fy) +
           S((mz[0] + mz[1] + mz[4] + mz[5]) - (mz[2] + mz[3] + mz[6] + mz[7]), fzf));
}

template <typename T, bool BT = false>
__global__ void __launch_bounds__(256) k_youngs(fz_dims d, const double *__restrict__ dzci,
                                                const double *__restrict__ dzfi, double dxi_d, double dyi_d,
                                                const T *__restrict__ psi, T *__restrict__ normx,
                                                T *__restrict__ normy, T *__restrict__ normz,
                                                T *__restrict__ kappa, fz_bref br, int xper,
                                                const unsigned *__restrict__ rflag) {
  CELL_IJKB(d, br, cs);
  if constexpr (BT) {  // this block's case of a batch
    psi = (const T *)fz_brole(br, cs, BR_PHI);
    normx = (T *)fz_brole(br, cs, BR_NX);
    normy = (T *)fz_brole(br, cs, BR_NY);
    normz = (T *)fz_brole(br, cs, BR_NZ);
    kappa = (T *)fz_brole(br, cs, BR_KAPPA);
  }
  (void)r;
  T nx, ny, nz, ka;
  bool ok = true;
  if (rflag && *rflag == 0u) {
    // every phi of the array is zero or of a magnitude in [2^-100, 2^20]
    // (k_acdi_phi), and the launcher has checked dxi, dyi and dzci against
    // [2^-20, 2^20]: then every vertex-normal component is zero or above
    // 2^-174 in magnitude (a non-zero difference of sums of such phi is a
    // multiple of 2^-152) and below 2^41, each normalised component zero or
    // above 2^-216, each cell-normal component zero or above 2^-271, all
    // within the range of the fast sequences, whose per-lane checks are
    // then left out
    youngs_cell<T, true, true>(buf_at(psi, d, q), dxi_d, dyi_d, dzci[k], dzci[k - 1], dzfi[k], nx, ny, nz, ka,
                               ok);
  } else {
    youngs_cell<T, true>(buf_at(psi, d, q), dxi_d, dyi_d, dzci[k], dzci[k - 1], dzfi[k], nx, ny, nz, ka, ok);
    if (!ok) {  // the plain sequences, from a fresh read of the phase field (an opaque copy of the
                // pointer, so that the 27 values need not stay in registers for this rare path)
      const T *ps = psi;
      asm volatile("" : "+s"(ps));
      youngs_cell<T>(buf_at(ps, d, q), dxi_d, dyi_d, dzci[k], dzci[k - 1], dzfi[k], nx, ny, nz, ka, ok);
    }
  }
  normx[q] = nx;
  normy[q] = ny;
  normz[q] = nz;
  kappa[q] = ka;
  if (xper) {  // the periodic x halo of the row
    const XHalo xh = xhalo_of(d, i);
    xhalo_put(xh, (unsigned)q, buf_out(normx, d), nx);
    xhalo_put(xh, (unsigned)q, buf_out(normy, d), ny);
    xhalo_put(xh, (unsigned)q, buf_out(normz, d), nz);
    xhalo_put(xh, (unsigned)q, buf_out(kappa, d), ka);
  }
}

// the 27 phase-field values of a cell held in registers: planes k-1, k, k+1
// of its 3 x 3 column neighbourhood
template <typename T>
struct RegAt27 {
  const T *p0, *p1, *p2;
  __device__ __forceinline__ T operator()(int a, int b, int c) const {
    const T *p = c < 0 ? p0 : (c == 0 ? p1 : p2);
    return p[(b + 1) * 3 + (a + 1)];
  }
};

// k_youngs with each thread marching up KZ cells in z: the 3 x 3 plane
// neighbourhoods of k-1, k, k+1 stay in registers, so a level loads the nine
// values of the plane k+2 (issued before the level's arithmetic) instead of
// 27. The cell arithmetic is youngs_cell's through a register accessor, the
// same expressions, the same values. MODE 0: the range checks per lane as in
// k_youngs (left out when the range flag allows); 1: only the form without
// them, and nothing when the range flag is raised or not given; 2: nothing
// unless the flag is raised, then mode 0 over every tile of the (gx, gy, gz)
// tile grid, the blocks striding over the tiles (so 1 then 2 compute the
// field in either case, the register-hungry checked form runs only when
// needed, and a few hundred blocks of mode 2 find the flag clear and exit).
template <typename T, int KZ, int MODE>
__device__ __forceinline__ void youngs_mz_tile(const fz_dims &d, const double *__restrict__ dzci,
                                               const double *__restrict__ dzfi, double dxi_d, double dyi_d,
                                               const T *__restrict__ psi, T *__restrict__ normx,
                                               T *__restrict__ normy, T *__restrict__ normz, T *__restrict__ kappa,
                                               int xper, bool nocheck, unsigned bx, unsigned by, unsigned bz) {
  const int i = (int)bx * blockDim.x + threadIdx.x + 1, j = (int)by * blockDim.y + threadIdx.y + 1;
  const int k0 = (int)bz * KZ + 1, kend = min(k0 + KZ - 1, d.n3);
  if (i > d.n1 || j > d.n2) return;  // (no barriers below)
  const BufFld<T> G = buf_fld(psi, d);
  const unsigned e00 = (unsigned)i + (unsigned)d.sy * (unsigned)j, sy = (unsigned)d.sy, sz = (unsigned)d.sz;
  auto load = [&](T *pl, int kp) {
#pragma unroll
    for (int b = -1; b <= 1; b++)
#pragma unroll
      for (int a = -1; a <= 1; a++) pl[(b + 1) * 3 + (a + 1)] = G(e00 + (unsigned)a + (unsigned)b * sy + (unsigned)kp * sz, 0);
  };
  T p0[9], p1[9], p2[9], pn[9];
  load(p0, k0 - 1);
  load(p1, k0);
  load(p2, k0 + 1);
  const XHalo xh = xhalo_of(d, i);
  for (int k = k0; k <= kend; k++) {
    if (k < kend) load(pn, k + 2);  // in flight during this level
    const long q = (long)e00 + (long)sz * k;
    T nx, ny, nz, ka;
    bool ok = true;
    const RegAt27<T> R{p0, p1, p2};
    if (MODE == 1 || nocheck) {
      youngs_cell<T, true, true>(R, dxi_d, dyi_d, dzci[k], dzci[k - 1], dzfi[k], nx, ny, nz, ka, ok);
    } else {
      youngs_cell<T, true>(R, dxi_d, dyi_d, dzci[k], dzci[k - 1], dzfi[k], nx, ny, nz, ka, ok);
      if (!ok) {  // the plain sequences, from a fresh read of the phase field (see k_youngs)
        const T *ps = psi;
        asm volatile("" : "+s"(ps));
        youngs_cell<T>(buf_at(ps, d, q), dxi_d, dyi_d, dzci[k], dzci[k - 1], dzfi[k], nx, ny, nz, ka, ok);
      }
    }
    normx[q] = nx;
    normy[q] = ny;
    normz[q] = nz;
    kappa[q] = ka;
    if (xper) {  // the periodic x halo of the row
      xhalo_put(xh, (unsigned)q, buf_out(normx, d), nx);
      xhalo_put(xh, (unsigned)q, buf_out(normy, d), ny);
      xhalo_put(xh, (unsigned)q, buf_out(normz, d), nz);
      xhalo_put(xh, (unsigned)q, buf_out(kappa, d), ka);
    }
#pragma unroll
    for (int m = 0; m < 9; m++) {
      p0[m] = p1[m];
      p1[m] = p2[m];
      p2[m] = pn[m];
    }
  }
}

template <typename T, int KZ, int MODE = 0>
__global__ void __launch_bounds__(256) k_youngs_mz(fz_dims d, const double *__restrict__ dzci,
                                                   const double *__restrict__ dzfi, double dxi_d, double dyi_d,
                                                   const T *__restrict__ psi, T *__restrict__ normx,
                                                   T *__restrict__ normy, T *__restrict__ normz,
                                                   T *__restrict__ kappa, int xper,
                                                   const unsigned *__restrict__ rflag, uint3 tiles) {
  const bool nocheck = rflag && *rflag == 0u;  // see k_youngs
  if constexpr (MODE == 1) {
    if (!nocheck) return;
  } else if constexpr (MODE == 2) {
    if (!rflag || nocheck) return;
  }
  if constexpr (MODE == 2) {
    const unsigned per_plane = tiles.x * tiles.y, nt = per_plane * tiles.z;
    for (unsigned t = blockIdx.x; t < nt; t += gridDim.x) {
      const unsigned bz = t / per_plane, r = t - bz * per_plane, by = r / tiles.x, bx = r - by * tiles.x;
      youngs_mz_tile<T, KZ, MODE>(d, dzci, dzfi, dxi_d, dyi_d, psi, normx, normy, normz, kappa, xper, false, bx, by,
                                  bz);
    }
  } else {
    unsigned cs_, bx, by, bz;
    fz_stencil_block_b(d.tile_y, gridDim.z, cs_, bx, by, bz);
    youngs_mz_tile<T, KZ, MODE>(d, dzci, dzfi, dxi_d, dyi_d, psi, normx, normy, normz, kappa, xper, nocheck, bx, by,
                                bz);
  }
}

// (a batch of ncase cases when br.bt is set: their grids stacked in z)
template <typename T>
static int launch_youngs(fz_ctx *c, const char *name, const T *psi, T *normx, T *normy, T *normz, T *kappa,
                         fz_bref br = kNoBRef, int ncase = 1, int xper = 0, const unsigned *rflag = nullptr) {
  FZ_TRY(fits_buffers(c, "cmpt_norm_curv_youngs"));
  // (the range flag only where the grid factors lie in [2^-20, 2^20])
  if (!c->youngs_fast_grid) rflag = nullptr;
  // 0: marching in z for FP32 (274 against 345-361 us at 256x256x512); for
  // FP64 the check-free marching kernel and its checked twin where the range
  // flag applies (the marching FP64 kernel with both forms needs 236 VGPRs,
  // two waves per SIMD, and took 758 against 705-745 us for one cell per
  // thread; capped, it spills), one cell per thread elsewhere
  const int yv = c->youngs_variant ? c->youngs_variant : (sizeof(T) == 4 ? 1 : 3);
  constexpr int KZ = 16;
  dim3 mgrid = fz_grid_interior(c->d);
  mgrid.z = (c->d.n3 + KZ - 1) / KZ;
  const uint3 tiles = {mgrid.x, mgrid.y, mgrid.z};
  if (!br.bt && yv == 1) {  // marching in z
    FZ_LAUNCH(c, name, (k_youngs_mz<T, KZ>), mgrid, fz_block3(), 0, c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1],
              psi, normx, normy, normz, kappa, xper, rflag, tiles);
    FZ_LAUNCHED();
    return FZ_OK;
  }
  if (!br.bt && yv == 3 && rflag) {  // the check-free marching kernel, then the checked one if the flag is raised
    FZ_LAUNCH(c, name, (k_youngs_mz<T, KZ, 1>), mgrid, fz_block3(), 0, c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1],
              psi, normx, normy, normz, kappa, xper, rflag, tiles);
    FZ_LAUNCHED();
    // (a grid-stride launch of at most 512 blocks: 2 per CU)
    const unsigned nt = mgrid.x * mgrid.y * mgrid.z, nb2 = nt < 512u ? nt : 512u;
    FZ_LAUNCH(c, name, (k_youngs_mz<T, KZ, 2>), dim3(nb2), fz_block3(), 0, c->d, c->dzci, c->dzfi, c->dli[0],
              c->dli[1], psi, normx, normy, normz, kappa, xper, rflag, tiles);
    FZ_LAUNCHED();
    return FZ_OK;
  }
  dim3 grid = fz_grid_interior(c->d);
  br.gz = grid.z;
  grid.z *= (unsigned)ncase;
  auto kern = br.bt ? k_youngs<T, true> : k_youngs<T, false>;
  FZ_LAUNCH(c, name, kern, grid, fz_block3(), 0, c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1], psi, normx,
            normy, normz, kappa, br, xper, rflag);
  FZ_LAUNCHED();
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// mom_xyz_ad at one cell (linear index q): advection with the phase-field mass
// flux and variable-viscosity (non-split) diffusion, src/mom.f90:545-850
// (PT: element type of the phase fields, FP32 in the mixed-precision step;
// the arithmetic is FP64 either way)
// (AD: accessor of the FP64 velocity, AP: of the phase fields, PtrAt or BufAt)
template <int MASK = 7, class AD, class AP>  // MASK: components computed (1 x, 2 y, 4 z)
__device__ __forceinline__ void mom_ad_cell(const fz_dims &d, int k, const double *__restrict__ dzci,
                                            const double *__restrict__ dzfi, double dxi, double dyi, double rho,
                                            double drho, double mu, double dmu, const AP &psiflx_x,
                                            const AP &psiflx_y, const AP &psiflx_z, const AD &u, const AD &v,
                                            const AD &w, const AP &psi, double &ru, double &rv, double &rw) {
  (void)d;
#define S_(f, a, b, c) ((double)f(a, b, c))
  const double u_ccm = S_(u, 0, 0, -1), u_cmc = S_(u, 0, -1, 0), u_mcc = S_(u, -1, 0, 0), u_ccc = S_(u, 0, 0, 0),
               u_pcc = S_(u, 1, 0, 0), u_mpc = S_(u, -1, 1, 0), u_cpc = S_(u, 0, 1, 0), u_mcp = S_(u, -1, 0, 1),
               u_ccp = S_(u, 0, 0, 1);
  const double v_ccm = S_(v, 0, 0, -1), v_cmc = S_(v, 0, -1, 0), v_pmc = S_(v, 1, -1, 0), v_mcc = S_(v, -1, 0, 0),
               v_ccc = S_(v, 0, 0, 0), v_pcc = S_(v, 1, 0, 0), v_cpc = S_(v, 0, 1, 0), v_cmp = S_(v, 0, -1, 1),
               v_ccp = S_(v, 0, 0, 1);
  const double w_ccm = S_(w, 0, 0, -1), w_pcm = S_(w, 1, 0, -1), w_cpm = S_(w, 0, 1, -1), w_cmc = S_(w, 0, -1, 0),
               w_mcc = S_(w, -1, 0, 0), w_ccc = S_(w, 0, 0, 0), w_pcc = S_(w, 1, 0, 0), w_cpc = S_(w, 0, 1, 0),
               w_ccp = S_(w, 0, 0, 1);
  const double c_ccm = S_(psi, 0, 0, -1), c_pcm = S_(psi, 1, 0, -1), c_cpm = S_(psi, 0, 1, -1),
               c_cmc = S_(psi, 0, -1, 0), c_pmc = S_(psi, 1, -1, 0), c_mcc = S_(psi, -1, 0, 0),
               c_ccc = S_(psi, 0, 0, 0), c_pcc = S_(psi, 1, 0, 0), c_mpc = S_(psi, -1, 1, 0),
               c_cpc = S_(psi, 0, 1, 0), c_cmp = S_(psi, 0, -1, 1), c_mcp = S_(psi, -1, 0, 1),
               c_ccp = S_(psi, 0, 0, 1), c_cpp = S_(psi, 0, 1, 1), c_ppc = S_(psi, 1, 1, 0),
               c_pcp = S_(psi, 1, 0, 1);
  const double flx_x_mcc = rho * u_mcc + drho * S_(psiflx_x, -1, 0, 0);
  const double flx_x_ccc = rho * u_ccc + drho * S_(psiflx_x, 0, 0, 0);
  const double flx_x_pcc = rho * u_pcc + drho * S_(psiflx_x, 1, 0, 0);
  const double flx_x_mpc = rho * u_mpc + drho * S_(psiflx_x, -1, 1, 0);
  const double flx_x_cpc = rho * u_cpc + drho * S_(psiflx_x, 0, 1, 0);
  const double flx_x_mcp = rho * u_mcp + drho * S_(psiflx_x, -1, 0, 1);
  const double flx_x_ccp = rho * u_ccp + drho * S_(psiflx_x, 0, 0, 1);
  const double flx_y_cmc = rho * v_cmc + drho * S_(psiflx_y, 0, -1, 0);
  const double flx_y_ccc = rho * v_ccc + drho * S_(psiflx_y, 0, 0, 0);
  const double flx_y_cpc = rho * v_cpc + drho * S_(psiflx_y, 0, 1, 0);
  const double flx_y_pmc = rho * v_pmc + drho * S_(psiflx_y, 1, -1, 0);
  const double flx_y_pcc = rho * v_pcc + drho * S_(psiflx_y, 1, 0, 0);
  const double flx_y_cmp = rho * v_cmp + drho * S_(psiflx_y, 0, -1, 1);
  const double flx_y_ccp = rho * v_ccp + drho * S_(psiflx_y, 0, 0, 1);
  const double flx_z_ccm = rho * w_ccm + drho * S_(psiflx_z, 0, 0, -1);
  const double flx_z_ccc = rho * w_ccc + drho * S_(psiflx_z, 0, 0, 0);
  const double flx_z_ccp = rho * w_ccp + drho * S_(psiflx_z, 0, 0, 1);
  const double flx_z_pcm = rho * w_pcm + drho * S_(psiflx_z, 1, 0, -1);
  const double flx_z_pcc = rho * w_pcc + drho * S_(psiflx_z, 1, 0, 0);
  const double flx_z_cpm = rho * w_cpm + drho * S_(psiflx_z, 0, 1, -1);
  const double flx_z_cpc = rho * w_cpc + drho * S_(psiflx_z, 0, 1, 0);
#undef S_
  const double dzci_c = dzci[k], dzci_m = dzci[k - 1], dzfi_c = dzfi[k], dzfi_p = dzfi[k + 1];
  double dudt_aux = 0.0, dvdt_aux = 0.0, dwdt_aux = 0.0;
  if constexpr ((MASK & 1) != 0) {
    const double uuip = 0.25 * (flx_x_pcc + flx_x_ccc) * (u_ccc + u_pcc);
    const double uuim = 0.25 * (flx_x_mcc + flx_x_ccc) * (u_ccc + u_mcc);
    const double vujp = 0.25 * (flx_y_pcc + flx_y_ccc) * (u_ccc + u_cpc);
    const double vujm = 0.25 * (flx_y_pmc + flx_y_cmc) * (u_ccc + u_cmc);
    const double wukp = 0.25 * (flx_z_pcc + flx_z_ccc) * (u_ccc + u_ccp);
    const double wukm = 0.25 * (flx_z_pcm + flx_z_ccm) * (u_ccc + u_ccm);
    dudt_aux = dxi * (-uuip + uuim) + dyi * (-vujp + vujm) + dzfi_c * (-wukp + wukm);
  }
  if constexpr ((MASK & 2) != 0) {
    const double uvip = 0.25 * (flx_x_ccc + flx_x_cpc) * (v_ccc + v_pcc);
    const double uvim = 0.25 * (flx_x_mcc + flx_x_mpc) * (v_ccc + v_mcc);
    const double vvjp = 0.25 * (flx_y_ccc + flx_y_cpc) * (v_ccc + v_cpc);
    const double vvjm = 0.25 * (flx_y_ccc + flx_y_cmc) * (v_ccc + v_cmc);
    const double wvkp = 0.25 * (flx_z_ccc + flx_z_cpc) * (v_ccc + v_ccp);
    const double wvkm = 0.25 * (flx_z_ccm + flx_z_cpm) * (v_ccc + v_ccm);
    dvdt_aux = dxi * (-uvip + uvim) + dyi * (-vvjp + vvjm) + dzfi_c * (-wvkp + wvkm);
  }
  if constexpr ((MASK & 4) != 0) {
    const double uwip = 0.25 * (flx_x_ccc + flx_x_ccp) * (w_ccc + w_pcc);
    const double uwim = 0.25 * (flx_x_mcc + flx_x_mcp) * (w_ccc + w_mcc);
    const double vwjp = 0.25 * (flx_y_ccc + flx_y_ccp) * (w_ccc + w_cpc);
    const double vwjm = 0.25 * (flx_y_cmc + flx_y_cmp) * (w_ccc + w_cmc);
    const double wwkp = 0.25 * (flx_z_ccc + flx_z_ccp) * (w_ccc + w_ccp);
    const double wwkm = 0.25 * (flx_z_ccc + flx_z_ccm) * (w_ccc + w_ccm);
    dwdt_aux = dxi * (-uwip + uwim) + dyi * (-vwjp + vwjm) + dzci_c * (-wwkp + wwkm);
  }
  if constexpr ((MASK & 1) != 0) {
    const double dudxp = (u_pcc - u_ccc) * dxi, dudxm = (u_ccc - u_mcc) * dxi;
    const double dvdxp = (v_pcc - v_ccc) * dxi, dvdxm = (v_pmc - v_cmc) * dxi;
    const double dudyp = (u_cpc - u_ccc) * dyi, dudym = (u_ccc - u_cmc) * dyi;
    const double dudzp = (u_ccp - u_ccc) * dzci_c, dudzm = (u_ccc - u_ccm) * dzci_m;
    const double dwdxp = (w_pcc - w_ccc) * dxi, dwdxm = (w_pcm - w_ccm) * dxi;
    const double muxp = mu + dmu * c_pcc, muxm = mu + dmu * c_ccc;
    const double muyp = mu + dmu * 0.25 * (c_ccc + c_cpc + c_ppc + c_pcc);
    const double muym = mu + dmu * 0.25 * (c_ccc + c_cmc + c_pmc + c_pcc);
    const double muzp = mu + dmu * 0.25 * (c_ccc + c_pcc + c_ccp + c_pcp);
    const double muzm = mu + dmu * 0.25 * (c_ccc + c_pcc + c_ccm + c_pcm);
    dudt_aux = dudt_aux + dxi * ((dudxp + dudxp) * muxp - (dudxm + dudxm) * muxm) +
               dyi * ((dudyp + dvdxp) * muyp - (dudym + dvdxm) * muym) +
               dzfi_c * ((dudzp + dwdxp) * muzp - (dudzm + dwdxm) * muzm);
  }
  if constexpr ((MASK & 2) != 0) {
    const double dvdxp = (v_pcc - v_ccc) * dxi, dvdxm = (v_ccc - v_mcc) * dxi;
    const double dudyp = (u_cpc - u_ccc) * dyi, dudym = (u_mpc - u_mcc) * dyi;
    const double dvdyp = (v_cpc - v_ccc) * dyi, dvdym = (v_ccc - v_cmc) * dyi;
    const double dvdzp = (v_ccp - v_ccc) * dzci_c, dvdzm = (v_ccc - v_ccm) * dzci_m;
    const double dwdyp = (w_cpc - w_ccc) * dyi, dwdym = (w_cpm - w_ccm) * dyi;
    const double muxp = mu + dmu * 0.25 * (c_ccc + c_pcc + c_ppc + c_cpc);
    const double muxm = mu + dmu * 0.25 * (c_ccc + c_mcc + c_mpc + c_cpc);
    const double muyp = mu + dmu * c_cpc, muym = mu + dmu * c_ccc;
    const double muzp = mu + dmu * 0.25 * (c_ccc + c_cpc + c_cpp + c_ccp);
    const double muzm = mu + dmu * 0.25 * (c_ccc + c_cpc + c_cpm + c_ccm);
    dvdt_aux = dvdt_aux + dxi * ((dvdxp + dudyp) * muxp - (dvdxm + dudym) * muxm) +
               dyi * ((dvdyp + dvdyp) * muyp - (dvdym + dvdym) * muym) +
               dzfi_c * ((dvdzp + dwdyp) * muzp - (dvdzm + dwdym) * muzm);
  }
  if constexpr ((MASK & 4) != 0) {
    const double dwdxp = (w_pcc - w_ccc) * dxi, dwdxm = (w_ccc - w_mcc) * dxi;
    const double dudzp = (u_ccp - u_ccc) * dzci_c, dudzm = (u_mcp - u_mcc) * dzci_c;
    const double dwdyp = (w_cpc - w_ccc) * dyi, dwdym = (w_ccc - w_cmc) * dyi;
    const double dvdzp = (v_ccp - v_ccc) * dzci_c, dvdzm = (v_cmp - v_cmc) * dzci_c;
    const double dwdzp = (w_ccp - w_ccc) * dzfi_p, dwdzm = (w_ccc - w_ccm) * dzfi_c;
    const double muxp = mu + dmu * 0.25 * (c_ccc + c_pcc + c_ccp + c_pcp);
    const double muxm = mu + dmu * 0.25 * (c_ccc + c_mcc + c_ccp + c_mcp);
    const double muyp = mu + dmu * 0.25 * (c_ccc + c_cpc + c_ccp + c_cpp);
    const double muym = mu + dmu * 0.25 * (c_ccc + c_cmc + c_ccp + c_cmp);
    const double muzp = mu + dmu * c_ccp, muzm = mu + dmu * c_ccc;
    dwdt_aux = dwdt_aux + dxi * ((dwdxp + dudzp) * muxp - (dwdxm + dudzm) * muxm) +
               dyi * ((dwdyp + dvdzp) * muyp - (dwdym + dvdzm) * muym) +
               dzci_c * ((dwdzp + dwdzp) * muzp - (dwdzm + dwdzm) * muzm);
  }
  ru = dudt_aux;
  rv = dvdt_aux;
  rw = dwdt_aux;
}

__global__ void __launch_bounds__(256) k_mom_ad(fz_dims d, const double *__restrict__ dzci,
                                                const double *__restrict__ dzfi, double dxi, double dyi, double rho,
                                                double drho, double mu, double dmu,
                                                const double *__restrict__ psiflx_x,
                                                const double *__restrict__ psiflx_y,
                                                const double *__restrict__ psiflx_z, const double *__restrict__ u,
                                                const double *__restrict__ v, const double *__restrict__ w,
                                                const double *__restrict__ psi, double *__restrict__ dudt,
                                                double *__restrict__ dvdt, double *__restrict__ dwdt) {
  CELL_IJK(d);
  double ru, rv, rw;
  mom_ad_cell(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, ptr_at(psiflx_x, d, q), ptr_at(psiflx_y, d, q),
              ptr_at(psiflx_z, d, q), ptr_at(u, d, q), ptr_at(v, d, q), ptr_at(w, d, q), ptr_at(psi, d, q), ru, rv,
              rw);
  dudt[r] = ru;
  dvdt[r] = rv;
  dwdt[r] = rw;
}

// The velocity update of the momentum step at one cell, src/rk.f90:96-98 and
// :125-127 fused: (a1 / d + a2 / d) + a3 / d per component, a1 = u rho_n,
// a2 = factor1 dudtrk + factor2 dudtrko, a3 = factor12 dudtrk (the second
// right-hand side), d = rho_p. The nine quotients share three refined
// reciprocals (div_inrange's sequence, then the device's division fix-up,
// which sets zero, infinite and NaN quotients as '/' does); where a finite
// non-zero operand of the cell lies outside [2^-300, 2^300] in magnitude,
// outside that sequence's range, the cell divides with '/'. Bit for bit '/'.
__device__ __forceinline__ void mom_update(double au1, double au2, double au3, double du, double av1, double av2,
                                           double av3, double dv, double aw1, double aw2, double aw3, double dw,
                                           double &uo, double &vo, double &wo) {
  int lo = 0, hi = 0;  // frexp exponents: 0 for zeros, infinities and NaN
  auto span = [&](double x) {
    const int e = __builtin_amdgcn_frexp_exp(x);
    lo = min(lo, e);
    hi = max(hi, e);
  };
  span(au1), span(au2), span(au3), span(du), span(av1), span(av2), span(av3), span(dv);
  span(aw1), span(aw2), span(aw3), span(dw);
  if (lo >= -299 && hi <= 301) {
    auto q3 = [](double a1, double a2, double a3, double d) {
      double y = __builtin_amdgcn_rcp(d);
      double e = __builtin_fma(-d, y, 1.0);
      y = __builtin_fma(y, e, y);
      e = __builtin_fma(-d, y, 1.0);
      y = __builtin_fma(y, e, y);
      auto quo = [&](double a) {
        const double q = a * y;
        return __builtin_amdgcn_div_fixup(__builtin_fma(__builtin_fma(-d, q, a), y, q), d, a);
      };
      return (quo(a1) + quo(a2)) + quo(a3);
    };
    uo = q3(au1, au2, au3, du);
    vo = q3(av1, av2, av3, dv);
    wo = q3(aw1, aw2, aw3, dw);
  } else {
    uo = (au1 / du + au2 / du) + au3 / du;
    vo = (av1 / dv + av2 / dv) + av3 / dv;
    wo = (aw1 / dw + aw2 / dw) + aw3 / dw;
  }
}

// mom_xyz_oth at one cell, src/mom.f90:912-1038 (constant-coefficient split, Brackbill)
template <class AD, class AP, class AK = AP>
__device__ __forceinline__ void mom_oth_cell(int k, const double *__restrict__ dzci, double dxi, double dyi,
                                             double dt_r, double rho, double drho, double surf_factor, double rho0,
                                             double bfx, double bfy, double bfz, double gx, double gy, double gz,
                                             double rax, double ray, double raz, const AD &p, const AP &psi,
                                             const AK &kappa, const AD &pn, const AD &po, double &ru, double &rv,
                                             double &rw) {
  const double dzci_c = dzci[k];
  const double p_ccc = p(0, 0, 0), p_pcc = p(1, 0, 0), p_cpc = p(0, 1, 0), p_ccp = p(0, 0, 1);
  const double q_ccc = (1.0 + dt_r) * pn(0, 0, 0) - dt_r * po(0, 0, 0);
  const double q_pcc = (1.0 + dt_r) * pn(1, 0, 0) - dt_r * po(1, 0, 0);
  const double q_cpc = (1.0 + dt_r) * pn(0, 1, 0) - dt_r * po(0, 1, 0);
  const double q_ccp = (1.0 + dt_r) * pn(0, 0, 1) - dt_r * po(0, 0, 1);
  const double k_ccc = kappa(0, 0, 0), k_pcc = kappa(1, 0, 0), k_cpc = kappa(0, 1, 0), k_ccp = kappa(0, 0, 1);
  const double c_ccc = psi(0, 0, 0), c_pcc = psi(1, 0, 0), c_cpc = psi(0, 1, 0), c_ccp = psi(0, 0, 1);
  const double psixp = 0.5 * (c_pcc + c_ccc), psiyp = 0.5 * (c_cpc + c_ccc), psizp = 0.5 * (c_ccp + c_ccc);
  const double rhoxp = rho + drho * psixp, rhoyp = rho + drho * psiyp, rhozp = rho + drho * psizp;
  const double dpdx = (p_pcc - p_ccc) * dxi, dpdy = (p_cpc - p_ccc) * dyi, dpdz = (p_ccp - p_ccc) * dzci_c;
  double dudt_aux = bfx + (rhoxp - rax) * gx;
  double dvdt_aux = bfy + (rhoyp - ray) * gy;
  double dwdt_aux = bfz + (rhozp - raz) * gz;
  double skx = 0.5 * (k_pcc + k_ccc) * surf_factor, sky = 0.5 * (k_cpc + k_ccc) * surf_factor,
         skz = 0.5 * (k_ccp + k_ccc) * surf_factor;
  skx = skx * rhoxp;
  sky = sky * rhoyp;
  skz = skz * rhozp;
  const double surfx = skx * (c_pcc - c_ccc) * dxi, surfy = sky * (c_cpc - c_ccc) * dyi,
               surfz = skz * (c_ccp - c_ccc) * dzci_c;
  dudt_aux = dudt_aux + surfx;
  dvdt_aux = dvdt_aux + surfy;
  dwdt_aux = dwdt_aux + surfz;
  const double dpdx_e = (q_pcc - q_ccc) * dxi, dpdy_e = (q_cpc - q_ccc) * dyi, dpdz_e = (q_ccp - q_ccc) * dzci_c;
  dudt_aux = dudt_aux - (dpdx_e + rhoxp / rho0 * (dpdx - dpdx_e));
  dvdt_aux = dvdt_aux - (dpdy_e + rhoyp / rho0 * (dpdy - dpdy_e));
  dwdt_aux = dwdt_aux - (dpdz_e + rhozp / rho0 * (dpdz - dpdz_e));
  ru = dudt_aux;
  rv = dvdt_aux;
  rw = dwdt_aux;
}

__global__ void __launch_bounds__(256) k_mom_oth(fz_dims d, const double *__restrict__ dzci, double dxi, double dyi,
                                                 double dt_r, double rho, double drho, double surf_factor,
                                                 double rho0, double bfx, double bfy, double bfz, double gx, double gy,
                                                 double gz, double rax, double ray, double raz,
                                                 const double *__restrict__ p, const double *__restrict__ psi,
                                                 const double *__restrict__ kappa, const double *__restrict__ pn,
                                                 const double *__restrict__ po, double *__restrict__ dudt,
                                                 double *__restrict__ dvdt, double *__restrict__ dwdt) {
  CELL_IJK(d);
  double ru, rv, rw;
  mom_oth_cell(k, dzci, dxi, dyi, dt_r, rho, drho, surf_factor, rho0, bfx, bfy, bfz, gx, gy, gz, rax, ray, raz,
               ptr_at(p, d, q), ptr_at(psi, d, q), ptr_at(kappa, d, q), ptr_at(pn, d, q), ptr_at(po, d, q), ru, rv,
               rw);
  dudt[r] = ru;
  dvdt[r] = rv;
  dwdt[r] = rw;
}

// The whole momentum part of rk (src/rk.f90:72-130) in one pass: mom_xyz_ad,
// the AB2 update with the density rescale, mom_xyz_oth (pn = p) and the
// second update, with the same expressions as k_mom_ad, k_rk_mom_ab2,
// k_mom_oth and k_rk_mom_oth in turn (no intermediate is rounded
// differently: each was a double in memory and is a double in a register
// here). The new velocity goes to un/vn/wn (the stencil reads the old one);
// the AB2 history duo/dvo/dwo is read and replaced by this step's
// advection-diffusion term in place (on the first step the history is the
// term itself, src/rk.f90:76-83).
// the body, shared by occupancy variants of the kernel
// (PT: phase fields, HT: the AB2 history; both FP32 in the mixed-precision
// step. STAGED: the three advection-diffusion components one after another,
// each with its own loads, behind scheduling barriers, so fewer values are
// live at once; the arithmetic is the same.)
template <typename PT, typename HT, bool STAGED, bool BT>
__device__ __forceinline__ void mom_fused_body(
    fz_dims d, const double *__restrict__ dzci, const double *__restrict__ dzfi, double dxi, double dyi, double rho,
    double drho, double mu, double dmu, double factor1, double factor2, int first, double factor12, double dt_r,
    double surf_factor, double rho0, double bfx, double bfy, double bfz, double gx, double gy, double gz, double rax,
    double ray, double raz, const PT *__restrict__ psiflx_x, const PT *__restrict__ psiflx_y,
    const PT *__restrict__ psiflx_z, const double *__restrict__ u, const double *__restrict__ v,
    const double *__restrict__ w, const PT *__restrict__ psi, const PT *__restrict__ psio,
    const double *__restrict__ p, const double *__restrict__ po, const PT *__restrict__ kappa,
    HT *__restrict__ duo, HT *__restrict__ dvo, HT *__restrict__ dwo, double *__restrict__ un,
    double *__restrict__ vn, double *__restrict__ wn, const fz_bref &br) {
  CELL_IJKB(d, br, cs);
  if constexpr (BT) {  // this block's case of a batch
    const fz_bcase &B = br.bt[cs];
    const double *stc = B.stc;
    factor1 = stc[FZ_STC_F1];
    factor2 = stc[FZ_STC_F2];
    factor12 = stc[FZ_STC_F12];
    dt_r = stc[FZ_STC_DTR];
    first = 0;
    rho = B.rho, drho = B.drho, mu = B.mu, dmu = B.dmu, surf_factor = B.surf_factor, rho0 = B.rho0;
    bfx = B.bforce[0], bfy = B.bforce[1], bfz = B.bforce[2], gx = B.gacc[0], gy = B.gacc[1], gz = B.gacc[2];
    rax = ray = raz = 0.0;
    psiflx_x = (const PT *)fz_brole(br, cs, BR_FX);
    psiflx_y = (const PT *)fz_brole(br, cs, BR_FY);
    psiflx_z = (const PT *)fz_brole(br, cs, BR_FZ);
    u = (const double *)fz_brole(br, cs, BR_U);
    v = (const double *)fz_brole(br, cs, BR_V);
    w = (const double *)fz_brole(br, cs, BR_W);
    psi = (const PT *)fz_brole(br, cs, BR_PSI_NEW);
    psio = (const PT *)fz_brole(br, cs, BR_PSI_OLD);
    p = (const double *)fz_brole(br, cs, BR_P);
    po = (const double *)fz_brole(br, cs, BR_PO);
    kappa = (const PT *)fz_brole(br, cs, BR_KAPPA);
    duo = (HT *)fz_brole(br, cs, BR_DUO);
    dvo = (HT *)fz_brole(br, cs, BR_DVO);
    dwo = (HT *)fz_brole(br, cs, BR_DWO);
    un = (double *)fz_brole(br, cs, BR_UN);
    vn = (double *)fz_brole(br, cs, BR_VN);
    wn = (double *)fz_brole(br, cs, BR_WN);
  }
  const BufAt<double> U = buf_at(u, d, q), V = buf_at(v, d, q), W = buf_at(w, d, q);
  const BufAt<PT> PSI = buf_at(psi, d, q), FX = buf_at(psiflx_x, d, q), FY = buf_at(psiflx_y, d, q),
                  FZ = buf_at(psiflx_z, d, q);
  double ru, rv, rw;
  if constexpr (STAGED) {
    double t0, t1;
    mom_ad_cell<1>(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, ru, t0, t1);
    asm volatile("" ::: "memory");  // each stage loads its own values
    __builtin_amdgcn_sched_barrier(0);
    mom_ad_cell<2>(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, t0, rv, t1);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    mom_ad_cell<4>(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, t0, t1, rw);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
  } else {
    mom_ad_cell(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, ru, rv, rw);
  }
  const double ruo = first ? ru : duo[r], rvo = first ? rv : dvo[r], rwo = first ? rw : dwo[r];
  duo[r] = ru;
  dvo[r] = rv;
  dwo[r] = rw;
  const BufAt<PT> PSIO = buf_at(psio, d, q), KAP = buf_at(kappa, d, q);
  const BufAt<double> P = buf_at(p, d, q), PO = buf_at(po, d, q);
  const double rhox_n = rho + drho * 0.5 * ((double)PSIO(0, 0, 0) + (double)PSIO(1, 0, 0));
  const double rhoy_n = rho + drho * 0.5 * ((double)PSIO(0, 0, 0) + (double)PSIO(0, 1, 0));
  const double rhoz_n = rho + drho * 0.5 * ((double)PSIO(0, 0, 0) + (double)PSIO(0, 0, 1));
  const double rhox_p = rho + drho * 0.5 * ((double)PSI(0, 0, 0) + (double)PSI(1, 0, 0));
  const double rhoy_p = rho + drho * 0.5 * ((double)PSI(0, 0, 0) + (double)PSI(0, 1, 0));
  const double rhoz_p = rho + drho * 0.5 * ((double)PSI(0, 0, 0) + (double)PSI(0, 0, 1));
  const double nu1 = U(0, 0, 0) * rhox_n, nu2 = factor1 * ru + factor2 * ruo;
  const double nv1 = V(0, 0, 0) * rhoy_n, nv2 = factor1 * rv + factor2 * rvo;
  const double nw1 = W(0, 0, 0) * rhoz_n, nw2 = factor1 * rw + factor2 * rwo;
  mom_oth_cell(k, dzci, dxi, dyi, dt_r, rho, drho, surf_factor, rho0, bfx, bfy, bfz, gx, gy, gz, rax, ray, raz, P, PSI,
               KAP, P, PO, ru, rv, rw);
  double uo, vo, wo;
  mom_update(nu1, nu2, factor12 * ru, rhox_p, nv1, nv2, factor12 * rv, rhoy_p, nw1, nw2, factor12 * rw, rhoz_p, uo, vo,
             wo);
  un[q] = uo;
  vn[q] = vo;
  wn[q] = wo;
}

#define FZ_MOM_FUSED_KERNEL(NAME, ATTR, STG)                                                                     \
  template <typename PT, typename HT, bool BT = false>                                                           \
  __global__ void __launch_bounds__(256) ATTR NAME(                                                              \
      fz_dims d, const double *__restrict__ dzci, const double *__restrict__ dzfi, double dxi, double dyi,       \
      double rho, double drho, double mu, double dmu, double factor1_v, double factor2_v, int first,            \
      double factor12_v, double dt_r_v, const double *__restrict__ stc, double surf_factor, double rho0,         \
      double bfx, double bfy, double bfz, double gx, double gy,                                                   \
      double gz, double rax, double ray, double raz, const PT *__restrict__ psiflx_x,                            \
      const PT *__restrict__ psiflx_y, const PT *__restrict__ psiflx_z, const double *__restrict__ u,            \
      const double *__restrict__ v, const double *__restrict__ w, const PT *__restrict__ psi,                    \
      const PT *__restrict__ psio, const double *__restrict__ p, const double *__restrict__ po,                  \
      const PT *__restrict__ kappa, HT *__restrict__ duo, HT *__restrict__ dvo, HT *__restrict__ dwo,            \
      double *__restrict__ un, double *__restrict__ vn, double *__restrict__ wn, fz_bref br) {                   \
    const double factor1 = stc ? stc[FZ_STC_F1] : factor1_v, factor2 = stc ? stc[FZ_STC_F2] : factor2_v;      \
    const double factor12 = stc ? stc[FZ_STC_F12] : factor12_v, dt_r = stc ? stc[FZ_STC_DTR] : dt_r_v;         \
    mom_fused_body<PT, HT, STG, BT>(d, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, factor1, factor2, first, factor12,    \
                                dt_r,                                                                            \
                           surf_factor, rho0, bfx, bfy, bfz, gx, gy, gz, rax, ray, raz, psiflx_x, psiflx_y,      \
                           psiflx_z, u, v, w, psi, psio, p, po, kappa, duo, dvo, dwo, un, vn, wn, br);           \
  }
FZ_MOM_FUSED_KERNEL(k_mom_fused, , false)
FZ_MOM_FUSED_KERNEL(k_mom_fused_st, __attribute__((amdgpu_waves_per_eu(4))), true)
FZ_MOM_FUSED_KERNEL(k_mom_fused_w4, __attribute__((amdgpu_waves_per_eu(4))), false)
FZ_MOM_FUSED_KERNEL(k_mom_fused_w5, __attribute__((amdgpu_waves_per_eu(5))), false)

// ---------------------------------------------------------------------------
// k_mom_fused with its stencil fields staged in LDS. The advection-diffusion
// stencil reads u, v, w, psi and the three psi fluxes at up to 16 neighbours
// each; read from HBM/L1 one by one, those loads keep the texture-address
// unit busy 94 % of the kernel (rocprofv3 TA_TA_BUSY at 256x256x512) while
// HBM moves 3 TB/s. Here a block of MX x MY cells marches up KZ planes in z
// and holds, for each of the seven fields, the planes k-1, k, k+1 of its
// tile plus a one-cell x/y frame in LDS (a ring of three plane slots); each
// plane is read from HBM once, in rows, the plane k+2 being loaded into
// registers while level k is computed. The cell arithmetic is
// mom_fused_body's through an LDS accessor (LdsAt), so the results are the
// same bit for bit.
constexpr int MOM_MX = 32, MOM_MY = 8, MOM_PX = MOM_MX + 2, MOM_PL = MOM_PX * (MOM_MY + 2);


// a field's three plane slots in LDS, seen from one cell
template <typename T>
struct LdsAt {
  const T *base;
  int c;           // the cell's position in a plane
  int s0, s1, s2;  // offsets of the planes k-1, k, k+1
  __device__ __forceinline__ T operator()(int a, int b, int cc) const {
    const int so = cc < 0 ? s0 : (cc == 0 ? s1 : s2);
    return base[so + c + a + b * MOM_PX];
  }
};

// WPE: waves per SIMD the registers must allow; STEADY: a step after the
// first (first = 0) with no bulk-density term (rho_av = 0), known at compile time
// (XPER: the cells 1 and n1 also store their new velocity at the periodic
// x-halo images n1+1 and 0 of their row)
template <typename PT, typename HT, int KZ, int WPE, bool BT = false, bool STEADY = false, bool XPER = false>
__global__ void __launch_bounds__(MOM_MX *MOM_MY) __attribute__((amdgpu_waves_per_eu(WPE))) k_mom_lds(
    fz_dims d, const double *__restrict__ dzci, const double *__restrict__ dzfi, double dxi, double dyi, double rho,
    double drho, double mu, double dmu, double factor1_v, double factor2_v, int first, double factor12_v,
    double dt_r_v, const double *__restrict__ stc, double surf_factor, double rho0, double bfx, double bfy,
    double bfz, double gx, double gy, double gz, double rax, double ray, double raz, const PT *__restrict__ psiflx_x,
    const PT *__restrict__ psiflx_y, const PT *__restrict__ psiflx_z, const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, const PT *__restrict__ psi,
    const PT *__restrict__ psio, const double *__restrict__ p, const double *__restrict__ po,
    const PT *__restrict__ kappa, HT *__restrict__ duo, HT *__restrict__ dvo, HT *__restrict__ dwo,
    double *__restrict__ un, double *__restrict__ vn, double *__restrict__ wn, fz_bref br) {
  constexpr int NT = MOM_MX * MOM_MY, PL = MOM_PL;
  constexpr int NPT = (PL + NT - 1) / NT;  // plane positions per thread
  // u, v, w, psi and the z flux are read at k-1, k, k+1 (three plane slots);
  // the x and y fluxes at k and k+1 only (two slots)
  __shared__ double s_u[3 * PL], s_v[3 * PL], s_w[3 * PL];
  __shared__ PT s_psi[3 * PL], s_fx[2 * PL], s_fy[2 * PL], s_fz[3 * PL];
  FZ_BLOCK_XYZB(d, br, cs, bx, by, bz);
  if constexpr (STEADY) {
    first = 0;
    rax = ray = raz = 0.0;
  }
  if constexpr (BT) {  // this block's case of a batch
    const fz_bcase &B = br.bt[cs];
    stc = B.stc;
    first = 0;
    rho = B.rho, drho = B.drho, mu = B.mu, dmu = B.dmu, surf_factor = B.surf_factor, rho0 = B.rho0;
    bfx = B.bforce[0], bfy = B.bforce[1], bfz = B.bforce[2], gx = B.gacc[0], gy = B.gacc[1], gz = B.gacc[2];
    rax = ray = raz = 0.0;
    psiflx_x = (const PT *)fz_brole(br, cs, BR_FX);
    psiflx_y = (const PT *)fz_brole(br, cs, BR_FY);
    psiflx_z = (const PT *)fz_brole(br, cs, BR_FZ);
    u = (const double *)fz_brole(br, cs, BR_U);
    v = (const double *)fz_brole(br, cs, BR_V);
    w = (const double *)fz_brole(br, cs, BR_W);
    psi = (const PT *)fz_brole(br, cs, BR_PSI_NEW);
    psio = (const PT *)fz_brole(br, cs, BR_PSI_OLD);
    p = (const double *)fz_brole(br, cs, BR_P);
    po = (const double *)fz_brole(br, cs, BR_PO);
    kappa = (const PT *)fz_brole(br, cs, BR_KAPPA);
    duo = (HT *)fz_brole(br, cs, BR_DUO);
    dvo = (HT *)fz_brole(br, cs, BR_DVO);
    dwo = (HT *)fz_brole(br, cs, BR_DWO);
    un = (double *)fz_brole(br, cs, BR_UN);
    vn = (double *)fz_brole(br, cs, BR_VN);
    wn = (double *)fz_brole(br, cs, BR_WN);
  }
  const double factor1 = stc ? stc[FZ_STC_F1] : factor1_v, factor2 = stc ? stc[FZ_STC_F2] : factor2_v;
  const double factor12 = stc ? stc[FZ_STC_F12] : factor12_v, dt_r = stc ? stc[FZ_STC_DTR] : dt_r_v;
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + MOM_MX * ty;
  const int i0 = (int)bx * MOM_MX + 1, j0 = (int)by * MOM_MY + 1, k0 = (int)bz * KZ + 1;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = i <= d.n1 && j <= d.n2;
  const int kend = min(k0 + KZ - 1, d.n3);
  const int c = (ty + 1) * MOM_PX + (tx + 1);
  const BufFld<double> GU = buf_fld(u, d), GV = buf_fld(v, d), GW = buf_fld(w, d);
  const BufFld<PT> GPSI = buf_fld(psi, d), GFX = buf_fld(psiflx_x, d), GFY = buf_fld(psiflx_y, d),
                   GFZ = buf_fld(psiflx_z, d);
  // halo'd index of plane position pos (a frame cell past the grid's halo
  // reads another row of the array, or zero past its end; no cell uses it)
  unsigned gpos[NPT];
#pragma unroll
  for (int r = 0; r < NPT; r++) {
    const int pos = min(t + r * NT, PL - 1), py = pos / MOM_PX, px = pos - py * MOM_PX;
    gpos[r] = (unsigned)(i0 - 1 + px) + (unsigned)d.sy * (unsigned)(j0 - 1 + py);
  }
  auto slot = [](int kp) { return (kp % 3) * PL; };
  auto slot2 = [](int kp) { return (kp % 2) * PL; };
  double qu[NPT], qv[NPT], qw[NPT];
  PT qp[NPT], qx[NPT], qy[NPT], qz[NPT];
  auto load_plane = [&](int kp) {
    const unsigned zo = (unsigned)d.sz * (unsigned)kp;
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      const unsigned e = gpos[r] + zo;
      qu[r] = GU(e, 0);
      qv[r] = GV(e, 0);
      qw[r] = GW(e, 0);
      qp[r] = GPSI(e, 0);
      qx[r] = GFX(e, 0);
      qy[r] = GFY(e, 0);
      qz[r] = GFZ(e, 0);
    }
  };
  auto store_plane = [&](int kp) {
    const int so = slot(kp), so2 = slot2(kp);
#pragma unroll
    for (int r = 0; r < NPT; r++) {
      const int pos = t + r * NT;
      if (pos < PL) {
        s_u[so + pos] = qu[r];
        s_v[so + pos] = qv[r];
        s_w[so + pos] = qw[r];
        s_psi[so + pos] = qp[r];
        s_fx[so2 + pos] = qx[r];
        s_fy[so2 + pos] = qy[r];
        s_fz[so + pos] = qz[r];
      }
    }
  };
  for (int kp = k0 - 1; kp <= k0 + 1; kp++) {
    load_plane(kp);
    store_plane(kp);
  }
  __syncthreads();
  for (int k = k0; k <= kend; k++) {
    if (k < kend) load_plane(k + 2);  // in flight during this level
    if (act) {
      const long q = i + d.sy * j + d.sz * (long)k;
      const long r = (i - 1) + (long)d.n1 * ((j - 1) + (long)d.n2 * (k - 1));
      const int sa = slot(k - 1), sb = slot(k), sc = slot(k + 1), tb = slot2(k), tc = slot2(k + 1);
      const LdsAt<double> U{s_u, c, sa, sb, sc}, V{s_v, c, sa, sb, sc}, W{s_w, c, sa, sb, sc};
      const LdsAt<PT> PSI{s_psi, c, sa, sb, sc}, FX{s_fx, c, tb, tb, tc}, FY{s_fy, c, tb, tb, tc},
          FZ{s_fz, c, sa, sb, sc};
      double ru, rv, rw;
      if constexpr (STEADY) {  // one component at a time: fewer stencil values live at once
        double t0, t1;
        mom_ad_cell<1>(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, ru, t0, t1);
        __builtin_amdgcn_sched_barrier(0);
        mom_ad_cell<2>(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, t0, rv, t1);
        __builtin_amdgcn_sched_barrier(0);
        mom_ad_cell<4>(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, t0, t1, rw);
        __builtin_amdgcn_sched_barrier(0);
      } else {
        mom_ad_cell(d, k, dzci, dzfi, dxi, dyi, rho, drho, mu, dmu, FX, FY, FZ, U, V, W, PSI, ru, rv, rw);
      }
      const double ruo = first ? ru : duo[r], rvo = first ? rv : dvo[r], rwo = first ? rw : dwo[r];
      duo[r] = ru;
      dvo[r] = rv;
      dwo[r] = rw;
      const BufAt<PT> PSIO = buf_at(psio, d, q), KAP = buf_at(kappa, d, q);
      const BufAt<double> P = buf_at(p, d, q), PO = buf_at(po, d, q);
      const double rhox_n = rho + drho * 0.5 * ((double)PSIO(0, 0, 0) + (double)PSIO(1, 0, 0));
      const double rhoy_n = rho + drho * 0.5 * ((double)PSIO(0, 0, 0) + (double)PSIO(0, 1, 0));
      const double rhoz_n = rho + drho * 0.5 * ((double)PSIO(0, 0, 0) + (double)PSIO(0, 0, 1));
      const double rhox_p = rho + drho * 0.5 * ((double)PSI(0, 0, 0) + (double)PSI(1, 0, 0));
      const double rhoy_p = rho + drho * 0.5 * ((double)PSI(0, 0, 0) + (double)PSI(0, 1, 0));
      const double rhoz_p = rho + drho * 0.5 * ((double)PSI(0, 0, 0) + (double)PSI(0, 0, 1));
      const double nu1 = U(0, 0, 0) * rhox_n, nu2 = factor1 * ru + factor2 * ruo;
      const double nv1 = V(0, 0, 0) * rhoy_n, nv2 = factor1 * rv + factor2 * rvo;
      const double nw1 = W(0, 0, 0) * rhoz_n, nw2 = factor1 * rw + factor2 * rwo;
      mom_oth_cell(k, dzci, dxi, dyi, dt_r, rho, drho, surf_factor, rho0, bfx, bfy, bfz, gx, gy, gz, rax, ray, raz,
                   P, PSI, KAP, P, PO, ru, rv, rw);
      double uo, vo, wo;
      mom_update(nu1, nu2, factor12 * ru, rhox_p, nv1, nv2, factor12 * rv, rhoy_p, nw1, nw2, factor12 * rw, rhoz_p, uo,
                 vo, wo);
      un[q] = uo;
      vn[q] = vo;
      wn[q] = wo;
      if constexpr (XPER) {  // (plain stores: buffer resources for three more fields spill scalar registers)
        if (i == d.n1) {
          un[q - d.n1] = uo;
          vn[q - d.n1] = vo;
          wn[q - d.n1] = wo;
        }
        if (i == 1) {
          un[q + d.n1] = uo;
          vn[q + d.n1] = vo;
          wn[q + d.n1] = wo;
        }
      }
    }
    __syncthreads();  // the slot of plane k-1 is free
    if (k < kend) store_plane(k + 2);
    __syncthreads();
  }
}

template <typename PT, typename HT>
int fz_launch_mom_fused(fz_ctx *c, const double rho12[2], const double mu12[2], double factor1, double factor2,
                        int first, double factor12, double dt_r, const double *stc, double sigma, double rho0,
                        const double bforce[3],
                        const double gacc[3], const double rho_av[3], const PT *fx, const PT *fy, const PT *fz,
                        const double *u, const double *v, const double *w, const PT *psi, const PT *psio,
                        const double *p, const double *po, const PT *kappa, HT *duo, HT *dvo, HT *dwo, double *un,
                        double *vn, double *wn, int xper) {
  FZ_TRY(fits_buffers(c, "mom_fused"));
  const double surf_factor = sigma * 2.0 / (rho12[0] + rho12[1]);
  // the context's variant; 0: LDS staging (at 256x256x512: FP64 1386
  // against 1741 us, FP32 phase fields 1134 against 1317 us)
  int variant = c->mom_variant;
  if (variant == 0) variant = 8;
  if (variant == 8) {  // stencil fields staged in LDS
    constexpr int KZ = 16;
    const dim3 grid((c->d.n1 + MOM_MX - 1) / MOM_MX, (c->d.n2 + MOM_MY - 1) / MOM_MY, (c->d.n3 + KZ - 1) / KZ);
    // (3 waves per SIMD fit FP32 phase fields' registers, 169 -> 168 VGPRs;
    // FP64 needs 188, and a 3-wave FP64 build spills 36 and runs 2.2x slower)
    // With FP32 phase fields, a step after the first without a bulk-density
    // term (every step of the workloads but the first) knows both at compile
    // time and computes the stencil's components one at a time: 136 VGPRs
    // instead of 167 plus spills, 1134 against 1317 us for the plain kernel.
    // In FP64 the same gives 133 VGPRs and three waves per SIMD but is
    // slower (1532 against 1403-1452 us at two waves, 188 VGPRs).
    const bool steady = first == 0 && rho_av[0] == 0.0 && rho_av[1] == 0.0 && rho_av[2] == 0.0;
    // (xper: the x-halo stores, compiled only into the instances that make them)
    constexpr int W = sizeof(PT) == 4 ? 3 : 1;
    auto kern = sizeof(PT) == 4 && steady ? (xper ? k_mom_lds<PT, HT, KZ, 3, false, true, true>
                                                  : k_mom_lds<PT, HT, KZ, 3, false, true, false>)
                                          : (xper ? k_mom_lds<PT, HT, KZ, W, false, false, true>
                                                  : k_mom_lds<PT, HT, KZ, W, false, false, false>);
    FZ_LAUNCH(c, sizeof(PT) == 4 ? "mom_fused_f32" : "mom_fused", kern,
              grid, dim3(MOM_MX, MOM_MY), 0, c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1], rho12[1],
              rho12[0] - rho12[1], mu12[1], mu12[0] - mu12[1], factor1, factor2, first, factor12, dt_r, stc,
              surf_factor, rho0, bforce[0], bforce[1], bforce[2], gacc[0], gacc[1], gacc[2], rho_av[0], rho_av[1],
              rho_av[2], fx, fy, fz, u, v, w, psi, psio, p, po, kappa, duo, dvo, dwo, un, vn, wn, kNoBRef);
    FZ_LAUNCHED();
    return FZ_OK;
  }
  auto kern = variant == 4   ? k_mom_fused_w4<PT, HT>
              : variant == 5 ? k_mom_fused_w5<PT, HT>
              : variant == 1 ? k_mom_fused_st<PT, HT>
                             : k_mom_fused<PT, HT>;
  FZ_LAUNCH(c, sizeof(PT) == 4 ? "mom_fused_f32" : "mom_fused", kern, fz_grid_interior(c->d), fz_block3(), 0, c->d,
            c->dzci, c->dzfi, c->dli[0], c->dli[1], rho12[1], rho12[0] - rho12[1], mu12[1], mu12[0] - mu12[1], factor1,
            factor2, first, factor12, dt_r, stc, surf_factor, rho0, bforce[0], bforce[1], bforce[2], gacc[0], gacc[1],
            gacc[2], rho_av[0], rho_av[1], rho_av[2], fx, fy, fz, u, v, w, psi, psio, p, po, kappa, duo, dvo, dwo, un,
            vn, wn, kNoBRef);
  FZ_LAUNCHED();
  if (xper) return xhalo_copy<double>(c, un, vn, wn, nullptr);
  return FZ_OK;
}
template int fz_launch_mom_fused<double, double>(fz_ctx *, const double *, const double *, double, double, int, double,
                                                 double, const double *, double, double, const double *,
                                                 const double *,
                                                 const double *, const double *, const double *, const double *,
                                                 const double *, const double *, const double *, const double *,
                                                 const double *, const double *, const double *, const double *,
                                                 double *, double *, double *, double *, double *, double *, int);
template int fz_launch_mom_fused<float, float>(fz_ctx *, const double *, const double *, double, double, int, double,
                                               double, const double *, double, double, const double *,
                                               const double *, const double *,
                                               const float *, const float *, const float *, const double *,
                                               const double *, const double *, const float *, const float *,
                                               const double *, const double *, const float *, float *, float *,
                                               float *, double *, double *, double *, int);

// AB2 momentum update with the density rescale, src/rk.f90:86-101
__global__ void k_rk_mom_ab2(fz_dims d, double factor1, double factor2, double rho, double drho,
                             const double *__restrict__ psio, const double *__restrict__ psi,
                             const double *__restrict__ dudt, const double *__restrict__ dvdt,
                             const double *__restrict__ dwdt, const double *__restrict__ dudto,
                             const double *__restrict__ dvdto, const double *__restrict__ dwdto,
                             double *__restrict__ u, double *__restrict__ v, double *__restrict__ w) {
  CELL_IJK(d);
  const long sy = d.sy, sz = d.sz;
  const double rhox_n = rho + drho * 0.5 * (psio[q] + psio[q + 1]);
  const double rhoy_n = rho + drho * 0.5 * (psio[q] + psio[q + sy]);
  const double rhoz_n = rho + drho * 0.5 * (psio[q] + psio[q + sz]);
  const double rhox_p = rho + drho * 0.5 * (psi[q] + psi[q + 1]);
  const double rhoy_p = rho + drho * 0.5 * (psi[q] + psi[q + sy]);
  const double rhoz_p = rho + drho * 0.5 * (psi[q] + psi[q + sz]);
  u[q] = u[q] * rhox_n / rhox_p + (factor1 * dudt[r] + factor2 * dudto[r]) / rhox_p;
  v[q] = v[q] * rhoy_n / rhoy_p + (factor1 * dvdt[r] + factor2 * dvdto[r]) / rhoy_p;
  w[q] = w[q] * rhoz_n / rhoz_p + (factor1 * dwdt[r] + factor2 * dwdto[r]) / rhoz_p;
}

// src/rk.f90:118-130
__global__ void k_rk_mom_oth(fz_dims d, double factor12, double rho, double drho, const double *__restrict__ psi,
                             const double *__restrict__ dudt, const double *__restrict__ dvdt,
                             const double *__restrict__ dwdt, double *__restrict__ u, double *__restrict__ v,
                             double *__restrict__ w) {
  CELL_IJK(d);
  const long sy = d.sy, sz = d.sz;
  const double rhox_p = rho + drho * 0.5 * (psi[q] + psi[q + 1]);
  const double rhoy_p = rho + drho * 0.5 * (psi[q] + psi[q + sy]);
  const double rhoz_p = rho + drho * 0.5 * (psi[q] + psi[q + sz]);
  u[q] = u[q] + factor12 * dudt[r] / rhox_p;
  v[q] = v[q] + factor12 * dvdt[r] / rhoy_p;
  w[q] = w[q] + factor12 * dwdt[r] / rhoz_p;
}

// src/rk.f90:240-247
__global__ void k_rk_psi_ab2(fz_dims d, double factor1, double factor2, const double *__restrict__ dpsidt,
                             const double *__restrict__ dpsidto, double *__restrict__ psi) {
  CELL_IJK(d);
  psi[q] = psi[q] + factor1 * dpsidt[r] + factor2 * dpsidto[r];
}

// lscale_forcing, src/forcing.f90:27-63
__global__ void k_forcing(fz_dims d, int ftype, double f0, double factor_v, const double *__restrict__ stc, double l1,
                          double l2, double l3, double dl1, double dl2, const double *__restrict__ zc,
                          double *__restrict__ u, double *__restrict__ v, double *__restrict__ w) {
  CELL_IJK(d);
  (void)r;
  const double factor = stc ? stc[FZ_STC_DT] : factor_v;  // the forcing is scaled by dt (src/main.f90:485)
  const double k0 = 2.0;
  const double xxc = (double)((float)i - 0.5f) * dl1 / l1;
  const double yyc = (double)((float)j - 0.5f) * dl2 / l2;
  const double zzc = zc[k] / l3;
  if (ftype == 1) {  // ABC, a=b=c=1
    u[q] = u[q] + factor * (f0 * (1.0 * fzm::sin(k0 * zzc) + 1.0 * fzm::cos(k0 * yyc)));
    v[q] = v[q] + factor * (f0 * (1.0 * fzm::sin(k0 * xxc) + 1.0 * fzm::cos(k0 * zzc)));
    w[q] = w[q] + factor * (f0 * (1.0 * fzm::sin(k0 * yyc) + 1.0 * fzm::cos(k0 * xxc)));
  } else {
    const double xxf = xxc + 0.5 * dl1 / l1;
    const double yyf = yyc + 0.5 * dl2 / l2;
    u[q] = u[q] + factor * (f0 * fzm::sin(k0 * xxf) * fzm::cos(k0 * yyc) * fzm::cos(k0 * zzc));
    v[q] = v[q] - factor * (f0 * fzm::cos(k0 * xxc) * fzm::sin(k0 * yyf) * fzm::cos(k0 * zzc));
  }
}

// The forcing's sines and cosines each depend on one coordinate: tables over
// i, j and k (FT_ROWS rows of nt = max(n1, n2, n3) + 2 values, index = cell
// index), made once per context with the per-cell kernel's expressions, so
// k_forcing_tab adds the same values without evaluating them per cell.
enum { FT_SXC, FT_CXC, FT_SXF, FT_CXF, FT_SYC, FT_CYC, FT_SYF, FT_CYF, FT_SZC, FT_CZC, FT_ROWS };
__global__ void k_forcing_tables(fz_dims d, int nt, double l1, double l2, double l3, double dl1, double dl2,
                                 const double *__restrict__ zc, double *__restrict__ tab) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < 1 || n >= nt) return;
  const double k0 = 2.0;
  if (n <= d.n1) {
    const double xxc = (double)((float)n - 0.5f) * dl1 / l1;
    const double xxf = xxc + 0.5 * dl1 / l1;
    tab[FT_SXC * nt + n] = fzm::sin(k0 * xxc);
    tab[FT_CXC * nt + n] = fzm::cos(k0 * xxc);
    tab[FT_SXF * nt + n] = fzm::sin(k0 * xxf);
    tab[FT_CXF * nt + n] = fzm::cos(k0 * xxf);
  }
  if (n <= d.n2) {
    const double yyc = (double)((float)n - 0.5f) * dl2 / l2;
    const double yyf = yyc + 0.5 * dl2 / l2;
    tab[FT_SYC * nt + n] = fzm::sin(k0 * yyc);
    tab[FT_CYC * nt + n] = fzm::cos(k0 * yyc);
    tab[FT_SYF * nt + n] = fzm::sin(k0 * yyf);
    tab[FT_CYF * nt + n] = fzm::cos(k0 * yyf);
  }
  if (n <= d.n3) {
    const double zzc = zc[n] / l3;
    tab[FT_SZC * nt + n] = fzm::sin(k0 * zzc);
    tab[FT_CZC * nt + n] = fzm::cos(k0 * zzc);
  }
}
__global__ void k_forcing_tab(fz_dims d, int ftype, double f0, double factor_v, const double *__restrict__ stc,
                              int nt, const double *__restrict__ tab, double *__restrict__ u,
                              double *__restrict__ v, double *__restrict__ w) {
  CELL_IJK(d);
  (void)r;
  const double factor = stc ? stc[FZ_STC_DT] : factor_v;
  auto T = [&](int row, int n) { return tab[row * nt + n]; };
  if (ftype == 1) {  // ABC, a=b=c=1
    u[q] = u[q] + factor * (f0 * (1.0 * T(FT_SZC, k) + 1.0 * T(FT_CYC, j)));
    v[q] = v[q] + factor * (f0 * (1.0 * T(FT_SXC, i) + 1.0 * T(FT_CZC, k)));
    w[q] = w[q] + factor * (f0 * (1.0 * T(FT_SYC, j) + 1.0 * T(FT_CXC, i)));
  } else {
    u[q] = u[q] + factor * (f0 * T(FT_SXF, i) * T(FT_CYC, j) * T(FT_CZC, k));
    v[q] = v[q] - factor * (f0 * T(FT_CXC, i) * T(FT_SYF, j) * T(FT_CZC, k));
  }
}

// init2fl sphere/cylinder/film branch, src/two_fluid.f90:392-436
__global__ void k_init_spheres(fz_dims d, const double *__restrict__ zc, double dl1, double dl2, double l1, double l2,
                               double l3, int ip1, int ip2, int ip3, int id1, int id2, int id3, double sdist0,
                               double xxc, double yyc, double zzc, double rad, double seps, double *__restrict__ psi) {
  CELL_IJK(d);
  (void)r;
  const double z = zc[k];
  const double y = (double)((float)j - 0.5f) * dl2;
  const double x = (double)((float)i - 0.5f) * dl1;
  double sdistmin = sdist0;
  for (int kk = -1; kk <= 1; kk++)
    for (int jj = -1; jj <= 1; jj++)
      for (int ii = -1; ii <= 1; ii++) {
        const double ax = x + (double)(ii * ip1) * l1 - xxc;
        const double ay = y + (double)(jj * ip2) * l2 - yyc;
        const double az = z + (double)(kk * ip3) * l3 - zzc;
        const double sd = sqrt(ax * ax * id1 + ay * ay * id2 + az * az * id3) - rad;
        if (fabs(sd) <= sdistmin) sdistmin = sd;
      }
  const double ps = .5 * (1.0 + fzm::tanh(sdistmin / (2.0 * seps)));
  psi[q] = fmin(psi[q], ps);
}

__global__ void k_init_film(fz_dims d, const double *__restrict__ zc, double zmid, double seps,
                            double *__restrict__ psi) {
  CELL_IJK(d);
  (void)r;
  psi[q] = .5 * (1.0 + fzm::tanh((zc[k] - zmid) / (2.0 * seps)));
}

__global__ void k_fill(long nh, double val, double *__restrict__ p) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nh; q += (long)gridDim.x * blockDim.x) p[q] = val;
}

__global__ void k_one_minus(long nh, double *__restrict__ p) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nh; q += (long)gridDim.x * blockDim.x)
    p[q] = 1.0 - p[q];
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
static inline int flat_grid(long n) {
  long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  return (int)(b < 1 ? 1 : b);
}

extern "C" int fz_fillps(fz_ctx *c, double dti, double rho0, const double *u, const double *v, const double *w,
                         double *p) {
  if (!c || !u || !v || !w || !p) return fz_set_error(FZ_ERR_ARG, "fz_fillps: null argument");
  FZ_LAUNCH(c, "fillps", k_fillps, fz_grid_interior(c->d), fz_block3(), 0, c->d, c->dzfi, dti, dti * c->dli[0],
                     dti * c->dli[1], rho0, u, v, w, p);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_correc(fz_ctx *c, double rho0, const double rho12[2], double dt, const double *p,
                         const double *psi, double *u, double *v, double *w) {
  (void)rho12;
  (void)psi;
  if (!c || !p || !u || !v || !w) return fz_set_error(FZ_ERR_ARG, "fz_correc: null argument");
  FZ_LAUNCH(c, "correc", k_correc, fz_grid_halo(c->d), fz_block3(), 0, c->d, c->dzci, dt * c->dli[0],
                     dt * c->dli[1], dt, rho0, p, u, v, w);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_updatep(fz_ctx *c, const double *pp, double *p) {
  if (!c || !pp || !p) return fz_set_error(FZ_ERR_ARG, "fz_updatep: null argument");
  FZ_LAUNCH(c, "updatep", k_updatep, dim3(flat_grid(c->nh)), dim3(256), 0, c->nh, pp, p);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_updt_rhs_b(fz_ctx *c, double *p) {
  if (!c || !p) return fz_set_error(FZ_ERR_ARG, "fz_updt_rhs_b: null argument");
  if (!c->pois) return fz_set_error(FZ_ERR_STATE, "fz_updt_rhs_b: fz_initsolver has not been called");
  const fz_dims &d = c->d;
  // x is a boundary direction of the x-pencil; y and z when not periodic
  const double *rb[3] = {c->pois->rhsbx, c->pois->rhsby, c->pois->rhsbz};
  const int na[3] = {d.n2, d.n1, d.n1}, nb[3] = {d.n3, d.n3, d.n2}, nn[3] = {d.n1, d.n2, d.n3};
  for (int dir = 0; dir < 3; dir++) {
    if (dir > 0 && c->per[dir]) continue;
    for (int ib = 0; ib < 2; ib++) {
      const double *rhs = rb[dir] + (long)na[dir] * nb[dir] * ib;
      FZ_LAUNCH(c, "add_face", k_add_face, dim3((na[dir] + 63) / 64, nb[dir]), dim3(64), 0, d, dir,
                         ib == 0 ? 1 : nn[dir], rhs, p);
      FZ_LAUNCHED();
    }
  }
  return FZ_OK;
}

int fz_acdi_transport_dev(fz_ctx *c, const double *d_gam, double gam, double seps, const double *u, const double *v,
                          const double *w, const double *normx, const double *normy, const double *normz,
                          const double *phi, const double *psi, double *dpsidt, double *flux_x, double *flux_y,
                          double *flux_z) {
  FZ_LAUNCH(c, "acdi_transport", k_acdi_transport, fz_grid_interior(c->d), fz_block3(), 0, c->d, c->dzci, c->dzfi,
                     c->dli[0], c->dli[1], d_gam, gam, seps, u, v, w, normx, normy, normz, phi, psi, dpsidt, flux_x,
                     flux_y, flux_z);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_acdi_transport_pf(fz_ctx *c, double gam, double seps, const double *u, const double *v,
                                    const double *w, const double *normx, const double *normy, const double *normz,
                                    const double *phi, const double *psi, double *dpsidt, double *flux_x,
                                    double *flux_y, double *flux_z) {
  if (!c || !u || !v || !w || !normx || !normy || !normz || !phi || !psi || !dpsidt || !flux_x || !flux_y || !flux_z)
    return fz_set_error(FZ_ERR_ARG, "fz_acdi_transport_pf: null argument");
  return fz_acdi_transport_dev(c, nullptr, gam, seps, u, v, w, normx, normy, normz, phi, psi, dpsidt, flux_x, flux_y,
                               flux_z);
}

// (rflag: or null; see k_acdi_phi)
int fz_acdi_cmpt_phi_rf(fz_ctx *c, double seps, const double *psi, double *phi, unsigned *rflag) {
  FZ_LAUNCH(c, "acdi_phi", k_acdi_phi<double>, dim3(flat_grid(c->nh)), dim3(256), 0, c->nh, seps, psi, phi,
            kNoBRef, rflag);
  FZ_LAUNCHED();
  return FZ_OK;
}
extern "C" int fz_acdi_cmpt_phi(fz_ctx *c, double seps, const double *psi, double *phi) {
  if (!c || !psi || !phi) return fz_set_error(FZ_ERR_ARG, "fz_acdi_cmpt_phi: null argument");
  return fz_acdi_cmpt_phi_rf(c, seps, psi, phi, nullptr);
}

// FP32 phase field of the mixed-precision step
int fz_acdi_cmpt_phi_f32(fz_ctx *c, double seps, const float *psi, float *phi) {
  FZ_LAUNCH(c, "acdi_phi_f32", k_acdi_phi<float>, dim3(flat_grid(c->nh)), dim3(256), 0, c->nh, seps, psi, phi,
            kNoBRef, (unsigned *)nullptr);
  FZ_LAUNCHED();
  return FZ_OK;
}

int fz_youngs_f32(fz_ctx *c, const float *psi, float *normx, float *normy, float *normz, float *kappa) {
  return launch_youngs<float>(c, "youngs_f32", psi, normx, normy, normz, kappa);
}
// the normals and curvature of a step, with the periodic x halo when xper
// (rflag: the range flag of k_acdi_phi for this phi, or null)
template <typename T>
int fz_youngs_step(fz_ctx *c, const T *psi, T *normx, T *normy, T *normz, T *kappa, int xper, const unsigned *rflag) {
  if (!psi || !normx || !normy || !normz || !kappa) return fz_set_error(FZ_ERR_ARG, "youngs: null argument");
  return launch_youngs<T>(c, sizeof(T) == 4 ? "youngs_f32" : "youngs", psi, normx, normy, normz, kappa, kNoBRef, 1,
                          xper, rflag);
}
template int fz_youngs_step<double>(fz_ctx *, const double *, double *, double *, double *, double *, int,
                                    const unsigned *);
template int fz_youngs_step<float>(fz_ctx *, const float *, float *, float *, float *, float *, int,
                                   const unsigned *);

extern "C" int fz_vof_thinc_cmpt_phi(fz_ctx *c, double beta, const double *psi, double *phi) {
  if (!c || !psi || !phi) return fz_set_error(FZ_ERR_ARG, "fz_vof_thinc_cmpt_phi: null argument");
  FZ_LAUNCH(c, "vof_phi", k_vof_phi<double>, dim3(flat_grid(c->nh)), dim3(256), 0, c->nh, beta, psi, phi);
  FZ_LAUNCHED();
  return FZ_OK;
}

template <typename T>
static int vof_transport(fz_ctx *c, double beta, const double *u, const double *v, const double *w,
                         const T *normx, const T *normy, const T *normz, const T *phi, T *dpsidt, T *flux_x,
                         T *flux_y, T *flux_z) {
  const fz_dims &d = c->d;
  dim3 g((d.n1 + 1 + 63) / 64, (d.n2 + 1 + 3) / 4, d.n3 + 1);
  FZ_LAUNCH(c, sizeof(T) == 4 ? "vof_flux_f32" : "vof_flux", k_vof_flux<T>, g, fz_block3(), 0, d, beta, u, v, w,
            normx, normy, normz, phi, flux_x, flux_y, flux_z);
  FZ_LAUNCHED();
  FZ_LAUNCH(c, sizeof(T) == 4 ? "vof_div_f32" : "vof_div", k_vof_div<T>, fz_grid_interior(d), fz_block3(), 0, d,
            c->dzfi, c->dli[0], c->dli[1], flux_x, flux_y, flux_z, dpsidt);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_vof_thinc_transport_psi(fz_ctx *c, double beta, const double *u, const double *v, const double *w,
                                          const double *normx, const double *normy, const double *normz,
                                          const double *phi, const double *psi, double *dpsidt, double *flux_x,
                                          double *flux_y, double *flux_z) {
  (void)psi;
  if (!c || !u || !v || !w || !normx || !normy || !normz || !phi || !dpsidt || !flux_x || !flux_y || !flux_z)
    return fz_set_error(FZ_ERR_ARG, "fz_vof_thinc_transport_psi: null argument");
  return vof_transport<double>(c, beta, u, v, w, normx, normy, normz, phi, dpsidt, flux_x, flux_y, flux_z);
}

int fz_vof_phi_f32(fz_ctx *c, double beta, const float *psi, float *phi) {
  FZ_LAUNCH(c, "vof_phi_f32", k_vof_phi<float>, dim3(flat_grid(c->nh)), dim3(256), 0, c->nh, beta, psi, phi);
  FZ_LAUNCHED();
  return FZ_OK;
}


extern "C" int fz_clip_field(fz_ctx *c, const double minmax[2], double *p) {
  if (!c || !minmax || !p) return fz_set_error(FZ_ERR_ARG, "fz_clip_field: null argument");
  FZ_LAUNCH(c, "clip", k_clip<double>, dim3(flat_grid(c->nh)), dim3(256), 0, c->nh, minmax[0], minmax[1], p);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_cmpt_norm_curv_youngs(fz_ctx *c, const double *psi, double *normx, double *normy, double *normz,
                                        double *kappa) {
  if (!c || !psi || !normx || !normy || !normz || !kappa)
    return fz_set_error(FZ_ERR_ARG, "fz_cmpt_norm_curv_youngs: null argument");
  return launch_youngs<double>(c, "youngs", psi, normx, normy, normz, kappa);
}

extern "C" int fz_mom_xyz_ad(fz_ctx *c, const double rho12[2], const double mu12[2], const double *psiflx_x,
                             const double *psiflx_y, const double *psiflx_z, const double *u, const double *v,
                             const double *w, const double *psi, double *dudt, double *dvdt, double *dwdt) {
  if (!c || !rho12 || !mu12 || !psiflx_x || !psiflx_y || !psiflx_z || !u || !v || !w || !psi || !dudt || !dvdt ||
      !dwdt)
    return fz_set_error(FZ_ERR_ARG, "fz_mom_xyz_ad: null argument");
  FZ_LAUNCH(c, "mom_ad", k_mom_ad, fz_grid_interior(c->d), fz_block3(), 0, c->d, c->dzci, c->dzfi, c->dli[0],
                     c->dli[1], rho12[1], rho12[0] - rho12[1], mu12[1], mu12[0] - mu12[1], psiflx_x, psiflx_y,
                     psiflx_z, u, v, w, psi, dudt, dvdt, dwdt);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_mom_xyz_oth(fz_ctx *c, double dt_r, const double rho12[2], const double beta12[2],
                              const double bforce[3], const double gacc[3], double sigma, double rho0,
                              const double rho_av[3], const double *p, const double *psi, const double *kappa,
                              const double *pn, const double *po, double *dudt, double *dvdt, double *dwdt) {
  (void)beta12;
  if (!c || !rho12 || !bforce || !gacc || !rho_av || !p || !psi || !kappa || !pn || !po || !dudt || !dvdt || !dwdt)
    return fz_set_error(FZ_ERR_ARG, "fz_mom_xyz_oth: null argument");
  const double surf_factor = sigma * 2.0 / (rho12[0] + rho12[1]);
  FZ_LAUNCH(c, "mom_oth", k_mom_oth, fz_grid_interior(c->d), fz_block3(), 0, c->d, c->dzci, c->dli[0], c->dli[1],
                     dt_r, rho12[1], rho12[0] - rho12[1], surf_factor, rho0, bforce[0], bforce[1], bforce[2], gacc[0],
                     gacc[1], gacc[2], rho_av[0], rho_av[1], rho_av[2], p, psi, kappa, pn, po, dudt, dvdt, dwdt);
  FZ_LAUNCHED();
  return FZ_OK;
}

// lscale_forcing with dt from device memory when stc is given (captured steps)
int fz_launch_forcing(fz_ctx *c, int ftype, double alpha, double dt, const double *stc, double *u, double *v,
                      double *w) {
  const double f0 = alpha * ((2.0 * FZ_PI / c->l[2]) * (2.0 * FZ_PI / c->l[2]));
  const int nt = std::max(c->d.n1, std::max(c->d.n2, c->d.n3)) + 2;
  if (!c->ftab) {  // the tables, made on the first call outside a graph capture
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    FZ_HIP(hipStreamIsCapturing(c->st, &cap));
    if (cap == hipStreamCaptureStatusNone) {
      FZ_HIP(hipMalloc((void **)&c->ftab, sizeof(double) * FT_ROWS * nt));
      FZ_LAUNCH(c, "forcing_tables", k_forcing_tables, dim3((nt + 255) / 256), dim3(256), 0, c->d, nt, c->l[0],
                c->l[1], c->l[2], c->dl[0], c->dl[1], c->zc, c->ftab);
      FZ_LAUNCHED();
    }
  }
  if (c->ftab)
    FZ_LAUNCH(c, "forcing", k_forcing_tab, fz_grid_interior(c->d), fz_block3(), 0, c->d, ftype, f0, dt, stc, nt,
              c->ftab, u, v, w);
  else
    FZ_LAUNCH(c, "forcing", k_forcing, fz_grid_interior(c->d), fz_block3(), 0, c->d, ftype, f0, dt, stc, c->l[0],
              c->l[1], c->l[2], c->dl[0], c->dl[1], c->zc, u, v, w);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_lscale_forcing(fz_ctx *c, int ftype, double alpha, double dt, double *u, double *v, double *w) {
  if (!c || !u || !v || !w) return fz_set_error(FZ_ERR_ARG, "fz_lscale_forcing: null argument");
  return fz_launch_forcing(c, ftype, alpha, dt, nullptr, u, v, w);
}

// rk_2fl / rk AB2 kernels for the case driver and fz_rk / fz_rk_2fl
int fz_launch_rk_psi(fz_ctx *c, double f1, double f2, const double *dpsidt, const double *dpsidto, double *psi) {
  FZ_LAUNCH(c, "rk_psi_ab2", k_rk_psi_ab2, fz_grid_interior(c->d), fz_block3(), 0, c->d, f1, f2, dpsidt, dpsidto, psi);
  FZ_LAUNCHED();
  return FZ_OK;
}
int fz_launch_rk_mom(fz_ctx *c, double f1, double f2, const double rho12[2], const double *psio, const double *psi,
                     const double *du, const double *dv, const double *dw, const double *duo, const double *dvo,
                     const double *dwo, double *u, double *v, double *w) {
  FZ_LAUNCH(c, "rk_mom_ab2", k_rk_mom_ab2, fz_grid_interior(c->d), fz_block3(), 0, c->d, f1, f2, rho12[1],
                     rho12[0] - rho12[1], psio, psi, du, dv, dw, duo, dvo, dwo, u, v, w);
  FZ_LAUNCHED();
  return FZ_OK;
}
int fz_launch_rk_oth(fz_ctx *c, double f12, const double rho12[2], const double *psi, const double *du,
                     const double *dv, const double *dw, double *u, double *v, double *w) {
  FZ_LAUNCH(c, "rk_mom_oth", k_rk_mom_oth, fz_grid_interior(c->d), fz_block3(), 0, c->d, f12, rho12[1],
                     rho12[0] - rho12[1], psi, du, dv, dw, u, v, w);
  FZ_LAUNCHED();
  return FZ_OK;
}
int fz_launch_fill(fz_ctx *c, long n, double val, double *p) {
  FZ_LAUNCH(c, "fill", k_fill, dim3(flat_grid(n)), dim3(256), 0, n, val, p);
  FZ_LAUNCHED();
  return FZ_OK;
}

extern "C" int fz_init2fl(fz_ctx *c, const fz_params *prm, double seps, double *psi) {
  if (!c || !prm || !psi) return fz_set_error(FZ_ERR_ARG, "fz_init2fl: null argument");
  const fz_dims &d = c->d;
  FZ_TRY(fz_launch_fill(c, c->nh, 1.0, psi));
  int is_sphere = 0, is_swap = 0, is_dim[3] = {1, 1, 1};
  switch (prm->inipsi) {
    case 0: break;
    case 1: FZ_TRY(fz_launch_fill(c, c->nh, 0.0, psi)); break;
    case 2: is_sphere = 1; break;
    case 3: is_sphere = 1; is_dim[1] = 0; break;
    case 4: is_sphere = 1; is_dim[0] = 0; is_dim[1] = 0; break;
    case 5: is_sphere = 1; is_swap = 1; break;
    case 6: is_sphere = 1; is_swap = 1; is_dim[1] = 0; break;
    case 7: is_sphere = 1; is_swap = 1; is_dim[0] = 0; is_dim[1] = 0; break;
    case 8:
      FZ_LAUNCH(c, "init_film", k_init_film, fz_grid_interior(d), fz_block3(), 0, d, c->zc, c->l[2] / 2.0, seps, psi);
      FZ_LAUNCHED();
      break;
    default: return fz_set_error(FZ_ERR_ARG, "fz_init2fl: unknown inipsi %d", prm->inipsi);
  }
  if (is_sphere) {
    int nsph = prm->nspheres;
    double def[4];
    const double *sph = prm->spheres;
    if (nsph < 0 || nsph > 64) return fz_set_error(FZ_ERR_ARG, "fz_init2fl: nspheres must be in 0..64");
    if (nsph == 0) {  // no spheres.in: one in the centre, src/two_fluid.f90:393-401
      double lmin = 1e300;
      for (int q = 0; q < 3; q++)
        if (is_dim[q] && c->l[q] < lmin) lmin = c->l[q];
      def[0] = c->l[0] / 2;
      def[1] = c->l[1] / 2;
      def[2] = c->l[2] / 2;
      def[3] = 0.25 * lmin;
      sph = def;
      nsph = 1;
    }
    int ip[3];
    for (int q = 0; q < 3; q++) ip[q] = (prm->cbcpsi[2 * q] == 'P' && prm->cbcpsi[2 * q + 1] == 'P');
    double lmax = 0.0;
    for (int q = 0; q < 3; q++)
      if (c->l[q] * is_dim[q] > lmax) lmax = c->l[q] * is_dim[q];
    for (int s = 0; s < nsph; s++) {
      FZ_LAUNCH(c, "init_spheres", k_init_spheres, fz_grid_interior(d), fz_block3(), 0, d, c->zc, c->dl[0], c->dl[1],
                         c->l[0], c->l[1], c->l[2], ip[0], ip[1], ip[2], is_dim[0], is_dim[1], is_dim[2], lmax * 2.0,
                         sph[4 * s], sph[4 * s + 1], sph[4 * s + 2], sph[4 * s + 3], seps, psi);
      FZ_LAUNCHED();
    }
  }
  if (is_swap) {
    FZ_LAUNCH(c, "one_minus", k_one_minus, dim3(flat_grid(c->nh)), dim3(256), 0, c->nh, psi);
    FZ_LAUNCHED();
  }
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// The kernels' FP64 routines evaluated on the device, element by element, for
// tests against the oracle's copies: fn 0 exp, 1 expm1, 2 log, 3 tanh, 4 sin,
// 5 cos, 6 the in-range division x / y, 7 the compiler's division x / y
__global__ void k_math_eval(int fn, long n, const double *__restrict__ x, const double *__restrict__ y,
                            double *__restrict__ out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double a = x[q];
  double r;
  switch (fn) {
    case 0: r = fzm::exp(a); break;
    case 1: r = fzm::expm1(a); break;
    case 2: r = fzm::log(a); break;
    case 3: r = fzm::tanh(a); break;
    case 4: r = fzm::sin(a); break;
    case 5: r = fzm::cos(a); break;
    case 6: r = fzm::div_inrange(a, y[q]); break;
    default: r = a / y[q]; break;
  }
  out[q] = r;
}

extern "C" int fz_math_eval(fz_ctx *c, int fn, long n, const double *x, const double *y, double *out) {
  if (!c || !x || !out || n < 0 || fn < 0 || fn > 7 || (fn >= 6 && !y))
    return fz_set_error(FZ_ERR_ARG, "fz_math_eval: bad argument");
  if (n == 0) return FZ_OK;
  const long nb = (n + 255) / 256;
  if (nb > 0x7fffffffL) return fz_set_error(FZ_ERR_ARG, "fz_math_eval: n too large");
  FZ_LAUNCH(c, "math_eval", k_math_eval, dim3((unsigned)nb), dim3(256), 0, fn, n, x, y, out);
  FZ_LAUNCHED();
  return FZ_OK;
}

__global__ void k_clear_word(unsigned *__restrict__ w) {
  if (threadIdx.x == 0) w[0] = 0u;
}

// (tests) acdi_cmpt_phi then cmpt_norm_curv_youngs as the FP64 step runs
// them: phi from psi, raising the range flag (the device word *flag, cleared
// first) where some phi is neither zero nor of a magnitude in [2^-100, 2^20],
// then the Youngs kernels the flag selects
extern "C" int fz_phase_geometry_eval(fz_ctx *c, double seps, const double *psi, double *phi, double *normx,
                                      double *normy, double *normz, double *kappa, unsigned *flag) {
  if (!c || !psi || !phi || !normx || !normy || !normz || !kappa || !flag)
    return fz_set_error(FZ_ERR_ARG, "fz_phase_geometry_eval: null argument");
  FZ_LAUNCH(c, "clear_word", k_clear_word, dim3(1), dim3(64), 0, flag);
  FZ_LAUNCHED();
  FZ_TRY(fz_acdi_cmpt_phi_rf(c, seps, psi, phi, flag));
  return fz_youngs_step<double>(c, phi, normx, normy, normz, kappa, 0, flag);
}

// ---------------------------------------------------------------------------
// Launches of the step's kernels for a batch of ncase same-shaped cases
// (fz_case_batch): the grids of the cases stacked in z (or one block row per
// case for the flat kernels), each block taking its case's buffers and
// scalars from the table br. c is the first case's context: the grid, the
// launch parameters and the stream. Each case's values are those of its own
// launches bit for bit (the same kernels and arithmetic).
template <typename T>
int fz_batch_acdi_step(fz_ctx *c, fz_bref br, int ncase) {
  FZ_TRY(fits_buffers(c, "acdi_step"));
  constexpr int KZ = 16;
  dim3 grid((c->d.n1 + ACD_MX - 1) / ACD_MX, (c->d.n2 + ACD_MY - 1) / ACD_MY, (c->d.n3 + KZ - 1) / KZ);
  br.gz = grid.z;
  grid.z *= (unsigned)ncase;
  FZ_LAUNCH(c, sizeof(T) == 4 ? "acdi_step_f32" : "acdi_step", (k_acdi_lds<T, KZ, true>), grid, dim3(ACD_MX, ACD_MY), 0,
            c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1], (const double *)nullptr, 0.0, 0.0, 0.0, 0.0,
            (const double *)nullptr, 0, (const double *)nullptr, (const double *)nullptr, (const double *)nullptr,
            (const T *)nullptr, (const T *)nullptr, (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
            (T *)nullptr, (T *)nullptr, (T *)nullptr, (T *)nullptr, (T *)nullptr, br);
  FZ_LAUNCHED();
  return FZ_OK;
}
template int fz_batch_acdi_step<double>(fz_ctx *, fz_bref, int);
template int fz_batch_acdi_step<float>(fz_ctx *, fz_bref, int);

template <typename T>
int fz_batch_phase_geometry(fz_ctx *c, fz_bref br, int ncase) {
  if (ncase > 65535) return fz_set_error(FZ_ERR_ARG, "batched step: at most 65535 cases");
  FZ_LAUNCH(c, sizeof(T) == 4 ? "acdi_phi_f32" : "acdi_phi", (k_acdi_phi<T, true>),
            dim3((unsigned)flat_grid(c->nh), (unsigned)ncase), dim3(256), 0, c->nh, 0.0, (const T *)nullptr,
            (T *)nullptr, br, (unsigned *)nullptr);
  FZ_LAUNCHED();
  return launch_youngs<T>(c, sizeof(T) == 4 ? "youngs_f32" : "youngs", (const T *)nullptr, (T *)nullptr,
                          (T *)nullptr, (T *)nullptr, (T *)nullptr, br, ncase);
}
template int fz_batch_phase_geometry<double>(fz_ctx *, fz_bref, int);
template int fz_batch_phase_geometry<float>(fz_ctx *, fz_bref, int);

// the momentum kernel of the first case's variant (the LDS kernel, or the
// plain one of the given occupancy)
template <typename PT, typename HT>
int fz_batch_mom(fz_ctx *c, fz_bref br, int ncase) {
  FZ_TRY(fits_buffers(c, "mom_fused"));
  int variant = c->mom_variant;
  if (variant == 0) variant = 8;
  const char *name = sizeof(PT) == 4 ? "mom_fused_f32" : "mom_fused";
#define FZ_MOM_NULLS                                                                                              \
  0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0, (const double *)nullptr, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, \
      0.0, 0.0, (const PT *)nullptr, (const PT *)nullptr, (const PT *)nullptr, (const double *)nullptr,            \
      (const double *)nullptr, (const double *)nullptr, (const PT *)nullptr, (const PT *)nullptr,                  \
      (const double *)nullptr, (const double *)nullptr, (const PT *)nullptr, (HT *)nullptr, (HT *)nullptr,         \
      (HT *)nullptr, (double *)nullptr, (double *)nullptr, (double *)nullptr, br
  if (variant == 8) {
    constexpr int KZ = 16;
    dim3 grid((c->d.n1 + MOM_MX - 1) / MOM_MX, (c->d.n2 + MOM_MY - 1) / MOM_MY, (c->d.n3 + KZ - 1) / KZ);
    br.gz = grid.z;
    grid.z *= (unsigned)ncase;
    FZ_LAUNCH(c, name, (k_mom_lds<PT, HT, KZ, sizeof(PT) == 4 ? 3 : 1, true, sizeof(PT) == 4>), grid,
              dim3(MOM_MX, MOM_MY), 0, c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1], FZ_MOM_NULLS);
  } else {
    auto kern = variant == 4   ? k_mom_fused_w4<PT, HT, true>
                : variant == 5 ? k_mom_fused_w5<PT, HT, true>
                : variant == 1 ? k_mom_fused_st<PT, HT, true>
                               : k_mom_fused<PT, HT, true>;
    dim3 grid = fz_grid_interior(c->d);
    br.gz = grid.z;
    grid.z *= (unsigned)ncase;
    FZ_LAUNCH(c, name, kern, grid, fz_block3(), 0, c->d, c->dzci, c->dzfi, c->dli[0], c->dli[1], FZ_MOM_NULLS);
  }
#undef FZ_MOM_NULLS
  FZ_LAUNCHED();
  return FZ_OK;
}
template int fz_batch_mom<double, double>(fz_ctx *, fz_bref, int);
template int fz_batch_mom<float, float>(fz_ctx *, fz_bref, int);

// the correction step (the LDS kernel of the first case's tile height) and
// gamma from its slots
int fz_batch_correc_gamma(fz_ctx *c, fz_bref br, int ncase, bool gamma) {
  FZ_TRY(fits_buffers(c, "correc_step"));
  constexpr int KZ = 16;
  const int my = c->correc_my;
  dim3 grid((c->d.n1 + 2 + COR_MX - 1) / COR_MX, (c->d.n2 + 2 + my - 1) / my, (c->d.n3 + 2 + KZ - 1) / KZ);
  fz_bref b = br;
  b.gz = grid.z;
  grid.z *= (unsigned)ncase;
#define FZ_CORREC_LDS_B(MY)                                                                                        \
  FZ_LAUNCH(c, "correc_step", (k_correc_lds<KZ, MY, true>), grid, dim3(COR_MX, MY), 0, c->d, c->dzci, 0.0, 0.0, 0.0, \
            c->dli[0], c->dli[1], (const double *)nullptr, 0.0, (const double *)nullptr, (const double *)nullptr,   \
            (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, (double *)nullptr,          \
            (double *)nullptr, (double *)nullptr, (double *)nullptr, (unsigned long long *)nullptr, b, 0, 0u)
  if (my == 16) FZ_CORREC_LDS_B(16);
  else if (my == 8) FZ_CORREC_LDS_B(8);
  else FZ_CORREC_LDS_B(4);
#undef FZ_CORREC_LDS_B
  FZ_LAUNCHED();
  if (gamma) {
    FZ_LAUNCH(c, "gamma_fin", k_gamma_fin<true>, dim3((unsigned)ncase), dim3(FZ_GAMMA_SLOTS), 0,
              (unsigned long long *)nullptr, 0.0, 0.0, (double *)nullptr, br, (unsigned *)nullptr);
    FZ_LAUNCHED();
  }
  return FZ_OK;
}

// fillps (src/fillps.f90:33-50) for a batch: the momentum step's velocity
// (the spare buffers) into each case's correction buffer, with its 1/dt and
// rho0; the expression of the fillps evaluated in the forward x transform
__global__ void __launch_bounds__(256) k_fillps_b(fz_dims d, const double *__restrict__ dzfi, double dxi, double dyi,
                                                  fz_bref br) {
  constexpr bool BT = true;
  CELL_IJKB(d, br, cs);
  (void)r;
  const fz_bcase &B = br.bt[cs];
  const double *__restrict__ u = (const double *)fz_brole(br, cs, BR_UN);
  const double *__restrict__ v = (const double *)fz_brole(br, cs, BR_VN);
  const double *__restrict__ w = (const double *)fz_brole(br, cs, BR_WN);
  double *__restrict__ p = (double *)fz_brole(br, cs, BR_DP);
  const double dti = B.stc[FZ_STC_DTI];
  const double dtidxi = dti * dxi, dtidyi = dti * dyi;
  p[q] = B.rho0 * ((w[q] - w[q - d.sz]) * dti * dzfi[k] + (v[q] - v[q - d.sy]) * dtidyi + (u[q] - u[q - 1]) * dtidxi);
}

int fz_batch_fillps(fz_ctx *c, fz_bref br, int ncase) {
  dim3 grid = fz_grid_interior(c->d);
  br.gz = grid.z;
  grid.z *= (unsigned)ncase;
  FZ_LAUNCH(c, "fillps", k_fillps_b, grid, fz_block3(), 0, c->d, c->dzfi, c->dli[0], c->dli[1], br);
  FZ_LAUNCHED();
  return FZ_OK;
}

// the next pointer state of every case of a batch
__global__ void k_batch_flip(int n, int *__restrict__ par) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) par[t] ^= 1;
}
int fz_batch_flip(fz_ctx *c, int *par, int ncase) {
  FZ_LAUNCH(c, "batch_flip", k_batch_flip, dim3((unsigned)((ncase + 255) / 256)), dim3(256), 0, ncase, par);
  FZ_LAUNCHED();
  return FZ_OK;
}
