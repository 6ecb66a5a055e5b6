// FFT-based pressure Poisson solver on one MI355X.
//
// Replaces initsolver (src/initsolver.f90:15-228), solver/solver_gpu
// (src/solver.f90:19-177, src/solver_gpu.f90:31-370) and the r2r transforms
// of src/fft.f90. The x and y transforms are batched line transforms done in
// LDS (Stockham autosort, mixed radix) with the FFTW r2r definitions the
// reference plans (find_fft, src/fft.f90:189-242); the z direction is the
// reference's Thomas recurrence, one column per thread. The pencil
// transposes of the reference are not needed: the transforms address the
// lines of one HBM-resident field directly.
#include <cmath>
#include <cstring>

#include "fz_internal.h"
#include "fz_math.h"

enum { K_R2HC = 0, K_HC2R = 1, K_REDFT01 = 4, K_REDFT10 = 5, K_REDFT11 = 6, K_RODFT01 = 8, K_RODFT10 = 9, K_RODFT11 = 10 };

struct RadList {
  int r[24];
};
struct LineMap {  // element e of line (l1,l2) sits at base + l1*s1 + l2*s2 + e*se
  long base, s1, s2, se;
};
// the arrays of a batch of same-shaped solves, one per blockIdx.z (device
// arrays of n pointers; src null: no batch)
struct Batch {
  const void *const *src;
  void *const *dst;
  int n;
};
// The line source of a forward x transform when it is not a plain array:
// fillps (src/fillps.f90:33-50) evaluated in the load (u, v, w halo'd
// fields; k = l2 + 1), or, for the residual iterations of the mixed-precision
// solve, a right-hand side minus the discrete operator applied to the current
// solution x (FP64). The optional slots take the maximum |value| (FP64 bits).
enum { SRC_ARRAY = 0, SRC_FILLPS = 1, SRC_FILLPS_RES = 2, SRC_B_RES = 3, SRC_B = 4 };
struct FillSrc {
  const double *u, *v, *w, *dzfi;
  double rho0, dti, dtidxi, dtidyi;
  long sy, sz;
  int on;
  const double *x, *b, *dzci;  // current solution (halo'd), right-hand side (halo'd layout)
  double dxi2, dyi2;
  unsigned long long *rmax;    // FZ_GAMMA_SLOTS slots, or null
  const double *stc;           // the case's device scalars (1/dt from there), or null
  double dxi, dyi;
};

// the FP64 value of the source at halo index a (z index k)
__device__ inline double src_value(const FillSrc &fs, long a, int k) {
  double val;
  if (fs.on == SRC_FILLPS || fs.on == SRC_FILLPS_RES) {
    const double dti = fs.stc ? fs.stc[FZ_STC_DTI] : fs.dti;
    const double dtidxi = fs.stc ? dti * fs.dxi : fs.dtidxi, dtidyi = fs.stc ? dti * fs.dyi : fs.dtidyi;
    val = fs.rho0 * ((fs.w[a] - fs.w[a - fs.sz]) * dti * fs.dzfi[k] + (fs.v[a] - fs.v[a - fs.sy]) * dtidyi +
                     (fs.u[a] - fs.u[a - 1]) * dtidxi);
  } else {
    val = fs.b[a];
  }
  if (fs.on == SRC_FILLPS_RES || fs.on == SRC_B_RES) {
    // the operator the transforms invert: 3-point second differences in x
    // and y, and in z the tridiagonal of src/initsolver.f90:153-182 (the
    // boundary rows follow from the halo values of homogeneous conditions)
    const double *x = fs.x;
    const double xc = x[a];
    const double lap = ((x[a + 1] - xc) - (xc - x[a - 1])) * fs.dxi2 +
                       ((x[a + fs.sy] - xc) - (xc - x[a - fs.sy])) * fs.dyi2 +
                       ((x[a + fs.sz] - xc) * fs.dzci[k] - (xc - x[a - fs.sz]) * fs.dzci[k - 1]) * fs.dzfi[k];
    val -= lap;
  }
  return val;
}

// maximum of m over the wave, one atomic per wave into the slots
__device__ inline void slots_max(unsigned long long *slots, double m, unsigned wave_id) {
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0)
    atomicMax(slots + (wave_id & (FZ_GAMMA_SLOTS - 1)), (unsigned long long)__double_as_longlong(m));
}

static __host__ __device__ inline bool kind_is_quarter(int kind) { return kind == K_REDFT11 || kind == K_RODFT11; }
static __host__ __device__ inline bool kind_is_forward(int kind) {
  return kind == K_R2HC || kind == K_REDFT10 || kind == K_RODFT10 || kind == K_REDFT11 || kind == K_RODFT11;
}

// Transforms are written once for the real type R: double for the plain
// solver, float for the mixed-precision one.
template <typename R>
struct CT;
template <>
struct CT<double> {
  using C = double2;
};
template <>
struct CT<float> {
  using C = float2;
};
__device__ inline double2 mkc(double x, double y) { return make_double2(x, y); }
__device__ inline float2 mkc(float x, float y) { return make_float2(x, y); }
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// radix-4 butterfly in place on twiddled inputs; sign -1 forward, +1 inverse
template <typename C, typename R>
__device__ inline void bfly4(C *v, R sign) {
  const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  const C s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
  const C r = mkc(-sign * d13.y, sign * d13.x);  // i*sign*d13
  v[0] = cadd(s02, s13);
  v[1] = cadd(d02, r);
  v[2] = csub(s02, s13);
  v[3] = csub(d02, r);
}

// One Stockham pass of radix R over nfft complex lines of length m (LDS to LDS).
template <typename T>
__device__ inline void stockham_pass(const typename CT<T>::C *__restrict__ X, typename CT<T>::C *__restrict__ Y,
                                     int nfft, int m, int R, int Ns, T sign, const typename CT<T>::C *__restrict__ tw) {
  using C = typename CT<T>::C;
  const int nb = m / R;
  const int total = nfft * nb;
  const int tstride = m / (Ns * R);
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int f = idx / nb, b = idx - f * nb;
    const C *x = X + (long)f * m;
    C *y = Y + (long)f * m;
    const int k = b % Ns;
    const int base = (b / Ns) * Ns * R + k;
    C v[16];
    for (int t = 0; t < R; t++) {
      C xv = x[b + t * nb];
      if (t > 0 && k > 0) {
        C w = tw[t * k * tstride];  // exp(-2 pi i j/m); the inverse uses its conjugate
        if (sign > 0) w.y = -w.y;
        xv = cmul(xv, w);
      }
      v[t] = xv;
    }
    if (R == 2) {
      y[base] = cadd(v[0], v[1]);
      y[base + Ns] = csub(v[0], v[1]);
    } else if (R == 4) {
      bfly4(v, sign);
      for (int t = 0; t < 4; t++) y[base + t * Ns] = v[t];
    } else if (R == 3) {
      const T c3 = (T)0.86602540378443864676;  // sqrt(3)/2
      const C t1 = cadd(v[1], v[2]);
      const C t2 = mkc(v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y);
      const C d = csub(v[1], v[2]);
      const C t3 = mkc(-sign * c3 * d.y, sign * c3 * d.x);
      y[base] = cadd(v[0], t1);
      y[base + Ns] = cadd(t2, t3);
      y[base + 2 * Ns] = csub(t2, t3);
    } else {
      const int st = m / R;
      for (int s = 0; s < R; s++) {
        C acc = v[0];
        for (int t = 1; t < R; t++) {
          const C w0 = tw[((t * s) % R) * st];
          const C w = mkc(w0.x, sign > 0 ? -w0.y : w0.y);
          acc = cadd(acc, cmul(v[t], w));
        }
        y[base + s * Ns] = acc;
      }
    }
  }
}

// complex input k of packed line f, from the real lines staged in Rw (n values
// each, rows rs apart)
template <typename R>
__device__ inline typename CT<R>::C pre_value(int kind, const R *__restrict__ Rw, int f, int k, int n,
                                              const typename CT<R>::C *__restrict__ q, int rs) {
  using C = typename CT<R>::C;
  C z;
  switch (kind) {
    case K_R2HC: z = mkc(Rw[(2 * f) * rs + k], Rw[(2 * f + 1) * rs + k]); break;
    case K_REDFT10:
    case K_RODFT10: {
      const int e = (k <= (n - 1) / 2) ? 2 * k : 2 * (n - 1 - k) + 1;
      const R sg = (kind == K_RODFT10 && (e & 1)) ? (R)-1.0 : (R)1.0;
      z = mkc(sg * Rw[(2 * f) * rs + e], sg * Rw[(2 * f + 1) * rs + e]);
      break;
    }
    case K_REDFT11:
    case K_RODFT11: {
      if (k < n) {
        const R sg = (kind == K_RODFT11 && (k & 1)) ? (R)-1.0 : (R)1.0;
        const R x = sg * Rw[f * rs + k];
        z = mkc(x * q[k].x, -x * q[k].y);
      } else {
        z = mkc((R)0.0, (R)0.0);
      }
      break;
    }
    case K_HC2R: {
      const R *Ha = Rw + (2 * f) * rs, *Hb = Rw + (2 * f + 1) * rs;
      R ra, ia, rb, ib;
      if (k == 0) { ra = Ha[0]; ia = 0.0; rb = Hb[0]; ib = 0.0; }
      else if (2 * k < n) { ra = Ha[k]; ia = Ha[n - k]; rb = Hb[k]; ib = Hb[n - k]; }
      else if (2 * k == n) { ra = Ha[k]; ia = 0.0; rb = Hb[k]; ib = 0.0; }
      else { ra = Ha[n - k]; ia = -Ha[k]; rb = Hb[n - k]; ib = -Hb[k]; }
      z = mkc(ra - ib, ia + rb);
      break;
    }
    case K_REDFT01:
    case K_RODFT01: {
      const R *Ha = Rw + (2 * f) * rs, *Hb = Rw + (2 * f + 1) * rs;
      R xa, xna, xb, xnb;
      if (kind == K_REDFT01) {
        xa = Ha[k]; xb = Hb[k];
        xna = (k == 0) ? (R)0.0 : Ha[n - k];
        xnb = (k == 0) ? (R)0.0 : Hb[n - k];
      } else {  // input reversed
        xa = Ha[n - 1 - k]; xb = Hb[n - 1 - k];
        xna = (k == 0) ? (R)0.0 : Ha[k - 1];
        xnb = (k == 0) ? (R)0.0 : Hb[k - 1];
      }
      const R cs = q[k].x, sn = q[k].y;
      const R rca = xa * cs + xna * sn, ica = xa * sn - xna * cs;
      const R rcb = xb * cs + xnb * sn, icb = xb * sn - xnb * cs;
      z = mkc(rca - icb, ica + rcb);
      break;
    }
    default: z = mkc((R)0.0, (R)0.0);
  }
  return z;
}

// real output k of line g from the transformed packed lines (length m each;
// Z(row, c) is complex c of row `row`)
template <typename R, class ZA>
__device__ inline R post_value(int kind, const ZA &Z, int g, int k, int n, int m,
                               const typename CT<R>::C *__restrict__ q, const typename CT<R>::C *__restrict__ t2) {
  using C = typename CT<R>::C;
  R val;
  switch (kind) {
    case K_R2HC: {
      const int kk = (2 * k <= n) ? k : n - k;
      const C zk = Z(g >> 1, kk), znk = Z(g >> 1, (m - kk) % m);
      R re, im;
      if ((g & 1) == 0) { re = (R)0.5 * (zk.x + znk.x); im = (R)0.5 * (zk.y - znk.y); }
      else { re = (R)0.5 * (zk.y + znk.y); im = (R)0.5 * (znk.x - zk.x); }
      val = (2 * k <= n) ? re : im;
      break;
    }
    case K_REDFT10:
    case K_RODFT10: {
      const int kk = (kind == K_REDFT10) ? k : n - 1 - k;
      const C zk = Z(g >> 1, kk), znk = Z(g >> 1, (m - kk) % m);
      R re, im;
      if ((g & 1) == 0) { re = (R)0.5 * (zk.x + znk.x); im = (R)0.5 * (zk.y - znk.y); }
      else { re = (R)0.5 * (zk.y + znk.y); im = (R)0.5 * (znk.x - zk.x); }
      val = (R)2.0 * (q[kk].x * re + q[kk].y * im);
      break;
    }
    case K_REDFT11:
    case K_RODFT11: {
      const int kk = (kind == K_REDFT11) ? k : n - 1 - k;
      const C zk = Z(g, kk);
      val = (R)2.0 * (t2[kk].x * zk.x + t2[kk].y * zk.y);
      break;
    }
    case K_HC2R: {
      const C zk = Z(g >> 1, k);
      val = (g & 1) ? zk.y : zk.x;
      break;
    }
    case K_REDFT01:
    case K_RODFT01: {
      const int srcm = (k & 1) ? n - 1 - (k >> 1) : (k >> 1);
      const C zk = Z(g >> 1, srcm);
      val = (g & 1) ? zk.y : zk.x;
      if (kind == K_RODFT01 && (k & 1)) val = -val;
      break;
    }
    default: val = 0.0;
  }
  return val;
}

// the stored value of an output element: scaled (in FP64 when the output is
// FP64), or added to what the output holds (accumulating corrections)
template <typename TO, typename R>
__device__ inline void store_out(TO *dst, long a, R val, double scale, int accum) {
  const double s = (double)val * scale;
  if (accum) dst[a] = (TO)((double)dst[a] + s);
  else dst[a] = (TO)s;
}

// Batched r2r line transform: G lines per block, all in LDS. R is the
// arithmetic type, TI/TO the element types of the input and output arrays.
template <typename R, typename TI, typename TO>
__global__ void __launch_bounds__(256) k_r2r(const TI *__restrict__ src, LineMap in, TO *__restrict__ dst,
                                             LineMap out, int L1, int n, int m, int kind, int G, double scale,
                                             int accum, const typename CT<R>::C *__restrict__ tw,
                                             const typename CT<R>::C *__restrict__ q,
                                             const typename CT<R>::C *__restrict__ t2, int nrad, RadList rl,
                                             Batch bt) {
  using C = typename CT<R>::C;
  if (bt.src) {
    src = (const TI *)bt.src[blockIdx.z];
    dst = (TO *)bt.dst[blockIdx.z];
  }
  extern __shared__ double2 smem_raw[];
  C *smem = (C *)smem_raw;
  const bool quarter = kind_is_quarter(kind);
  const int nfft = quarter ? G : G / 2;
  C *A = smem;
  C *B = smem + (long)nfft * m;
  R *Rw = (R *)B;  // raw staging, G*n values
  const int l1_0 = blockIdx.x * G;
  const long l2 = blockIdx.y;
  const int gcount = min(G, L1 - l1_0);
  const int tot = G * n;
  // 1. load the G lines
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    int g, e;
    if (in.se == 1) { g = idx / n; e = idx - g * n; }
    else { e = idx / G; g = idx - e * G; }
    Rw[g * n + e] = (g < gcount) ? (R)src[in.base + (long)(l1_0 + g) * in.s1 + l2 * in.s2 + (long)e * in.se] : (R)0.0;
  }
  __syncthreads();
  // 2. pre-processing into complex lines
  const int totc = nfft * m;
  for (int idx = threadIdx.x; idx < totc; idx += blockDim.x) {
    const int f = idx / m, k = idx - f * m;
    A[idx] = pre_value<R>(kind, Rw, f, k, n, q, n);
  }
  __syncthreads();
  // 3. complex FFT passes
  const R sign = kind_is_forward(kind) ? (R)-1.0 : (R)1.0;
  C *cur = A, *oth = B;
  int Ns = 1;
  for (int s = 0; s < nrad; s++) {
    const int Rd = rl.r[s];
    stockham_pass<R>(cur, oth, nfft, m, Rd, Ns, sign, tw);
    __syncthreads();
    C *t = cur; cur = oth; oth = t;
    Ns *= Rd;
  }
  // 4. post-processing into real lines
  R *O = (R *)oth;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int g = idx / n, k = idx - g * n;
    O[idx] = post_value<R>(kind, [&](int row, int c) { return cur[(long)row * m + c]; }, g, k, n, m, q, t2);
  }
  __syncthreads();
  // 5. store
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    int g, e;
    if (out.se == 1) { g = idx / n; e = idx - g * n; }
    else { e = idx / G; g = idx - e * G; }
    if (g < gcount) store_out(dst, out.base + (long)(l1_0 + g) * out.s1 + l2 * out.s2 + (long)e * out.se, O[g * n + e],
                              scale, accum);
  }
}

// Power-of-two line transforms (complex length M = 2^p, 16 <= M <= 2048).
// The same Stockham factorisation as k_r2r (radix 4 passes, then one radix 2
// pass when p is odd) with identical butterfly and twiddle arithmetic, so the
// results are bitwise those of k_r2r; here each thread keeps its four points
// in registers and the passes exchange through one LDS buffer, with all
// index arithmetic resolved at compile time. G real lines per block, M/4
// threads per complex line.
// (R: the arithmetic type, which sets the LDS bytes. FP32 blocks keep the
// FP64 line counts: twice the lines per block, at 1024 threads, measured
// slower for both the x and the y transforms.)
template <int M, bool QUARTER, typename R = double, int GDIV = 1>
struct P2Shape {
  static constexpr int LG = __builtin_ctz(M);
  static constexpr int NP4 = LG / 2;           // radix-4 passes
  static constexpr bool HAS2 = (LG & 1) != 0;  // closing radix-2 pass
  static constexpr int N = QUARTER ? M / 2 : M;  // real line length
  static constexpr int G0 = QUARTER ? (2048 / M < 64 ? 2048 / M : 64) : (4096 / M < 64 ? (4096 / M > 2 ? 4096 / M : 2) : 64);
  // strided lines (GDIV 1, the y lines) whose row segments of G0 values would
  // span less than 128 bytes take twice the lines (up to 1024 threads): 512
  // FP64 y lines 702 and 688 -> 675 and 656 us, FP32 ones 608 and 620 -> 530
  // and 529 us, FP32 256 lines 115 -> 102 us backward; 256 FP64 lines, whose
  // 16-line segments are 128 bytes already, measured slower doubled
  static constexpr int G64 =
      GDIV == 1 && !QUARTER && G0 * (int)sizeof(R) < 128 && G0 * M / 4 <= 1024 ? 2 * G0 : G0;
  // lines whose elements are contiguous in the source and the destination
  // (x lines) run GDIV times fewer lines per block: more, smaller blocks
  static constexpr int G = G64 / GDIV >= (QUARTER ? 1 : 2) ? G64 / GDIV : (QUARTER ? 1 : 2);
  static constexpr int NFFT = QUARTER ? G : G / 2;
  static constexpr int NT = NFFT * (M / 4);
  // LDS rows, padded against bank conflicts: one complex every 128 bytes
  // (SH: log2 of the complex values per 128 bytes). The staged real lines of
  // a strided source are written and read G lanes to a row (lane l: line
  // l % G, element l / G); the LDS serves 256 bytes (64 banks of 4 bytes) a
  // pass, SL values of type R, so a row stride RS = (SL/G) x (an odd number)
  // puts the lanes of one pass on distinct banks
  static constexpr int SH = sizeof(typename CT<R>::C) == 16 ? 3 : 4;
  static constexpr int MP = M + (M >> SH);  // complex row stride
  static constexpr int SL = 256 / (int)sizeof(R);
  static constexpr int RSM = SL / G > 1 ? SL / G : 1;
  static constexpr int RSJ = (N + RSM - 1) / RSM;
  static constexpr int RS = RSM * (RSJ | 1);  // real row stride
  static constexpr size_t LDSC = (size_t)NFFT * MP * sizeof(typename CT<R>::C);
  static constexpr size_t LDSR = (size_t)G * RS * sizeof(R);
  static constexpr size_t LDS = LDSC > LDSR ? LDSC : LDSR;
};
// padded position of complex c in an LDS row
template <int SH>
__device__ __forceinline__ int lds_pad(int c) {
  return c + (c >> SH);
}

// (WPE: the waves per SIMD its registers must allow; 8 for array sources, a
// level the LDS allows too, 1 (no cap) for computed sources, whose loads
// need more registers: only the WPE 1 instances read fs. KIND: the r2r kind
// when fixed at compile time, else -1 and the kind argument)
template <int M, bool QUARTER, typename R, typename TI, typename TO, int WPE, int GDIV, int KIND = -1>
__global__ void __launch_bounds__((P2Shape<M, QUARTER, R, GDIV>::NT)) __attribute__((amdgpu_waves_per_eu(WPE)))
    k_r2r_p2(const TI *__restrict__ src, LineMap in, TO *__restrict__ dst,
                                                LineMap out, int L1, int kind_a, double scale, int accum,
                                                const typename CT<R>::C *__restrict__ tw,
                                                const typename CT<R>::C *__restrict__ q,
                                                const typename CT<R>::C *__restrict__ t2, FillSrc fs,
                                                Batch bt) {
  using S = P2Shape<M, QUARTER, R, GDIV>;
  using C = typename CT<R>::C;
  constexpr int N = S::N, G = S::G, NT = S::NT, Q4 = M / 4, SH = S::SH, MP = S::MP, RS = S::RS;
  const int kind = KIND >= 0 ? KIND : kind_a;  // (KIND: the kind fixed at compile time)
  if (bt.src) {  // case blockIdx.z of a batch of same-shaped solves
    src = (const TI *)bt.src[blockIdx.z];
    dst = (TO *)bt.dst[blockIdx.z];
  }
  extern __shared__ double2 smem_raw[];
  C *smem = (C *)smem_raw;
  R *Rw = (R *)smem;  // G staged real lines (RS apart), overlaid by the complex lines (MP apart)
  const int tid = threadIdx.x;
  const int l1_0 = blockIdx.x * G;
  const long l2 = blockIdx.y;
  const int gcount = min(G, L1 - l1_0);
  // 1. coalesced load of the G lines
  double vmax = 0.0;
#pragma unroll
  for (int r = 0; r < G * N / NT; r++) {
    const int idx = tid + r * NT;
    int g, e;
    if (in.se == 1) { g = idx / N; e = idx - g * N; }
    else { e = idx / G; g = idx - e * G; }
    R val = 0.0;
    if (g < gcount) {
      const long a = in.base + (long)(l1_0 + g) * in.s1 + l2 * in.s2 + (long)e * in.se;
      if (WPE == 1 && fs.on) {
        const double sv = src_value(fs, a, (int)l2 + 1);
        vmax = fmax(vmax, fabs(sv));
        val = (R)sv;
      } else {
        val = (R)src[a];
      }
    }
    Rw[g * RS + e] = val;
  }
  if (WPE == 1 && fs.rmax) slots_max(fs.rmax, vmax, (blockIdx.x + gridDim.x * blockIdx.y) * (NT / 64) + tid / 64);
  __syncthreads();
  const int f = tid / Q4, b = tid - f * Q4;
  C v[4];
#pragma unroll
  for (int t = 0; t < 4; t++) v[t] = pre_value<R>(kind, Rw, f, b + t * Q4, N, q, RS);
  __syncthreads();
  const R sign = kind_is_forward(kind) ? (R)-1.0 : (R)1.0;
  C *Z = smem + (long)f * MP;
  // 2. radix-4 passes
#pragma unroll
  for (int s = 0; s < S::NP4; s++) {
    const int Ns = 1 << (2 * s);
    if (s > 0) {
#pragma unroll
      for (int t = 0; t < 4; t++) v[t] = Z[lds_pad<SH>(b + t * Q4)];
      __syncthreads();
    }
    const int k = b & (Ns - 1);
    const int base = (b - k) * 4 + k;
    const int tstride = M / (Ns * 4);
    if (k > 0) {
#pragma unroll
      for (int t = 1; t < 4; t++) {
        C w = tw[t * k * tstride];
        if (sign > 0) w.y = -w.y;
        v[t] = cmul(v[t], w);
      }
    }
    bfly4(v, sign);
#pragma unroll
    for (int t = 0; t < 4; t++) Z[lds_pad<SH>(base + t * Ns)] = v[t];
    __syncthreads();
  }
  // 3. closing radix-2 pass (Ns = M/2): butterflies b and b + M/4
  if (S::HAS2) {
#pragma unroll
    for (int t = 0; t < 4; t++) v[t] = Z[lds_pad<SH>(b + t * Q4)];
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int k = b + h * Q4;
      C x0 = v[h], x1 = v[h + 2];
      if (k > 0) {
        C w = tw[k];
        if (sign > 0) w.y = -w.y;
        x1 = cmul(x1, w);
      }
      Z[lds_pad<SH>(k)] = cadd(x0, x1);
      Z[lds_pad<SH>(k + M / 2)] = csub(x0, x1);
    }
    __syncthreads();
  }
  // 4. post-processing and store
  constexpr int PER = G * N / NT;
  R val[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int idx = tid + r * NT;
    const int g = idx / N, k = idx - g * N;
    val[r] = post_value<R>(kind, [&](int row, int c) { return smem[row * MP + lds_pad<SH>(c)]; }, g, k, N, M, q,
                           t2);
  }
  if (out.se == 1) {
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int idx = tid + r * NT;
      const int g = idx / N, e = idx - g * N;
      if (g < gcount) store_out(dst, out.base + (long)(l1_0 + g) * out.s1 + l2 * out.s2 + e, val[r], scale, accum);
    }
    return;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int idx = tid + r * NT;
    const int g = idx / N, k = idx - g * N;
    Rw[g * RS + k] = val[r];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int idx = tid + r * NT;
    const int e = idx / G, g = idx - e * G;
    if (g < gcount)
      store_out(dst, out.base + (long)(l1_0 + g) * out.s1 + l2 * out.s2 + (long)e * out.se, Rw[g * RS + e], scale,
                accum);
  }
}

// the transform tables in the arithmetic type R
template <typename R>
struct PlanTabs {
  const typename CT<R>::C *tw, *q, *t2;
};
static PlanTabs<double> tabs_of(const fz_fftplan &P, double) { return {P.tw, P.q, P.t2}; }
static PlanTabs<float> tabs_of(const fz_fftplan &P, float) { return {P.tw32, P.q32, P.t232}; }

template <int M, bool QUARTER, typename R, typename TI, typename TO, int GDIV>
static void launch_p2_g(fz_ctx *ctx, const char *name, const fz_fftplan &P, int kind, const TI *src, LineMap in,
                        TO *dst, LineMap out, int L1, int L2, double scale, int accum, const FillSrc &fs,
                        const Batch &bt) {
  using S = P2Shape<M, QUARTER, R, GDIV>;
  dim3 grid((L1 + S::G - 1) / S::G, L2, bt.src ? bt.n : 1);
  const PlanTabs<R> t = tabs_of(P, R());
#define FZ_P2_LAUNCH(WPE, KIND)                                                                                   \
  FZ_LAUNCH(ctx, name, (k_r2r_p2<M, QUARTER, R, TI, TO, WPE, GDIV, KIND>), grid, dim3(S::NT), S::LDS, src, in, dst, \
            out, L1, kind, scale, accum, t.tw, t.q, t.t2, fs, bt)
  // the periodic kinds of lengths from 64 with the kind fixed at compile
  // time: the pre- and post-processing without the other kinds' branches
  // (y lines of 256: 174 and 163 -> 152 and 141 us forward and backward)
  constexpr bool SPEC = !QUARTER && M >= 64;
  if (fs.on) {
    if (SPEC && kind == K_R2HC) FZ_P2_LAUNCH(1, SPEC ? K_R2HC : -1);
    else FZ_P2_LAUNCH(1, -1);
  } else if (SPEC && kind == K_R2HC) {
    FZ_P2_LAUNCH(8, SPEC ? K_R2HC : -1);
  } else if (SPEC && kind == K_HC2R) {
    FZ_P2_LAUNCH(8, SPEC ? K_HC2R : -1);
  } else {
    FZ_P2_LAUNCH(8, -1);
  }
#undef FZ_P2_LAUNCH
}
template <int M, bool QUARTER, typename R, typename TI, typename TO>
static void launch_p2(fz_ctx *ctx, const char *name, const fz_fftplan &P, int kind, const TI *src, LineMap in,
                      TO *dst, LineMap out, int L1, int L2, double scale, int accum, const FillSrc &fs,
                      const Batch &bt) {
  // x lines (contiguous on both sides): half the lines per block (x and y
  // lines of 256 at 16 lines per block: 242 and 168 us; x lines at 8 per
  // block 223 us, at 4 per block 228 us; y lines at 8 or 32 per block 177
  // and 175 us)
  if (in.se == 1 && out.se == 1 && ctx->p2_xdiv == 2)
    launch_p2_g<M, QUARTER, R, TI, TO, 2>(ctx, name, P, kind, src, in, dst, out, L1, L2, scale, accum, fs, bt);
  else
    launch_p2_g<M, QUARTER, R, TI, TO, 1>(ctx, name, P, kind, src, in, dst, out, L1, L2, scale, accum, fs, bt);
}

static const Batch kNoBatch = {nullptr, nullptr, 0};

static bool p2_length(int m) { return m >= 16 && m <= 2048 && (m & (m - 1)) == 0; }

template <bool QUARTER, typename R, typename TI, typename TO>
static bool dispatch_p2(fz_ctx *ctx, const char *name, const fz_fftplan &P, int kind, const TI *src, LineMap in,
                        TO *dst, LineMap out, int L1, int L2, double scale, int accum, const FillSrc &fs,
                        const Batch &bt = kNoBatch) {
#define FZ_P2_CASE(MM) \
  case MM: launch_p2<MM, QUARTER, R, TI, TO>(ctx, name, P, kind, src, in, dst, out, L1, L2, scale, accum, fs, bt); return true;
  switch (P.m) {
    FZ_P2_CASE(16)
    FZ_P2_CASE(32)
    FZ_P2_CASE(64)
    FZ_P2_CASE(128)
    FZ_P2_CASE(256)
    FZ_P2_CASE(512)
    FZ_P2_CASE(1024)
    FZ_P2_CASE(2048)
    default: return false;
  }
#undef FZ_P2_CASE
}

// Thomas recurrence of dgtsv_homebrewed (src/solver.f90:151-177) on one
// column of the spectral buffer; pivots go to piv (same layout, FP64). The
// recurrence runs in FP64 whatever the column's element type.
template <typename T>
__device__ inline void thomas_col(T *__restrict__ col, double *__restrict__ dcol, long ps, int n,
                                  const double *__restrict__ a, const double *__restrict__ b,
                                  const double *__restrict__ c, double lam) {
  double bb = b[0] + lam;
  double z = 1.0 / (bb + FZ_EPS);
  double d = c[0] * z;
  double pprev = (double)col[0] * z;
  dcol[0] = d;
  col[0] = (T)pprev;
  for (int l = 1; l < n; l++) {
    bb = b[l] + lam;
    z = 1.0 / (bb - a[l] * d + FZ_EPS);
    d = c[l] * z;
    pprev = ((double)col[l * ps] - a[l] * pprev) * z;
    dcol[l * ps] = d;
    col[l * ps] = (T)pprev;
  }
  double pnext = pprev;
  for (int l = n - 2; l >= 0; l--) {
    pnext = (double)col[l * ps] - dcol[l * ps] * pnext;
    col[l * ps] = (T)pnext;
  }
}

// The chunked recurrence of one column over n rows (pin: the last unknown
// set to zero, for the gauge column of the mixed-precision solve)
template <int TK, typename T>
__device__ __forceinline__ void thomas_chunked(T *__restrict__ col, double *__restrict__ dcol, long ps, int n3,
                                               double lam, const double *__restrict__ a,
                                               const double *__restrict__ b, const double *__restrict__ c,
                                               bool pin) {
  double d = 0.0, pprev = 0.0;
  for (int l0 = 0; l0 < n3; l0 += TK) {
    double r[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < n3) r[t] = (double)col[(long)(l0 + t) * ps];
#pragma unroll
    for (int t = 0; t < TK; t++) {
      const int l = l0 + t;
      if (l < n3) {
        const double bb = b[l] + lam;
        double z;
        if (l == 0) {
          z = 1.0 / (bb + FZ_EPS);
          d = c[0] * z;
          pprev = r[t] * z;
        } else {
          z = 1.0 / (bb - a[l] * d + FZ_EPS);
          d = c[l] * z;
          pprev = (r[t] - a[l] * pprev) * z;
        }
        if (pin && l == n3 - 1) pprev = 0.0;
        dcol[(long)l * ps] = d;
        col[(long)l * ps] = (T)pprev;
      }
    }
  }
  double pnext = pprev;
  for (int h0 = n3 - 2; h0 >= 0; h0 -= TK) {
    double r[TK], dd[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        r[t] = (double)col[(long)(h0 - t) * ps];
        dd[t] = dcol[(long)(h0 - t) * ps];
      }
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        pnext = r[t] - dd[t] * pnext;
        col[(long)(h0 - t) * ps] = (T)pnext;
      }
  }
}

// The same recurrence as thomas_col, in chunks of TK rows: each chunk's
// loads are issued together so one wave keeps TK loads in flight (one
// column per lane leaves few waves per CU to hide latency otherwise).
// Column `gauge` (the singular mode of an all-Neumann/periodic operator, or
// -1) has its last unknown set to zero instead of divided by the eps-shifted
// zero pivot: the mixed-precision solve fixes the null mode this way, since
// the O(1/eps) constant of the reference's pivots would swamp FP32 values.
template <int TK, typename T>
__global__ void __launch_bounds__(64) k_thomas(T *__restrict__ W, double *__restrict__ D,
                                               const double *__restrict__ lambdaxy, const double *__restrict__ a,
                                               const double *__restrict__ b, const double *__restrict__ c, int n1,
                                               int n2, int n3, long gauge, Batch bt) {
  if (bt.src) {  // case blockIdx.y of a batch: its spectral array and pivots
    W = (T *)bt.src[blockIdx.y];
    D = (double *)bt.dst[blockIdx.y];
  }
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n1 * n2) return;
  const long ps = (long)n1 * n2;
  thomas_chunked<TK, T>(W + idx, D + idx, ps, n3, lambdaxy[idx], a, b, c, idx == gauge);
}

// k_thomas without a stored pivot array. A first sweep runs the forward
// elimination and keeps only its state (the last pivot d and eliminated
// value) at the end of every chunk of TK rows, in CK (two FP64 values per
// chunk and column). The backward sweep takes the chunks from the last:
// it re-runs a chunk's elimination from the saved state, with the same
// operations and so the same values, holds that chunk's pivots and
// eliminated values in registers and substitutes back. A column is read
// twice and written once (about 25 bytes per FP64 cell), against three
// reads and three writes (48 bytes) with a pivot array. The values are
// k_thomas<TK, T>'s bit for bit: the eliminated values are rounded to T
// where k_thomas stores and reloads them, and the last row keeps its FP64
// value as the start of the substitution, as there.
template <int TK, typename T>
__global__ void __launch_bounds__(64) k_thomas_rc(T *__restrict__ W, double *__restrict__ CK,
                                                  const double *__restrict__ lambdaxy, const double *__restrict__ a,
                                                  const double *__restrict__ b, const double *__restrict__ c, int n1,
                                                  int n2, int n3, long gauge, Batch bt) {
  if (bt.src) {  // case blockIdx.y of a batch: its spectral array and checkpoint area
    W = (T *)bt.src[blockIdx.y];
    CK = (double *)bt.dst[blockIdx.y];
  }
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n1 * n2) return;
  const long ps = (long)n1 * n2;
  T *col = W + idx;
  double *ck = CK + idx;
  const double lam = lambdaxy[idx];
  const bool pin = idx == gauge;
  const int nch = (n3 + TK - 1) / TK;
  double d = 0.0, pprev = 0.0;
  for (int ch = 0; ch < nch; ch++) {
    const int l0 = ch * TK;
    double r[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < n3) r[t] = (double)col[(long)(l0 + t) * ps];
#pragma unroll
    for (int t = 0; t < TK; t++) {
      const int l = l0 + t;
      if (l < n3) {
        const double bb = b[l] + lam;
        double z;
        if (l == 0) {
          z = 1.0 / (bb + FZ_EPS);
          d = c[0] * z;
          pprev = r[t] * z;
        } else {
          z = 1.0 / (bb - a[l] * d + FZ_EPS);
          d = c[l] * z;
          pprev = (r[t] - a[l] * pprev) * z;
        }
        if (pin && l == n3 - 1) pprev = 0.0;
      }
    }
    if (ch + 1 < nch) {
      ck[(long)(2 * ch) * ps] = d;
      ck[(long)(2 * ch + 1) * ps] = pprev;
    }
  }
  double pnext = 0.0;
  for (int ch = nch - 1; ch >= 0; ch--) {
    const int l0 = ch * TK;
    double dd = 0.0, pp = 0.0;
    if (ch > 0) {
      dd = ck[(long)(2 * ch - 2) * ps];
      pp = ck[(long)(2 * ch - 1) * ps];
    }
    double r[TK], dv[TK], pv[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < n3) r[t] = (double)col[(long)(l0 + t) * ps];
#pragma unroll
    for (int t = 0; t < TK; t++) {
      const int l = l0 + t;
      if (l < n3) {
        const double bb = b[l] + lam;
        double z;
        if (l == 0) {
          z = 1.0 / (bb + FZ_EPS);
          dd = c[0] * z;
          pp = r[t] * z;
        } else {
          z = 1.0 / (bb - a[l] * dd + FZ_EPS);
          dd = c[l] * z;
          pp = (r[t] - a[l] * pp) * z;
        }
        if (pin && l == n3 - 1) pp = 0.0;
        dv[t] = dd;
        pv[t] = l == n3 - 1 ? pp : (double)(T)pp;
      }
    }
#pragma unroll
    for (int t = TK - 1; t >= 0; t--) {
      const int l = l0 + t;
      if (l < n3) {
        pnext = l == n3 - 1 ? pv[t] : pv[t] - dv[t] * pnext;
        col[(long)l * ps] = (T)pnext;
      }
    }
  }
}

// The pivots of dgtsv_homebrewed (src/solver.f90:151-177) depend on the
// operator only: k_piv_pre makes them once, with thomas_chunked's operations
// (Z: the reciprocal pivot z of every row, D: d = c z), and k_thomas_pre
// solves with them. The values are k_thomas's bit for bit; the bytes moved
// are the same (Z read in the forward sweep where k_thomas writes D), but no
// division lies on the per-row dependency chain.
__global__ void k_piv_pre(double *__restrict__ Z, double *__restrict__ D, const double *__restrict__ lambdaxy,
                          const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
                          int n1, int n2, int n3) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ps = (long)n1 * n2;
  if (idx >= ps) return;
  const double lam = lambdaxy[idx];
  double d = 0.0;
  for (int l = 0; l < n3; l++) {
    const double bb = b[l] + lam;
    const double z = l == 0 ? 1.0 / (bb + FZ_EPS) : 1.0 / (bb - a[l] * d + FZ_EPS);
    d = (l == 0 ? c[0] : c[l]) * z;
    Z[idx + (long)l * ps] = z;
    D[idx + (long)l * ps] = d;
  }
}

template <int TK, typename T>
__global__ void __launch_bounds__(64) k_thomas_pre(T *__restrict__ W, const double *__restrict__ Z,
                                                   const double *__restrict__ D, const double *__restrict__ a,
                                                   int n1, int n2, int n3, long gauge) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ps = (long)n1 * n2;
  if (idx >= ps) return;
  T *col = W + idx;
  const double *zc = Z + idx, *dc = D + idx;
  const bool pin = idx == gauge;
  double pprev = 0.0;
  for (int l0 = 0; l0 < n3; l0 += TK) {
    double r[TK], z[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < n3) {
        r[t] = (double)col[(long)(l0 + t) * ps];
        z[t] = zc[(long)(l0 + t) * ps];
      }
#pragma unroll
    for (int t = 0; t < TK; t++) {
      const int l = l0 + t;
      if (l < n3) {
        pprev = l == 0 ? r[t] * z[t] : (r[t] - a[l] * pprev) * z[t];
        if (pin && l == n3 - 1) pprev = 0.0;
        col[(long)l * ps] = (T)pprev;
      }
    }
  }
  double pnext = pprev;
  for (int h0 = n3 - 2; h0 >= 0; h0 -= TK) {
    double r[TK], dd[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        r[t] = (double)col[(long)(h0 - t) * ps];
        dd[t] = dc[(long)(h0 - t) * ps];
      }
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        pnext = r[t] - dd[t] * pnext;
        col[(long)(h0 - t) * ps] = (T)pnext;
      }
  }
}

// k_thomas for FP32 columns, two adjacent columns per lane: each load moves
// 8 bytes of W and 16 of pivots per lane, as the FP64 kernel does with one
// column. The per-column arithmetic is k_thomas<TK, float>'s exactly.
// Needs an even column count n1*n2.
template <int TK>
__global__ void __launch_bounds__(64) k_thomas_x2(float *__restrict__ W, double *__restrict__ D,
                                                  const double *__restrict__ lambdaxy, const double *__restrict__ a,
                                                  const double *__restrict__ b, const double *__restrict__ c, int n1,
                                                  int n2, int n3, long gauge) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;  // column pair
  const long ps2 = (long)n1 * n2 / 2;
  if (idx >= ps2) return;
  float2 *col = (float2 *)W + idx;
  double2 *dcol = (double2 *)D + idx;
  const double lam0 = lambdaxy[2 * idx], lam1 = lambdaxy[2 * idx + 1];
  const bool g0 = 2 * idx == gauge, g1 = 2 * idx + 1 == gauge;
  double d0 = 0.0, d1 = 0.0, p0 = 0.0, p1 = 0.0;
  for (int l0 = 0; l0 < n3; l0 += TK) {
    float2 r[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < n3) r[t] = col[(long)(l0 + t) * ps2];
#pragma unroll
    for (int t = 0; t < TK; t++) {
      const int l = l0 + t;
      if (l < n3) {
        const double bb0 = b[l] + lam0, bb1 = b[l] + lam1;
        double z0, z1;
        if (l == 0) {
          z0 = 1.0 / (bb0 + FZ_EPS);
          z1 = 1.0 / (bb1 + FZ_EPS);
          d0 = c[0] * z0;
          d1 = c[0] * z1;
          p0 = (double)r[t].x * z0;
          p1 = (double)r[t].y * z1;
        } else {
          z0 = 1.0 / (bb0 - a[l] * d0 + FZ_EPS);
          z1 = 1.0 / (bb1 - a[l] * d1 + FZ_EPS);
          d0 = c[l] * z0;
          d1 = c[l] * z1;
          p0 = ((double)r[t].x - a[l] * p0) * z0;
          p1 = ((double)r[t].y - a[l] * p1) * z1;
        }
        if (l == n3 - 1) {
          if (g0) p0 = 0.0;
          if (g1) p1 = 0.0;
        }
        dcol[(long)l * ps2] = make_double2(d0, d1);
        col[(long)l * ps2] = make_float2((float)p0, (float)p1);
      }
    }
  }
  double q0 = p0, q1 = p1;
  for (int h0 = n3 - 2; h0 >= 0; h0 -= TK) {
    float2 r[TK];
    double2 dd[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        r[t] = col[(long)(h0 - t) * ps2];
        dd[t] = dcol[(long)(h0 - t) * ps2];
      }
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        q0 = (double)r[t].x - dd[t].x * q0;
        q1 = (double)r[t].y - dd[t].y * q1;
        col[(long)(h0 - t) * ps2] = make_float2((float)q0, (float)q1);
      }
  }
}

// k_thomas_pre for FP32 columns, two adjacent columns per lane (an even
// column count), with the pivots made once at set-up (k_piv_pre): each row
// moves 8 bytes of W and 16 of pivots per lane, as k_thomas_x2 does, without
// the division on the per-row dependency chain. The per-column arithmetic is
// k_thomas_pre<TK, float>'s, and so k_thomas<TK, float>'s, exactly.
template <int TK>
__global__ void __launch_bounds__(64) k_thomas_pre_x2(float *__restrict__ W, const double *__restrict__ Z,
                                                      const double *__restrict__ D, const double *__restrict__ a,
                                                      int n1, int n2, int n3, long gauge) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;  // column pair
  const long ps2 = (long)n1 * n2 / 2;
  if (idx >= ps2) return;
  float2 *col = (float2 *)W + idx;
  const double2 *zc = (const double2 *)Z + idx, *dc = (const double2 *)D + idx;
  const bool g0 = 2 * idx == gauge, g1 = 2 * idx + 1 == gauge;
  double p0 = 0.0, p1 = 0.0;
  for (int l0 = 0; l0 < n3; l0 += TK) {
    float2 r[TK];
    double2 z[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < n3) {
        r[t] = col[(long)(l0 + t) * ps2];
        z[t] = zc[(long)(l0 + t) * ps2];
      }
#pragma unroll
    for (int t = 0; t < TK; t++) {
      const int l = l0 + t;
      if (l < n3) {
        p0 = l == 0 ? (double)r[t].x * z[t].x : ((double)r[t].x - a[l] * p0) * z[t].x;
        p1 = l == 0 ? (double)r[t].y * z[t].y : ((double)r[t].y - a[l] * p1) * z[t].y;
        if (l == n3 - 1) {
          if (g0) p0 = 0.0;
          if (g1) p1 = 0.0;
        }
        col[(long)l * ps2] = make_float2((float)p0, (float)p1);
      }
    }
  }
  double q0 = p0, q1 = p1;
  for (int h0 = n3 - 2; h0 >= 0; h0 -= TK) {
    float2 r[TK];
    double2 dd[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        r[t] = col[(long)(h0 - t) * ps2];
        dd[t] = dc[(long)(h0 - t) * ps2];
      }
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        q0 = (double)r[t].x - dd[t].x * q0;
        q1 = (double)r[t].y - dd[t].y * q1;
        col[(long)(h0 - t) * ps2] = make_float2((float)q0, (float)q1);
      }
  }
}

// the second solve of gaussel_periodic (src/solver.f90:107-149): it
// depends on the operator only, so it is made once per solver set-up, with
// the operations of the per-solve form (pivots scratch in D)
__global__ void k_periodic_p2(double *__restrict__ D, double *__restrict__ P2, const double *__restrict__ lambdaxy,
                              const double *__restrict__ a, const double *__restrict__ b,
                              const double *__restrict__ c, int n1, int n2, int n3) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n1 * n2) return;
  const long ps = (long)n1 * n2;
  const int m = n3 - 1;
  double *p2 = P2 + idx;
  for (int l = 0; l < m; l++) p2[l * ps] = 0.0;
  p2[0] = -a[0];
  p2[(m - 1) * ps] = -c[m - 1];
  thomas_col(p2, D + idx, ps, m, a, b, c, lambdaxy[idx]);
}

// gaussel_periodic with the second solve P2 made at set-up: the chunked
// recurrence on the first n3-1 rows, then the cyclic unknown and the update
// (the gauge column as in k_thomas: the cyclic unknown is set to zero)
template <int TK, typename T>
__global__ void __launch_bounds__(64) k_thomas_periodic_pc(T *__restrict__ W, double *__restrict__ D,
                                                           const double *__restrict__ P2,
                                                           const double *__restrict__ lambdaxy,
                                                           const double *__restrict__ a, const double *__restrict__ b,
                                                           const double *__restrict__ c, int n1, int n2, int n3,
                                                           long gauge) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n1 * n2) return;
  const long ps = (long)n1 * n2;
  const double lam = lambdaxy[idx];
  const int m = n3 - 1;
  T *col = W + idx;
  const double *p2 = P2 + idx;
  thomas_chunked<TK, T>(col, D + idx, ps, m, lam, a, b, c, false);
  const double bbm = b[m] + lam;
  double pn = ((double)col[m * ps] - c[m] * (double)col[0] - a[m] * (double)col[(m - 1) * ps]) /
              (bbm + c[m] * p2[0] + a[m] * p2[(m - 1) * ps] + FZ_EPS);
  if (idx == gauge) pn = 0.0;
  col[m * ps] = (T)pn;
  for (int l0 = 0; l0 < m; l0 += TK) {
    double r[TK], q[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < m) {
        r[t] = (double)col[(long)(l0 + t) * ps];
        q[t] = p2[(long)(l0 + t) * ps];
      }
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < m) col[(long)(l0 + t) * ps] = (T)(r[t] + q[t] * pn);
  }
}

// the source of a forward x transform evaluated into x lines of W (the
// general path for lengths k_r2r_p2 does not cover)
template <typename T>
__global__ void k_src_lines(fz_dims d, FillSrc fs, T *__restrict__ W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  double v = 0.0;
  if (i <= d.n1 && j <= d.n2) {
    v = src_value(fs, i + d.sy * j + d.sz * k, k);
    W[(i - 1) + (long)d.n1 * ((j - 1) + (long)d.n2 * (k - 1))] = (T)v;
  }
  if (fs.rmax) {
    const unsigned t = threadIdx.x + blockDim.x * threadIdx.y;
    const unsigned blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    double m = fabs(v);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((t & 63) == 0)
      atomicMax(fs.rmax + ((blk * 4 + t / 64) & (FZ_GAMMA_SLOTS - 1)), (unsigned long long)__double_as_longlong(m));
  }
}

// k_thomas_periodic_pc for FP32 columns, two adjacent columns per lane (the
// per-column arithmetic of k_thomas_periodic_pc<TK, float>; even n1*n2)
template <int TK>
__global__ void __launch_bounds__(64) k_thomas_periodic_x2(float *__restrict__ W, double *__restrict__ D,
                                                           const double *__restrict__ P2,
                                                           const double *__restrict__ lambdaxy,
                                                           const double *__restrict__ a, const double *__restrict__ b,
                                                           const double *__restrict__ c, int n1, int n2, int n3,
                                                           long gauge) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;  // column pair
  const long ps2 = (long)n1 * n2 / 2;
  if (idx >= ps2) return;
  float2 *col = (float2 *)W + idx;
  double2 *dcol = (double2 *)D + idx;
  const double2 *p2 = (const double2 *)P2 + idx;
  const double lam0 = lambdaxy[2 * idx], lam1 = lambdaxy[2 * idx + 1];
  const int m = n3 - 1;
  double d0 = 0.0, d1 = 0.0, p0 = 0.0, p1 = 0.0;
  for (int l0 = 0; l0 < m; l0 += TK) {
    float2 r[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < m) r[t] = col[(long)(l0 + t) * ps2];
#pragma unroll
    for (int t = 0; t < TK; t++) {
      const int l = l0 + t;
      if (l < m) {
        const double bb0 = b[l] + lam0, bb1 = b[l] + lam1;
        double z0, z1;
        if (l == 0) {
          z0 = 1.0 / (bb0 + FZ_EPS);
          z1 = 1.0 / (bb1 + FZ_EPS);
          d0 = c[0] * z0;
          d1 = c[0] * z1;
          p0 = (double)r[t].x * z0;
          p1 = (double)r[t].y * z1;
        } else {
          z0 = 1.0 / (bb0 - a[l] * d0 + FZ_EPS);
          z1 = 1.0 / (bb1 - a[l] * d1 + FZ_EPS);
          d0 = c[l] * z0;
          d1 = c[l] * z1;
          p0 = ((double)r[t].x - a[l] * p0) * z0;
          p1 = ((double)r[t].y - a[l] * p1) * z1;
        }
        dcol[(long)l * ps2] = make_double2(d0, d1);
        col[(long)l * ps2] = make_float2((float)p0, (float)p1);
      }
    }
  }
  double q0 = p0, q1 = p1;
  for (int h0 = m - 2; h0 >= 0; h0 -= TK) {
    float2 r[TK];
    double2 dd[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        r[t] = col[(long)(h0 - t) * ps2];
        dd[t] = dcol[(long)(h0 - t) * ps2];
      }
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (h0 - t >= 0) {
        q0 = (double)r[t].x - dd[t].x * q0;
        q1 = (double)r[t].y - dd[t].y * q1;
        col[(long)(h0 - t) * ps2] = make_float2((float)q0, (float)q1);
      }
  }
  const float2 cm = col[(long)m * ps2], c0 = col[0], cm1 = col[(long)(m - 1) * ps2];
  const double2 p2f = p2[0], p2l = p2[(long)(m - 1) * ps2];
  double pn0 = ((double)cm.x - c[m] * (double)c0.x - a[m] * (double)cm1.x) /
               (b[m] + lam0 + c[m] * p2f.x + a[m] * p2l.x + FZ_EPS);
  double pn1 = ((double)cm.y - c[m] * (double)c0.y - a[m] * (double)cm1.y) /
               (b[m] + lam1 + c[m] * p2f.y + a[m] * p2l.y + FZ_EPS);
  if (2 * idx == gauge) pn0 = 0.0;
  if (2 * idx + 1 == gauge) pn1 = 0.0;
  col[(long)m * ps2] = make_float2((float)pn0, (float)pn1);
  for (int l0 = 0; l0 < m; l0 += TK) {
    float2 r[TK];
    double2 q[TK];
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < m) {
        r[t] = col[(long)(l0 + t) * ps2];
        q[t] = p2[(long)(l0 + t) * ps2];
      }
#pragma unroll
    for (int t = 0; t < TK; t++)
      if (l0 + t < m)
        col[(long)(l0 + t) * ps2] = make_float2((float)((double)r[t].x + q[t].x * pn0),
                                                (float)((double)r[t].y + q[t].y * pn1));
  }
}

// ---------------------------------------------------------------------------
// host side: operator set-up
// ---------------------------------------------------------------------------
static void eigenvalues_c(int n, char b0, char b1, double *lambda) {
  for (int l = 1; l <= n; l++) {  // src/initsolver.f90:89-138, cell-centred
    double arg;
    if (b0 == 'P' && b1 == 'P') arg = ((double)(2 * (l - 1)) * FZ_PI) / (double)n;
    else if (b0 == 'N' && b1 == 'N') arg = ((double)(l - 1) * FZ_PI) / (double)n;
    else if (b0 == 'D' && b1 == 'D') arg = ((double)l * FZ_PI) / (double)n;
    else arg = ((double)(2 * l - 1) * FZ_PI) / (2.0 * n);
    lambda[l - 1] = -2.0 * (1.0 - cos(arg));
  }
}

static void find_fft_c(char b0, char b1, int *kf, int *kb, double *norm) {  // src/fft.f90:195-217
  if (b0 == 'P' && b1 == 'P') { *kf = K_R2HC; *kb = K_HC2R; *norm = 1.0; }
  else if (b0 == 'N' && b1 == 'N') { *kf = K_REDFT10; *kb = K_REDFT01; *norm = 2.0; }
  else if (b0 == 'D' && b1 == 'D') { *kf = K_RODFT10; *kb = K_RODFT01; *norm = 2.0; }
  else if (b0 == 'N' && b1 == 'D') { *kf = K_REDFT11; *kb = K_REDFT11; *norm = 2.0; }
  else { *kf = K_RODFT11; *kb = K_RODFT11; *norm = 2.0; }
}

static double bc_factor_c(char cbc, int ib, double bc, double dlc) {  // src/initsolver.f90:196-209
  if (cbc == 'P') return 0.0;
  if (cbc == 'D') return -2.0 * bc;
  const double sgn = (ib == 0) ? 1.0 : -1.0;
  return sgn * dlc * bc;
}

static int factorize(int m, int *rad, int *nrad) {
  int k = 0;
  while (m % 4 == 0) { rad[k++] = 4; m /= 4; }
  while (m % 2 == 0) { rad[k++] = 2; m /= 2; }
  for (int p = 3; m > 1; p += 2) {
    while (m % p == 0) {
      if (p > 13 || k >= 24) return -1;
      rad[k++] = p;
      m /= p;
    }
  }
  *nrad = k;
  return 0;
}

static int fftplan_init_kinds(fz_fftplan *P, int kind_f, int kind_b, int n) {
  P->kind_f = kind_f;
  P->kind_b = kind_b;
  P->n = n;
  P->m = kind_is_quarter(P->kind_f) ? 2 * n : n;
  if (P->m > 2048)
    return fz_set_error(FZ_ERR_UNSUPPORTED, "line transforms longer than 2048 points are not supported (n=%d)", n);
  if (factorize(P->m, P->rad, &P->nrad) != 0)
    return fz_set_error(FZ_ERR_UNSUPPORTED, "transform length %d has a prime factor above 13", P->m);
  std::vector<double2> tw(P->m), q(n + 1), t2(n + 1);
  for (int j = 0; j < P->m; j++) tw[j] = make_double2(fzm::cos(2.0 * FZ_PI * (double)j / (double)P->m),
                                                      -fzm::sin(2.0 * FZ_PI * (double)j / (double)P->m));
  for (int k = 0; k <= n; k++) {
    q[k] = make_double2(fzm::cos(FZ_PI * (double)k / (2.0 * n)), fzm::sin(FZ_PI * (double)k / (2.0 * n)));
    t2[k] = make_double2(fzm::cos(FZ_PI * (double)(2 * k + 1) / (4.0 * n)), fzm::sin(FZ_PI * (double)(2 * k + 1) / (4.0 * n)));
  }
  std::vector<float2> tw32(P->m), q32(n + 1), t232(n + 1);
  for (int j = 0; j < P->m; j++) tw32[j] = make_float2((float)tw[j].x, (float)tw[j].y);
  for (int k = 0; k <= n; k++) {
    q32[k] = make_float2((float)q[k].x, (float)q[k].y);
    t232[k] = make_float2((float)t2[k].x, (float)t2[k].y);
  }
  FZ_HIP(hipMalloc((void **)&P->tw32, sizeof(float2) * P->m));
  FZ_HIP(hipMalloc((void **)&P->q32, sizeof(float2) * (n + 1)));
  FZ_HIP(hipMalloc((void **)&P->t232, sizeof(float2) * (n + 1)));
  FZ_HIP(hipMemcpy(P->tw32, tw32.data(), sizeof(float2) * P->m, hipMemcpyHostToDevice));
  FZ_HIP(hipMemcpy(P->q32, q32.data(), sizeof(float2) * (n + 1), hipMemcpyHostToDevice));
  FZ_HIP(hipMemcpy(P->t232, t232.data(), sizeof(float2) * (n + 1), hipMemcpyHostToDevice));
  FZ_HIP(hipMalloc((void **)&P->tw, sizeof(double2) * P->m));
  FZ_HIP(hipMalloc((void **)&P->q, sizeof(double2) * (n + 1)));
  FZ_HIP(hipMalloc((void **)&P->t2, sizeof(double2) * (n + 1)));
  FZ_HIP(hipMemcpy(P->tw, tw.data(), sizeof(double2) * P->m, hipMemcpyHostToDevice));
  FZ_HIP(hipMemcpy(P->q, q.data(), sizeof(double2) * (n + 1), hipMemcpyHostToDevice));
  FZ_HIP(hipMemcpy(P->t2, t2.data(), sizeof(double2) * (n + 1), hipMemcpyHostToDevice));
  return FZ_OK;
}

static int fftplan_init(fz_fftplan *P, char b0, char b1, int n) {
  double nrm;
  int kf, kb;
  find_fft_c(b0, b1, &kf, &kb, &nrm);
  return fftplan_init_kinds(P, kf, kb, n);
}

static void fftplan_free(fz_fftplan *P) {
  if (P->tw) (void)hipFree(P->tw);
  if (P->q) (void)hipFree(P->q);
  if (P->t2) (void)hipFree(P->t2);
  P->tw = P->q = P->t2 = nullptr;
  if (P->tw32) (void)hipFree(P->tw32);
  if (P->q32) (void)hipFree(P->q32);
  if (P->t232) (void)hipFree(P->t232);
  P->tw32 = P->q32 = P->t232 = nullptr;
}

int fz_poisson_destroy(fz_poisson *p) {
  if (!p) return FZ_OK;
  fftplan_free(&p->px);
  fftplan_free(&p->py);
  double *bufs[9] = {p->lambdaxy, p->a, p->b, p->c, p->rhsbx, p->rhsby, p->rhsbz, p->work, p->piv};
  for (int q = 0; q < 9; q++)
    if (bufs[q]) (void)hipFree(bufs[q]);
  if (p->aux) (void)hipFree(p->aux);
  if (p->zpre) (void)hipFree(p->zpre);
  if (p->dpre) (void)hipFree(p->dpre);
  if (p->work32) (void)hipFree(p->work32);
  if (p->bcopy) (void)hipFree(p->bcopy);
  if (p->rslots) (void)hipFree(p->rslots);
  if (p->batch_ptrs) (void)hipFree(p->batch_ptrs);
  delete p;
  return FZ_OK;
}

extern "C" int fz_initsolver(fz_ctx *ctx, const double bcpre[6], double *lambdaxy_h, double *a_h, double *b_h,
                             double *c_h, double *normfft_h, double *rhsbx_h, double *rhsby_h, double *rhsbz_h) {
  if (!ctx) return fz_set_error(FZ_ERR_ARG, "fz_initsolver: null context");
  const int n1 = ctx->d.n1, n2 = ctx->d.n2, n3 = ctx->d.n3;
  const char *cbc = ctx->cbcpre;
  double bc[6] = {0, 0, 0, 0, 0, 0};
  if (bcpre) memcpy(bc, bcpre, sizeof(bc));
  std::vector<double> lx(n1), ly(n2), lxy((size_t)n1 * n2), a(n3), b(n3), c(n3);
  eigenvalues_c(n1, cbc[0], cbc[1], lx.data());
  for (int i = 0; i < n1; i++) lx[i] = lx[i] * (ctx->dli[0] * ctx->dli[0]);
  eigenvalues_c(n2, cbc[2], cbc[3], ly.data());
  for (int j = 0; j < n2; j++) ly[j] = ly[j] * (ctx->dli[1] * ctx->dli[1]);
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n1; i++) lxy[i + (size_t)n1 * j] = lx[i] + ly[j];
  const double *dzci = ctx->h_dzci.data(), *dzfi = ctx->h_dzfi.data();
  for (int k = 1; k <= n3; k++) {  // tridmatrix, src/initsolver.f90:153-182
    a[k - 1] = dzfi[k] * dzci[k - 1];
    c[k - 1] = dzfi[k] * dzci[k];
  }
  for (int k = 0; k < n3; k++) b[k] = -(a[k] + c[k]);
  double f0 = 0.0, f1 = 0.0;
  if (cbc[4] == 'D') f0 = -1.0; else if (cbc[4] == 'N') f0 = 1.0;
  if (cbc[5] == 'D') f1 = -1.0; else if (cbc[5] == 'N') f1 = 1.0;
  b[0] = b[0] + f0 * a[0];
  b[n3 - 1] = b[n3 - 1] + f1 * c[n3 - 1];
  // boundary right-hand sides, src/initsolver.f90:65-74
  const double dl0 = 1.0 / ctx->dli[0], dl1 = 1.0 / ctx->dli[1];
  const double dzc0 = 1.0 / dzci[0], dzcn = 1.0 / dzci[n3], dzf1 = 1.0 / dzfi[1], dzfn = 1.0 / dzfi[n3];
  std::vector<double> rbx((size_t)n2 * n3 * 2), rby((size_t)n1 * n3 * 2), rbz((size_t)n1 * n2 * 2);
  for (int ib = 0; ib < 2; ib++) {
    const double fx = bc_factor_c(cbc[0 + ib], ib, bc[0 + ib], dl0) / dl0 / dl0;
    const double fy = bc_factor_c(cbc[2 + ib], ib, bc[2 + ib], dl1) / dl1 / dl1;
    const double dlc = ib == 0 ? dzc0 : dzcn, dlf = ib == 0 ? dzf1 : dzfn;
    const double fz = bc_factor_c(cbc[4 + ib], ib, bc[4 + ib], dlc) / dlc / dlf;
    for (size_t q = 0; q < (size_t)n2 * n3; q++) rbx[q + (size_t)n2 * n3 * ib] = fx;
    for (size_t q = 0; q < (size_t)n1 * n3; q++) rby[q + (size_t)n1 * n3 * ib] = fy;
    for (size_t q = 0; q < (size_t)n1 * n2; q++) rbz[q + (size_t)n1 * n2 * ib] = fz;
  }
  double nx, ny;
  int kf, kb;
  find_fft_c(cbc[0], cbc[1], &kf, &kb, &nx);
  find_fft_c(cbc[2], cbc[3], &kf, &kb, &ny);
  double nf = 1.0;  // normfft, src/fft.f90:59-139
  nf = nf * nx * (double)n1;
  nf = nf * ny * (double)n2;
  const double normfft = 1.0 / nf;
  if (lambdaxy_h) memcpy(lambdaxy_h, lxy.data(), sizeof(double) * lxy.size());
  if (a_h) memcpy(a_h, a.data(), sizeof(double) * n3);
  if (b_h) memcpy(b_h, b.data(), sizeof(double) * n3);
  if (c_h) memcpy(c_h, c.data(), sizeof(double) * n3);
  if (normfft_h) *normfft_h = normfft;
  if (rhsbx_h) memcpy(rhsbx_h, rbx.data(), sizeof(double) * rbx.size());
  if (rhsby_h) memcpy(rhsby_h, rby.data(), sizeof(double) * rby.size());
  if (rhsbz_h) memcpy(rhsbz_h, rbz.data(), sizeof(double) * rbz.size());
  // device plan
  if (ctx->pois) {
    FZ_HIP(hipStreamSynchronize(ctx->st));
    fz_poisson_destroy(ctx->pois);
    ctx->pois = nullptr;
  }
  bool zero_b = true;
  for (double x : rbx) zero_b = zero_b && (x == 0.0);
  for (double x : rby) zero_b = zero_b && (x == 0.0);
  for (double x : rbz) zero_b = zero_b && (x == 0.0);
  if (cbc[4] == 'P' && cbc[5] == 'P' && n3 < 2)
    return fz_set_error(FZ_ERR_UNSUPPORTED, "fz_initsolver: periodic z needs n3 >= 2 (gaussel_periodic)");
  fz_poisson *P = new fz_poisson();
  memcpy(P->cbc, cbc, 6);
  P->normfft = normfft;
  P->periodic_z = (cbc[4] == 'P' && cbc[5] == 'P');
  P->rhsb_zero = zero_b ? 1 : 0;
  const bool z_singular = (cbc[4] == 'N' && cbc[5] == 'N') || P->periodic_z;
  P->null_col = (lxy[0] == 0.0 && z_singular) ? 0 : -1;
  int rc = fftplan_init(&P->px, cbc[0], cbc[1], n1);
  if (rc == FZ_OK) rc = fftplan_init(&P->py, cbc[2], cbc[3], n2);
  if (rc != FZ_OK) {
    fz_poisson_destroy(P);
    return rc;
  }
  const size_t N = (size_t)n1 * n2 * n3;
  struct {
    double **d;
    const std::vector<double> *h;
  } up[7] = {{&P->lambdaxy, &lxy}, {&P->a, &a}, {&P->b, &b}, {&P->c, &c}, {&P->rhsbx, &rbx}, {&P->rhsby, &rby},
             {&P->rhsbz, &rbz}};
  hipError_t e = hipSuccess;
  for (int q = 0; q < 7 && e == hipSuccess; q++) {
    e = hipMalloc((void **)up[q].d, sizeof(double) * up[q].h->size());
    if (e == hipSuccess) e = hipMemcpy(*up[q].d, up[q].h->data(), sizeof(double) * up[q].h->size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMalloc((void **)&P->work, sizeof(double) * N);
  if (e == hipSuccess) e = hipMalloc((void **)&P->piv, sizeof(double) * N);
  if (e == hipSuccess && P->periodic_z) e = hipMalloc((void **)&P->aux, sizeof(double) * N);
  if (e == hipSuccess && !P->periodic_z && ctx->thomas_pre) {
    e = hipMalloc((void **)&P->zpre, sizeof(double) * N);
    if (e == hipSuccess) e = hipMalloc((void **)&P->dpre, sizeof(double) * N);
  }
  if (e != hipSuccess) {
    fz_poisson_destroy(P);
    return fz_set_error(FZ_ERR_HIP, "fz_initsolver: %s", hipGetErrorString(e));
  }
  ctx->pois = P;
  if (P->zpre) {  // the pivots, once
    const long ps = (long)n1 * n2;
    FZ_LAUNCH(ctx, "piv_pre", k_piv_pre, dim3((int)((ps + 255) / 256)), dim3(256), 0, P->zpre, P->dpre, P->lambdaxy,
              P->a, P->b, P->c, n1, n2, n3);
    FZ_LAUNCHED();
  }
  if (P->periodic_z) {  // the operator's second solve of gaussel_periodic, once
    const long ps = (long)n1 * n2;
    FZ_LAUNCH(ctx, "periodic_p2", k_periodic_p2, dim3((int)((ps + 255) / 256)), dim3(256), 0, P->piv, P->aux,
              P->lambdaxy, P->a, P->b, P->c, n1, n2, n3);
    FZ_LAUNCHED();
  }
  return FZ_OK;
}

template <typename R, typename TI, typename TO>
static int launch_r2r(fz_ctx *ctx, const char *name, const fz_fftplan &P, int kind, const TI *src, LineMap in,
                      TO *dst, LineMap out, int L1, int L2, double scale, int accum = 0,
                      const Batch &bt = kNoBatch) {
  const bool quarter = kind_is_quarter(kind);
  if (p2_length(P.m)) {
    FillSrc off;
    memset(&off, 0, sizeof(off));
    const bool done =
        quarter ? dispatch_p2<true, R, TI, TO>(ctx, name, P, kind, src, in, dst, out, L1, L2, scale, accum, off, bt)
                : dispatch_p2<false, R, TI, TO>(ctx, name, P, kind, src, in, dst, out, L1, L2, scale, accum, off, bt);
    if (done) {
      FZ_LAUNCHED();
      return FZ_OK;
    }
  }
  int G;
  size_t lds;
  if (quarter) {
    G = 2048 / P.m;
    if (G < 1) G = 1;
    if (G > 64) G = 64;
    lds = (size_t)G * P.m * sizeof(typename CT<R>::C) * 2;
  } else {
    G = 4096 / P.m;
    if (G < 2) G = 2;
    if (G > 64) G = 64;
    G &= ~1;
    lds = (size_t)(G / 2) * P.m * sizeof(typename CT<R>::C) * 2;
  }
  RadList rl;
  for (int s = 0; s < 24; s++) rl.r[s] = P.rad[s];
  dim3 grid((L1 + G - 1) / G, L2, bt.src ? bt.n : 1);
  const PlanTabs<R> t = tabs_of(P, R());
  FZ_LAUNCH(ctx, name, (k_r2r<R, TI, TO>), grid, dim3(256), lds, src, in, dst, out, L1, P.n, P.m, kind, G, scale,
            accum, t.tw, t.q, t.t2, P.nrad, rl, bt);
  FZ_LAUNCHED();
  return FZ_OK;
}

// everything after the forward x transform: y transform, z solve, inverses
// (T = double: the reference's solver; T = float: one FP32 solve of the
// mixed-precision form, whose inverse x transform writes or adds into FP64 p)
template <typename T>
static int solver_rest(fz_ctx *ctx, T *W, double *p, int accum);

int fz_solver_dev(fz_ctx *ctx, double *p) {
  fz_poisson *P = ctx->pois;
  if (!P) return fz_set_error(FZ_ERR_STATE, "fz_solver: fz_initsolver has not been called");
  const int n1 = ctx->d.n1, n2 = ctx->d.n2, n3 = ctx->d.n3;
  const long sy = ctx->d.sy, sz = ctx->d.sz;
  const long ps = (long)n1 * n2;
  const LineMap p_x = {1 + sy + sz, sy, sz, 1};  // x-lines of the halo'd field, (j,k) -> line
  const LineMap w_x = {0, n1, ps, 1};
  FZ_TRY((launch_r2r<double, double, double>(ctx, "fft_x_fwd", P->px, P->px.kind_f, p, p_x, P->work, w_x, n2, n3,
                                             1.0)));
  return solver_rest<double>(ctx, P->work, p, 0);
}

int fz_fillps(fz_ctx *c, double dti, double rho0, const double *u, const double *v, const double *w, double *p);
int fz_updt_rhs_b(fz_ctx *c, double *p);

static FillSrc fill_src(fz_ctx *ctx, double dti, double rho0, const double *u, const double *v, const double *w) {
  FillSrc fs;
  memset(&fs, 0, sizeof(fs));
  fs.u = u;
  fs.v = v;
  fs.w = w;
  fs.dzfi = ctx->dzfi;
  fs.dzci = ctx->dzci;
  fs.rho0 = rho0;
  fs.dti = dti;
  fs.dtidxi = dti * ctx->dli[0];  // the scalars of fz_fillps
  fs.dtidyi = dti * ctx->dli[1];
  fs.dxi2 = ctx->dli[0] * ctx->dli[0];
  fs.dyi2 = ctx->dli[1] * ctx->dli[1];
  fs.dxi = ctx->dli[0];
  fs.dyi = ctx->dli[1];
  fs.sy = ctx->d.sy;
  fs.sz = ctx->d.sz;
  fs.on = SRC_FILLPS;
  return fs;
}

// fillps + updt_rhs_b + solver (src/main.f90:492-496). When every boundary
// right-hand side is zero, updt_rhs_b adds exact zeros and is skipped, and
// fillps is evaluated inside the load of the forward x transform (power-of-two
// x lengths); p then receives the correction pressure only. Otherwise the
// three routines run as they are.
int fz_projection_rhs_solve(fz_ctx *ctx, double dti, double rho0, const double *u, const double *v, const double *w,
                            double *p, const double *stc) {
  fz_poisson *P = ctx->pois;
  if (!P) return fz_set_error(FZ_ERR_STATE, "fz_solver: fz_initsolver has not been called");
  if (!P->rhsb_zero || !p2_length(P->px.m) || kind_is_quarter(P->px.kind_f)) {
    if (stc) return fz_set_error(FZ_ERR_STATE, "fz_projection_rhs_solve: device scalars need the fused path");
    FZ_TRY(fz_fillps(ctx, dti, rho0, u, v, w, p));
    FZ_TRY(fz_updt_rhs_b(ctx, p));
    return fz_solver_dev(ctx, p);
  }
  const int n1 = ctx->d.n1, n2 = ctx->d.n2, n3 = ctx->d.n3;
  const long sy = ctx->d.sy, sz = ctx->d.sz;
  const long ps = (long)n1 * n2;
  const LineMap p_x = {1 + sy + sz, sy, sz, 1};
  const LineMap w_x = {0, n1, ps, 1};
  FillSrc fs = fill_src(ctx, dti, rho0, u, v, w);
  fs.stc = stc;
  dispatch_p2<false, double, double, double>(ctx, "fillps_fft_x_fwd", P->px, P->px.kind_f, (const double *)nullptr,
                                             p_x, P->work, w_x, n2, n3, 1.0, 0, fs);
  FZ_LAUNCHED();
  return solver_rest<double>(ctx, P->work, p, 0);
}

template <typename T>
static int solver_rest(fz_ctx *ctx, T *W, double *p, int accum) {
  fz_poisson *P = ctx->pois;
  const int n1 = ctx->d.n1, n2 = ctx->d.n2, n3 = ctx->d.n3;
  const long sy = ctx->d.sy, sz = ctx->d.sz;
  const long ps = (long)n1 * n2;
  const bool mixed = sizeof(T) == sizeof(float);
  const long gauge = mixed ? P->null_col : -1;
  const LineMap p_x = {1 + sy + sz, sy, sz, 1};
  const LineMap w_x = {0, n1, ps, 1};
  const LineMap w_y = {0, 1, ps, n1};
  FZ_TRY((launch_r2r<T, T, T>(ctx, mixed ? "fft_y_fwd_f32" : "fft_y_fwd", P->py, P->py.kind_f, W, w_y, W, w_y, n1, n3,
                              1.0)));
  const int tpb = 256;
  const int nb = (int)((ps + tpb - 1) / tpb);
  if (!P->periodic_z) {
    // rows per chunk: the context's choice (FP32 columns move half the
    // bytes per load, so their default keeps twice the rows in flight)
    const int tk = mixed ? ctx->tk32 : ctx->tk64;
    const dim3 grid((int)((ps + 63) / 64)), blk(64);
    const char *name = mixed ? "thomas_f32" : "thomas";
    if (P->zpre && tk > 0) {  // pivots made at set-up
      if (tk == 8) FZ_LAUNCH(ctx, name, (k_thomas_pre<8, T>), grid, blk, 0, W, P->zpre, P->dpre, P->a, n1, n2, n3,
                             gauge);
      else FZ_LAUNCH(ctx, name, (k_thomas_pre<16, T>), grid, blk, 0, W, P->zpre, P->dpre, P->a, n1, n2, n3, gauge);
    } else if (ctx->thomas_rc && tk > 0) {  // no pivot array: the elimination is re-run per chunk
      if (tk == 8) FZ_LAUNCH(ctx, name, (k_thomas_rc<8, T>), grid, blk, 0, W, P->piv, P->lambdaxy, P->a, P->b, P->c,
                             n1, n2, n3, gauge, kNoBatch);
      else FZ_LAUNCH(ctx, name, (k_thomas_rc<16, T>), grid, blk, 0, W, P->piv, P->lambdaxy, P->a, P->b, P->c, n1, n2,
                     n3, gauge, kNoBatch);
    } else if (mixed && tk < 0 && ps % 2 == 0) {  // two FP32 columns per lane (tk = -rows per chunk)
      const dim3 g2((int)((ps / 2 + 63) / 64));
      float *W32 = (float *)(void *)W;
      if (P->zpre) {  // pivots made at set-up
        if (tk == -8) FZ_LAUNCH(ctx, name, k_thomas_pre_x2<8>, g2, blk, 0, W32, P->zpre, P->dpre, P->a, n1, n2, n3,
                                gauge);
        else FZ_LAUNCH(ctx, name, k_thomas_pre_x2<16>, g2, blk, 0, W32, P->zpre, P->dpre, P->a, n1, n2, n3, gauge);
      } else if (tk == -8) {
        FZ_LAUNCH(ctx, name, k_thomas_x2<8>, g2, blk, 0, W32, P->piv, P->lambdaxy, P->a, P->b, P->c, n1, n2, n3,
                  gauge);
      } else {
        FZ_LAUNCH(ctx, name, k_thomas_x2<16>, g2, blk, 0, W32, P->piv, P->lambdaxy, P->a, P->b, P->c, n1, n2, n3,
                  gauge);
      }
    } else if (tk == 8) FZ_LAUNCH(ctx, name, (k_thomas<8, T>), grid, blk, 0, W, P->piv, P->lambdaxy, P->a, P->b, P->c, n1, n2,
                           n3, gauge, kNoBatch);
    else if (tk == 32) FZ_LAUNCH(ctx, name, (k_thomas<32, T>), grid, blk, 0, W, P->piv, P->lambdaxy, P->a, P->b, P->c,
                                 n1, n2, n3, gauge, kNoBatch);
    else FZ_LAUNCH(ctx, name, (k_thomas<16, T>), grid, blk, 0, W, P->piv, P->lambdaxy, P->a, P->b, P->c, n1, n2, n3,
                   gauge, kNoBatch);
  } else {
    (void)nb;
    (void)tpb;
    if (mixed && ctx->tk32 < 0 && ps % 2 == 0)  // two FP32 columns per lane
      FZ_LAUNCH(ctx, "thomas_periodic_f32", k_thomas_periodic_x2<16>, dim3((int)((ps / 2 + 63) / 64)), dim3(64), 0,
                (float *)(void *)W, P->piv, P->aux, P->lambdaxy, P->a, P->b, P->c, n1, n2, n3, gauge);
    else
      FZ_LAUNCH(ctx, mixed ? "thomas_periodic_f32" : "thomas_periodic", (k_thomas_periodic_pc<16, T>),
                dim3((int)((ps + 63) / 64)), dim3(64), 0, W, P->piv, P->aux, P->lambdaxy, P->a, P->b, P->c, n1, n2,
                n3, gauge);
  }
  FZ_LAUNCHED();
  FZ_TRY((launch_r2r<T, T, T>(ctx, mixed ? "fft_y_bwd_f32" : "fft_y_bwd", P->py, P->py.kind_b, W, w_y, W, w_y, n1, n3,
                              1.0)));
  FZ_TRY((launch_r2r<T, T, double>(ctx, mixed ? "fft_x_bwd_f32" : "fft_x_bwd", P->px, P->px.kind_b, W, w_x, p, p_x,
                                   n2, n3, P->normfft, accum)));
  return FZ_OK;
}

extern "C" int fz_solver(fz_ctx *ctx, double *p) {
  if (!ctx || !p) return fz_set_error(FZ_ERR_ARG, "fz_solver: null argument");
  return fz_solver_dev(ctx, p);
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 transforms and line solves, FP64 residual refinement
// ---------------------------------------------------------------------------
static int mixed_buffers(fz_ctx *ctx) {
  fz_poisson *P = ctx->pois;
  if (!P->work32) FZ_HIP(hipMalloc((void **)&P->work32, sizeof(float) * ctx->ni));
  if (!P->rslots) {
    FZ_HIP(hipMalloc((void **)&P->rslots, sizeof(unsigned long long) * FZ_GAMMA_SLOTS));
    FZ_HIP(hipMemsetAsync(P->rslots, 0, sizeof(unsigned long long) * FZ_GAMMA_SLOTS, ctx->st));
  }
  return FZ_OK;
}

// one FP32 solve of the source fs: forward x transform of the source (fused
// into the load for power-of-two lengths), the rest in FP32; the result is
// written into p (accum 0) or added to it (accum 1)
static int mixed_pass(fz_ctx *ctx, const FillSrc &fs, double *p, int accum) {
  fz_poisson *P = ctx->pois;
  const int n1 = ctx->d.n1, n2 = ctx->d.n2, n3 = ctx->d.n3;
  const long sy = ctx->d.sy, sz = ctx->d.sz;
  const long ps = (long)n1 * n2;
  const LineMap p_x = {1 + sy + sz, sy, sz, 1};
  const LineMap w_x = {0, n1, ps, 1};
  float *W = P->work32;
  const char *name = fs.on == SRC_FILLPS ? "fillps_fft_x_fwd_f32" : (fs.on == SRC_B ? "fft_x_fwd_f32" : "residual_fft_x_fwd_f32");
  bool fused = false;
  if (p2_length(P->px.m)) {
    fused = kind_is_quarter(P->px.kind_f)
                ? dispatch_p2<true, float, double, float>(ctx, name, P->px, P->px.kind_f, (const double *)nullptr, p_x,
                                                          W, w_x, n2, n3, 1.0, 0, fs)
                : dispatch_p2<false, float, double, float>(ctx, name, P->px, P->px.kind_f, (const double *)nullptr,
                                                           p_x, W, w_x, n2, n3, 1.0, 0, fs);
    FZ_LAUNCHED();
  }
  if (!fused) {
    FZ_LAUNCH(ctx, "src_lines_f32", k_src_lines<float>, fz_grid_interior(ctx->d), fz_block3(), 0, ctx->d, fs, W);
    FZ_LAUNCHED();
    FZ_TRY((launch_r2r<float, float, float>(ctx, "fft_x_fwd_f32", P->px, P->px.kind_f, W, w_x, W, w_x, n2, n3, 1.0)));
  }
  return solver_rest<float>(ctx, W, p, accum);
}

// the homogeneous pressure conditions on the halo of x (the residual's
// operator is the solver's, whose boundary values sit in the right-hand side)
static int halo_homogeneous(fz_ctx *ctx, double *x) {
  const double zero[6] = {0, 0, 0, 0, 0, 0};
  return fz_boundp(ctx, ctx->pois->cbc, zero, x);
}

// x = solve(fs) with FP32 transforms, then nref refinements
// x += solve32(fs_b - A x) whose residual is formed in FP64
static int mixed_solve(fz_ctx *ctx, FillSrc fs, double *x, int nref) {
  FZ_TRY(mixed_buffers(ctx));
  fs.rmax = nullptr;
  FZ_TRY(mixed_pass(ctx, fs, x, 0));
  FillSrc fr = fs;
  fr.on = fs.on == SRC_FILLPS ? SRC_FILLPS_RES : SRC_B_RES;
  fr.x = x;
  for (int it = 0; it < nref; it++) {
    FZ_TRY(halo_homogeneous(ctx, x));
    FZ_TRY(mixed_pass(ctx, fr, x, 1));
  }
  return FZ_OK;
}

// fz_projection_rhs_solve in mixed precision: fillps in FP64 evaluated in the
// load of the first FP32 transform, nref FP64-residual refinements
int fz_projection_rhs_solve_mixed(fz_ctx *ctx, double dti, double rho0, const double *u, const double *v,
                                  const double *w, double *p, int nref, const double *stc) {
  fz_poisson *P = ctx->pois;
  if (!P) return fz_set_error(FZ_ERR_STATE, "fz_solver: fz_initsolver has not been called");
  if (!P->rhsb_zero)
    return fz_set_error(FZ_ERR_UNSUPPORTED, "mixed-precision projection: nonzero boundary pressure values");
  FillSrc fs = fill_src(ctx, dti, rho0, u, v, w);
  fs.stc = stc;
  return mixed_solve(ctx, fs, p, nref);
}

// whether the projection can take 1/dt from device memory (the fused
// forward transform; graph capture of the step depends on it)
int fz_projection_device_scalars_ok(fz_ctx *ctx, int mixed) {
  fz_poisson *P = ctx->pois;
  if (!P) return 0;
  if (mixed) return 1;
  return P->rhsb_zero && p2_length(P->px.m) && !kind_is_quarter(P->px.kind_f);
}

__global__ void k_slots_max(unsigned long long *__restrict__ slots, double *__restrict__ out) {
  __shared__ unsigned long long s[FZ_GAMMA_SLOTS];
  const int t = threadIdx.x;
  s[t] = slots[t];
  slots[t] = 0ull;
  __syncthreads();
  for (int o = FZ_GAMMA_SLOTS / 2; o > 0; o >>= 1) {
    if (t < o) s[t] = s[t] > s[t + o] ? s[t] : s[t + o];
    __syncthreads();
  }
  if (t == 0) *out = __longlong_as_double((long long)s[0]);
}

// max |value| of the slots (cleared for the next use), synchronising
static int read_slots(fz_ctx *ctx, double *out) {
  fz_poisson *P = ctx->pois;
  FZ_LAUNCH(ctx, "slots_max", k_slots_max, dim3(1), dim3(FZ_GAMMA_SLOTS), 0, P->rslots, ctx->red);
  FZ_LAUNCHED();
  FZ_HIP(hipMemcpyAsync(ctx->h_small + 12, ctx->red, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
  FZ_HIP(hipStreamSynchronize(ctx->st));
  *out = ctx->h_small[12];
  return FZ_OK;
}

// max |b - A x| over the interior (x's halo receives the homogeneous
// conditions); b in the halo'd layout
static int residual_max(fz_ctx *ctx, const double *b, double *x, double *res) {
  FZ_TRY(halo_homogeneous(ctx, x));
  FillSrc fr = fill_src(ctx, 0.0, 0.0, nullptr, nullptr, nullptr);
  fr.on = SRC_B_RES;
  fr.b = b;
  fr.x = x;
  fr.rmax = ctx->pois->rslots;
  FZ_LAUNCH(ctx, "residual_max", k_src_lines<float>, fz_grid_interior(ctx->d), fz_block3(), 0, ctx->d, fr,
            ctx->pois->work32);
  FZ_LAUNCHED();
  return read_slots(ctx, res);
}

// fz_solver with FP32 transforms and FP64-residual iterative refinement:
// p holds the right-hand side on entry and the solution on exit (interior);
// refinements continue until max|b - A x| <= tol * max|b| or maxit of them
extern "C" int fz_solver_mixed(fz_ctx *ctx, double *p, double tol, int maxit, int *iters, double *relres) {
  if (!ctx || !p || maxit < 0) return fz_set_error(FZ_ERR_ARG, "fz_solver_mixed: bad argument");
  fz_poisson *P = ctx->pois;
  if (!P) return fz_set_error(FZ_ERR_STATE, "fz_solver_mixed: fz_initsolver has not been called");
  FZ_TRY(mixed_buffers(ctx));
  FZ_TRY(fz_reduce_ensure(ctx, 1));
  if (!P->bcopy) FZ_HIP(hipMalloc((void **)&P->bcopy, sizeof(double) * ctx->nh));
  FZ_HIP(hipMemcpyAsync(P->bcopy, p, sizeof(double) * ctx->nh, hipMemcpyDeviceToDevice, ctx->st));
  FillSrc fb = fill_src(ctx, 0.0, 0.0, nullptr, nullptr, nullptr);
  fb.on = SRC_B;
  fb.b = P->bcopy;
  fb.rmax = P->rslots;
  FZ_TRY(mixed_pass(ctx, fb, p, 0));
  double bmax;
  FZ_TRY(read_slots(ctx, &bmax));
  FillSrc fr = fb;
  fr.on = SRC_B_RES;
  fr.x = p;
  int it = 0;
  double res = bmax, prev = 0.0;
  while (it < maxit) {
    FZ_TRY(halo_homogeneous(ctx, p));
    FZ_TRY(mixed_pass(ctx, fr, p, 1));  // the residual of the current x, then its correction
    it++;
    FZ_TRY(read_slots(ctx, &res));
    if (res <= tol * bmax) break;
    if (it > 1 && res > 0.5 * prev) break;  // at the FP64 rounding floor of the residual
    prev = res;
  }
  if (relres) {
    FZ_TRY(residual_max(ctx, P->bcopy, p, &res));
    *relres = bmax > 0.0 ? res / bmax : res;
  }
  if (iters) *iters = it;
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// batched solves: the solvers of n contexts of one shape in one launch per stage
// ---------------------------------------------------------------------------
static bool same_operator(const fz_ctx *a, const fz_ctx *b) {
  if (a->d.n1 != b->d.n1 || a->d.n2 != b->d.n2 || a->d.n3 != b->d.n3) return false;
  if (memcmp(a->pois->cbc, b->pois->cbc, 6) != 0) return false;
  if (a->dli[0] != b->dli[0] || a->dli[1] != b->dli[1]) return false;
  return a->h_dzc == b->h_dzc && a->h_dzf == b->h_dzf;
}

extern "C" int fz_solver_batched(fz_ctx *const *ctxs, double *const *ps, int n) {
  if (!ctxs || !ps || n < 0) return fz_set_error(FZ_ERR_ARG, "fz_solver_batched: bad argument");
  if (n == 0) return FZ_OK;
  fz_ctx *c0 = ctxs[0];
  for (int q = 0; q < n; q++) {
    if (!ctxs[q] || !ps[q]) return fz_set_error(FZ_ERR_ARG, "fz_solver_batched: null context or field %d", q);
    if (!ctxs[q]->pois)
      return fz_set_error(FZ_ERR_STATE, "fz_solver_batched: fz_initsolver has not been called on context %d", q);
    if (!same_operator(c0, ctxs[q]))
      return fz_set_error(FZ_ERR_ARG, "fz_solver_batched: context %d differs in grid or pressure conditions", q);
  }
  fz_poisson *P = c0->pois;
  if (P->periodic_z) return fz_set_error(FZ_ERR_UNSUPPORTED, "fz_solver_batched: periodic z is not batched");
  if (n > 65535) return fz_set_error(FZ_ERR_ARG, "fz_solver_batched: at most 65535 solves per call");
  // each context's pending work on its fields first
  for (int q = 0; q < n; q++) FZ_HIP(hipStreamSynchronize(ctxs[q]->st));
  // device arrays: p, work, pivots of every case
  std::vector<void *> h(3 * (size_t)n);
  for (int q = 0; q < n; q++) {
    h[q] = ps[q];
    h[n + q] = ctxs[q]->pois->work;
    h[2 * n + q] = ctxs[q]->pois->piv;
  }
  if (P->batch_cap < 3 * (size_t)n) {
    if (P->batch_ptrs) FZ_HIP(hipFree(P->batch_ptrs));
    P->batch_ptrs = nullptr;
    FZ_HIP(hipMalloc((void **)&P->batch_ptrs, sizeof(void *) * 3 * (size_t)n));
    P->batch_cap = 3 * (size_t)n;
  }
  FZ_HIP(hipMemcpy(P->batch_ptrs, h.data(), sizeof(void *) * 3 * (size_t)n, hipMemcpyHostToDevice));
  FZ_TRY(fz_batch_solve(c0, P->batch_ptrs, n));
  FZ_HIP(hipStreamSynchronize(c0->st));
  return FZ_OK;
}

// The stages of n same-shaped solves (non-periodic z) in one launch each, on
// c0's stream: dptrs is a device array of the n right-hand sides (halo'd,
// overwritten by the solutions), then the n work arrays, then the n pivot
// arrays. No synchronisation and no allocation (a batched step enqueues it).
int fz_batch_solve(fz_ctx *c0, void **dptrs, int n) {
  fz_poisson *P = c0->pois;
  if (!P || P->periodic_z) return fz_set_error(FZ_ERR_UNSUPPORTED, "batched solve: periodic z is not batched");
  void **dp = dptrs;
  const Batch p_w = {(const void *const *)dp, dp + n, n};              // p -> work
  const Batch w_w = {(const void *const *)(dp + n), dp + n, n};        // work in place
  const Batch w_p = {(const void *const *)(dp + n), dp, n};            // work -> p
  const Batch w_d = {(const void *const *)(dp + n), dp + 2 * n, n};    // Thomas: work, pivots
  const int n1 = c0->d.n1, n2 = c0->d.n2, n3 = c0->d.n3;
  const long sy = c0->d.sy, sz = c0->d.sz, ps1 = (long)n1 * n2;
  const LineMap p_x = {1 + sy + sz, sy, sz, 1};
  const LineMap w_x = {0, n1, ps1, 1};
  const LineMap w_y = {0, 1, ps1, n1};
  FZ_TRY((launch_r2r<double, double, double>(c0, "fft_x_fwd_batch", P->px, P->px.kind_f, (const double *)nullptr,
                                             p_x, (double *)nullptr, w_x, n2, n3, 1.0, 0, p_w)));
  FZ_TRY((launch_r2r<double, double, double>(c0, "fft_y_fwd_batch", P->py, P->py.kind_f, (const double *)nullptr,
                                             w_y, (double *)nullptr, w_y, n1, n3, 1.0, 0, w_w)));
  const dim3 grid((int)((ps1 + 63) / 64), n), blk(64);
  FZ_LAUNCH(c0, "thomas_batch", (k_thomas<16, double>), grid, blk, 0, (double *)nullptr, (double *)nullptr,
            P->lambdaxy, P->a, P->b, P->c, n1, n2, n3, -1L, w_d);
  FZ_LAUNCHED();
  FZ_TRY((launch_r2r<double, double, double>(c0, "fft_y_bwd_batch", P->py, P->py.kind_b, (const double *)nullptr,
                                             w_y, (double *)nullptr, w_y, n1, n3, 1.0, 0, w_w)));
  FZ_TRY((launch_r2r<double, double, double>(c0, "fft_x_bwd_batch", P->px, P->px.kind_b, (const double *)nullptr,
                                             w_x, (double *)nullptr, p_x, n2, n3, P->normfft, 0, w_p)));
  return FZ_OK;
}

// fft, src/fft.f90:176-187: one r2r kind over `howmany` contiguous lines of
// length n (in place, unnormalised); builds a transient plan
extern "C" int fz_r2r(fz_ctx *ctx, int kind, int n, int howmany, double *x) {
  if (!ctx || !x || n < 1 || howmany < 0) return fz_set_error(FZ_ERR_ARG, "fz_r2r: bad argument");
  const int kinds[8] = {K_R2HC, K_HC2R, K_REDFT01, K_REDFT10, K_REDFT11, K_RODFT01, K_RODFT10, K_RODFT11};
  bool ok = false;
  for (int q = 0; q < 8; q++) ok = ok || (kinds[q] == kind);
  if (!ok) return fz_set_error(FZ_ERR_UNSUPPORTED, "fz_r2r: unsupported kind %d", kind);
  if (howmany == 0) return FZ_OK;
  fz_fftplan P;
  FZ_TRY(fftplan_init_kinds(&P, kind, kind, n));
  const LineMap lines = {0, n, 0, 1};
  int rc = launch_r2r<double, double, double>(ctx, "fft_lines", P, kind, x, lines, x, lines, howmany, 1, 1.0);
  if (rc == FZ_OK && hipStreamSynchronize(ctx->st) != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "fz_r2r: sync");
  fftplan_free(&P);
  return rc;
}
