/*
 * CPU oracle: a plain-C restatement of the reference's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see fz_oracle.h). Every routine cites the
 * reference file:line it restates; the arithmetic keeps the reference's
 * operation order so that, built with -ffp-contract=off, it rounds like the
 * reference's own CPU build.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *  - or_r2r: every kind against scipy.fft golden vectors (tests/golden/r2r.npz)
 *    and O(n^2) brute-force sums; round trip equals 1/normfft.
 *  - or_initsolver/or_solver: discrete eigenfunctions solved exactly, the
 *    chk_poisson residual, dense numpy.linalg.solve of the 3-D operator on
 *    tiny grids, PP/NN/DD in x/y and NN/DD/PP in z.
 *  - fillps/solver/correc: chkdiv after projection at round-off.
 *  - stencils: closed forms (uniform fields give zero RHS, Taylor-Green
 *    single-phase momentum RHS against a direct finite-difference sum,
 *    planar interface normals, sphere curvature), conservation of sum(psi)
 *    under periodic flux-form transport, mirror symmetry.
 *  - whole steps: the Hysing rising-bubble contour fixture of the reference's
 *    own test (tests/rising_bubble/c1g2l3s.txt), run on the GPU path.
 */
#include "fz_oracle.h"
#include "fz_oracle_math.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS 2.220446049250313080847e-16 /* epsilon(1._rp), src/param.f90:17 */
#define PI 3.141592653589793115998      /* acos(-1._rp), src/param.f90:15 */

/* Planted mistakes. The pin tests show that each of their checks fails on a
 * plausible slip in this restatement: a second build of this file,
 * liboracle_planted.so (-DOR_PLANTED), carries numbered slips, and its
 * global or_planted selects one of them (0: none). In liboracle.so every
 * PLANTED(k) is the constant 0, so the restatement is unchanged. */
#ifdef OR_PLANTED
int or_planted = 0;
#define PLANTED(k) (or_planted == (k))
#else
#define PLANTED(k) 0
#endif
enum {
  PL_ACDI_SHARP_SIGN = 1,  /* sharpening added to the diffusive flux instead of subtracted */
  PL_ACDI_QUARTER,         /* 0.5*gam instead of 0.25*gam in front of the sharpening */
  PL_ACDI_TANH_ARG,        /* tanh(phi_face/seps) instead of tanh(phi_face/(2 seps)) */
  PL_ACDI_NORMAL_COMP,     /* the y normal on the x+ face */
  PL_ACDI_DZCI_M,          /* dzci(k) instead of dzci(k-1) on the z- face */
  PL_THINC_UPWIND,         /* the downwind cell's normal and phi for u > 0 */
  PL_THINC_OFFSET,         /* the face at -1/2 instead of +1/2 of the upwind cell for u > 0 */
  PL_THINC_NORMAL_COMP,    /* the y normal on x faces */
  PL_OTH_BRACKBILL,        /* the face density missing from the Brackbill surface tension */
  PL_OTH_KAPPA_FACE,       /* the cell's own curvature instead of the face average */
  PL_OTH_EXTRAP,           /* (1+dt_r)*po - dt_r*pn: the two pressure levels swapped */
  PL_OTH_SPLIT,            /* rho0/rho_face instead of rho_face/rho0 */
  PL_OTH_SURF_FACTOR,      /* sigma/(rho1+rho2) instead of 2 sigma/(rho1+rho2) */
  PL_DT_VISC,              /* the factor 2 of the viscous limit dropped */
  PL_DT_CAP_RHO,           /* the larger density in the capillary limit */
  PL_DT_GRAV,              /* |g|/dlmin**2 for the gravity limit */
  PL_DT_DLMIN,             /* dlmin over dx and dy only */
  PL_DT_COMBINE,           /* the 0.5 of the combined limit dropped */
  PL_INIT_WIDTH,           /* tanh(r/eps) instead of tanh(r/(2 eps)) in the smooth step */
  PL_INIT_PERIODIC,        /* no periodic images of the spheres */
  PL_INIT_FLM_INDEX,       /* the film's distance taken at zc(k-1) instead of zc(k) */
};

static inline long HX(const int *n, int i, int j, int k) {
  return i + (long)(n[0] + 2) * (j + (long)(n[1] + 2) * k);
}
static inline long IX(const int *n, int i, int j, int k) {
  return (i - 1) + (long)n[0] * ((j - 1) + (long)n[1] * (k - 1));
}
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* ======================================================================
 * r2r transforms with FFTW's unnormalised definitions (the reference plans
 * them in src/fft.f90:63-139 with the kinds chosen by find_fft,
 * src/fft.f90:189-242).
 *
 * Algorithm: a complex Stockham autosort FFT (mixed radix 4,2,3,5,7,11,13)
 * with two real lines packed into one complex line; DCT-II/III by Makhoul's
 * reordering, DST-II/III by sign alternation and reversal, DCT-IV/DST-IV by
 * a zero-padded transform of length 2n. The operation order is the one of
 * the HIP kernel k_r2r, so that the two agree to the last bit; correctness
 * is pinned independently by scipy golden vectors and brute-force sums.
 * ====================================================================== */
typedef struct { double x, y; } cplx;

typedef struct {
  int kind, n, m, nrad;
  int rad[24];
  cplx *tw; /* (cos, -sin)(2 pi j / m) */
  cplx *q;  /* (cos, sin)(pi k / 2n), k <= n */
  cplx *t2; /* (cos, sin)(pi (2k+1) / 4n) */
} rplan;

static int is_quarter(int kind) { return kind == OR_REDFT11 || kind == OR_RODFT11; }
static int is_forward(int kind) {
  return kind == OR_R2HC || kind == OR_REDFT10 || kind == OR_RODFT10 || kind == OR_REDFT11 || kind == OR_RODFT11;
}

static rplan *rplan_new(int kind, int n) {
  rplan *p = (rplan *)calloc(1, sizeof(rplan));
  p->kind = kind;
  p->n = n;
  p->m = is_quarter(kind) ? 2 * n : n;
  int m = p->m, k = 0;
  while (m % 4 == 0) { p->rad[k++] = 4; m /= 4; }
  while (m % 2 == 0) { p->rad[k++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2)
    while (m % f == 0) {
      if (f > 13 || k >= 24) { fprintf(stderr, "or_r2r: length %d has a prime factor above 13\n", p->m); abort(); }
      p->rad[k++] = f;
      m /= f;
    }
  p->nrad = k;
  p->tw = (cplx *)malloc(sizeof(cplx) * p->m);
  p->q = (cplx *)malloc(sizeof(cplx) * (n + 1));
  p->t2 = (cplx *)malloc(sizeof(cplx) * (n + 1));
  for (int j = 0; j < p->m; j++) {
    p->tw[j].x = om_cos(2.0 * PI * (double)j / (double)p->m);
    p->tw[j].y = -om_sin(2.0 * PI * (double)j / (double)p->m);
  }
  for (int j = 0; j <= n; j++) {
    p->q[j].x = om_cos(PI * (double)j / (2.0 * n));
    p->q[j].y = om_sin(PI * (double)j / (2.0 * n));
    p->t2[j].x = om_cos(PI * (double)(2 * j + 1) / (4.0 * n));
    p->t2[j].y = om_sin(PI * (double)(2 * j + 1) / (4.0 * n));
  }
  return p;
}
static void rplan_free(rplan *p) {
  if (!p) return;
  free(p->tw); free(p->q); free(p->t2);
  free(p);
}

static inline cplx cmul(cplx a, cplx b) { cplx r = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; return r; }
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.x + b.x, a.y + b.y}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.x - b.x, a.y - b.y}; return r; }

/* one Stockham pass of radix R on one complex line of length m */
static void stockham_pass(const cplx *X, cplx *Y, int m, int R, int Ns, double sign, const cplx *tw) {
  const int nb = m / R, tstride = m / (Ns * R);
  for (int b = 0; b < nb; b++) {
    const int k = b % Ns, base = (b / Ns) * Ns * R + k;
    cplx v[16];
    for (int t = 0; t < R; t++) {
      cplx xv = X[b + t * nb];
      if (t > 0 && k > 0) {
        cplx w = tw[t * k * tstride];
        if (sign > 0) w.y = -w.y;
        xv = cmul(xv, w);
      }
      v[t] = xv;
    }
    if (R == 2) {
      Y[base] = cadd(v[0], v[1]);
      Y[base + Ns] = csub(v[0], v[1]);
    } else if (R == 4) {
      const cplx s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
      const cplx s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
      const cplx r = {-sign * d13.y, sign * d13.x};
      Y[base] = cadd(s02, s13);
      Y[base + Ns] = cadd(d02, r);
      Y[base + 2 * Ns] = csub(s02, s13);
      Y[base + 3 * Ns] = csub(d02, r);
    } else if (R == 3) {
      const double c3 = 0.86602540378443864676;
      const cplx t1 = cadd(v[1], v[2]);
      const cplx t2 = {v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y};
      const cplx d = csub(v[1], v[2]);
      const cplx t3 = {-sign * c3 * d.y, sign * c3 * d.x};
      Y[base] = cadd(v[0], t1);
      Y[base + Ns] = cadd(t2, t3);
      Y[base + 2 * Ns] = csub(t2, t3);
    } else {
      const int st = m / R;
      for (int s = 0; s < R; s++) {
        cplx acc = v[0];
        for (int t = 1; t < R; t++) {
          const cplx w0 = tw[((t * s) % R) * st];
          const cplx w = {w0.x, sign > 0 ? -w0.y : w0.y};
          acc = cadd(acc, cmul(v[t], w));
        }
        Y[base + s * Ns] = acc;
      }
    }
  }
}

/* complex FFT of one line in A (B scratch); returns the buffer holding the result */
static cplx *cfft(const rplan *p, double sign, cplx *A, cplx *B) {
  cplx *cur = A, *oth = B;
  int Ns = 1;
  for (int s = 0; s < p->nrad; s++) {
    stockham_pass(cur, oth, p->m, p->rad[s], Ns, sign, p->tw);
    cplx *t = cur; cur = oth; oth = t;
    Ns *= p->rad[s];
  }
  return cur;
}

/* unpack line g (0 = real part, 1 = imaginary part) of two packed transforms */
static inline void unpack(const cplx *Z, int m, int kk, int g, double *re, double *im) {
  const cplx zk = Z[kk], znk = Z[(m - kk) % m];
  if (g == 0) { *re = 0.5 * (zk.x + znk.x); *im = 0.5 * (zk.y - znk.y); }
  else { *re = 0.5 * (zk.y + znk.y); *im = 0.5 * (znk.x - zk.x); }
}

/* transform lines a and b (b may be NULL: a zero line) of a packed kind;
 * w is scratch of 4n cplx */
static void r2r_pair(const rplan *p, double *a, double *b, cplx *w) {
  const int n = p->n, m = p->m, kind = p->kind;
  cplx *A = w, *B = w + m;
  double *lines[2] = {a, b};
  for (int k = 0; k < m; k++) {
    double ra, rb;
    cplx z;
#define RAW(g, e) (lines[g] ? lines[g][e] : 0.0)
    switch (kind) {
    case OR_R2HC: z.x = RAW(0, k); z.y = RAW(1, k); break;
    case OR_REDFT10:
    case OR_RODFT10: {
      const int e = (k <= (n - 1) / 2) ? 2 * k : 2 * (n - 1 - k) + 1;
      const double sg = (kind == OR_RODFT10 && (e & 1)) ? -1.0 : 1.0;
      z.x = sg * RAW(0, e); z.y = sg * RAW(1, e);
      break;
    }
    case OR_HC2R: {
      double ia, ib;
      if (k == 0) { ra = RAW(0, 0); ia = 0.0; rb = RAW(1, 0); ib = 0.0; }
      else if (2 * k < n) { ra = RAW(0, k); ia = RAW(0, n - k); rb = RAW(1, k); ib = RAW(1, n - k); }
      else if (2 * k == n) { ra = RAW(0, k); ia = 0.0; rb = RAW(1, k); ib = 0.0; }
      else { ra = RAW(0, n - k); ia = -RAW(0, k); rb = RAW(1, n - k); ib = -RAW(1, k); }
      z.x = ra - ib; z.y = ia + rb;
      break;
    }
    default: { /* REDFT01, RODFT01 */
      double xa, xna, xb, xnb;
      if (kind == OR_REDFT01) {
        xa = RAW(0, k); xb = RAW(1, k);
        xna = (k == 0) ? 0.0 : RAW(0, n - k);
        xnb = (k == 0) ? 0.0 : RAW(1, n - k);
      } else {
        xa = RAW(0, n - 1 - k); xb = RAW(1, n - 1 - k);
        xna = (k == 0) ? 0.0 : RAW(0, k - 1);
        xnb = (k == 0) ? 0.0 : RAW(1, k - 1);
      }
      const double cs = p->q[k].x, sn = p->q[k].y;
      const double rca = xa * cs + xna * sn, ica = xa * sn - xna * cs;
      const double rcb = xb * cs + xnb * sn, icb = xb * sn - xnb * cs;
      z.x = rca - icb; z.y = ica + rcb;
      break;
    }
    }
#undef RAW
    A[k] = z;
  }
  const cplx *Z = cfft(p, is_forward(kind) ? -1.0 : 1.0, A, B);
  for (int g = 0; g < 2; g++) {
    double *out = lines[g];
    if (!out) continue;
    for (int k = 0; k < n; k++) {
      double re, im, val;
      switch (kind) {
      case OR_R2HC: {
        const int kk = (2 * k <= n) ? k : n - k;
        unpack(Z, m, kk, g, &re, &im);
        val = (2 * k <= n) ? re : im;
        break;
      }
      case OR_REDFT10:
      case OR_RODFT10: {
        const int kk = (kind == OR_REDFT10) ? k : n - 1 - k;
        unpack(Z, m, kk, g, &re, &im);
        val = 2.0 * (p->q[kk].x * re + p->q[kk].y * im);
        break;
      }
      case OR_HC2R: val = g ? Z[k].y : Z[k].x; break;
      default: { /* REDFT01, RODFT01 */
        const int srcm = (k & 1) ? n - 1 - (k >> 1) : (k >> 1);
        val = g ? Z[srcm].y : Z[srcm].x;
        if (kind == OR_RODFT01 && (k & 1)) val = -val;
        break;
      }
      }
      if (g == 0) ((double *)(w + 2 * m))[k] = val; else ((double *)(w + 2 * m))[n + k] = val;
    }
  }
  for (int g = 0; g < 2; g++)
    if (lines[g]) memcpy(lines[g], ((double *)(w + 2 * m)) + g * n, sizeof(double) * n);
}

/* DCT-IV / DST-IV of one line (zero-padded complex transform of length 2n) */
static void r2r_quarter(const rplan *p, double *x, cplx *w) {
  const int n = p->n, m = p->m, kind = p->kind;
  cplx *A = w, *B = w + m;
  for (int k = 0; k < m; k++) {
    if (k < n) {
      const double sg = (kind == OR_RODFT11 && (k & 1)) ? -1.0 : 1.0;
      const double xv = sg * x[k];
      A[k].x = xv * p->q[k].x;
      A[k].y = -xv * p->q[k].y;
    } else {
      A[k].x = 0.0;
      A[k].y = 0.0;
    }
  }
  const cplx *Z = cfft(p, -1.0, A, B);
  for (int k = 0; k < n; k++) {
    const int kk = (kind == OR_REDFT11) ? k : n - 1 - k;
    x[k] = 2.0 * (p->t2[kk].x * Z[kk].x + p->t2[kk].y * Z[kk].y);
  }
}

/* lines (l, l+1) are packed together for l even, the last odd one alone */
static void r2r_lines(const rplan *p, int howmany, double *x, long ldist, cplx *w) {
  if (is_quarter(p->kind)) {
    for (int l = 0; l < howmany; l++) r2r_quarter(p, x + (long)l * ldist, w);
  } else {
    for (int l = 0; l < howmany; l += 2)
      r2r_pair(p, x + (long)l * ldist, (l + 1 < howmany) ? x + (long)(l + 1) * ldist : NULL, w);
  }
}

void or_r2r(int kind, int n, int howmany, double *x) {
  if (n <= 0) return;
  rplan *p = rplan_new(kind, n);
  const int npair = (howmany + 1) / 2;
#pragma omp parallel
  {
    cplx *w = (cplx *)malloc(sizeof(cplx) * (4 * p->m + 4));
#pragma omp for schedule(static)
    for (int g = 0; g < npair; g++) {
      const int l = 2 * g;
      r2r_lines(p, (l + 1 < howmany) ? 2 : 1, x + (long)l * n, n, w);
    }
    free(w);
  }
  rplan_free(p);
}

/* find_fft for cell-centred grids, src/fft.f90:195-217 */
static void find_fft_c(char b0, char b1, int *kf, int *kb, double *norm) {
  if (b0 == 'P' && b1 == 'P') { *kf = OR_R2HC; *kb = OR_HC2R; *norm = 1.0; }
  else if (b0 == 'N' && b1 == 'N') { *kf = OR_REDFT10; *kb = OR_REDFT01; *norm = 2.0; }
  else if (b0 == 'D' && b1 == 'D') { *kf = OR_RODFT10; *kb = OR_RODFT01; *norm = 2.0; }
  else if (b0 == 'N' && b1 == 'D') { *kf = OR_REDFT11; *kb = OR_REDFT11; *norm = 2.0; }
  else { *kf = OR_RODFT11; *kb = OR_RODFT11; *norm = 2.0; } /* 'DN' */
}

/* ======================================================================
 * Grid, src/initgrid.f90:14-78 (uniform and the four clustering laws)
 * ====================================================================== */
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

void or_initgrid(int gtype, int n, double gr, double lz, double *dzc, double *dzf, double *zc, double *zf) {
  zf[0] = 0.0;
  for (int k = 1; k <= n; k++) {
    /* (k-0.)/(1.*n) is evaluated in default (single) precision, src/initgrid.f90:38 */
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
}

/* ======================================================================
 * Poisson operator set-up, src/initsolver.f90:15-228
 * ====================================================================== */
static void eigenvalues_c(int n, char b0, char b1, double *lambda) {
  for (int l = 1; l <= n; l++) {
    double arg;
    if (b0 == 'P' && b1 == 'P') arg = ((double)(2 * (l - 1)) * PI) / (double)n;
    else if (b0 == 'N' && b1 == 'N') arg = ((double)(l - 1) * PI) / (double)n;
    else if (b0 == 'D' && b1 == 'D') arg = ((double)l * PI) / (double)n;
    else arg = ((double)(2 * l - 1) * PI) / (2.0 * n);
    lambda[l - 1] = -2.0 * (1.0 - cos(arg));
  }
}

static double bc_factor_c(char cbc, int ib, double bc, double dlc) {
  if (cbc == 'P') return 0.0;
  if (cbc == 'D') return -2.0 * bc;
  const double sgn = (ib == 0) ? 1.0 : -1.0;
  return sgn * dlc * bc; /* 'N' */
}

void or_initsolver(const int ng[3], const double dli[3], const double *dzci, const double *dzfi,
                   const char cbc[6], const double bc[6], double *lambdaxy, double *a, double *b, double *c,
                   double *normfft, double *rhsbx, double *rhsby, double *rhsbz) {
  const int n1 = ng[0], n2 = ng[1], n3 = ng[2];
  double *lx = (double *)malloc(sizeof(double) * n1), *ly = (double *)malloc(sizeof(double) * n2);
  eigenvalues_c(n1, cbc[0], cbc[1], lx);
  for (int i = 0; i < n1; i++) lx[i] = lx[i] * (dli[0] * dli[0]);
  eigenvalues_c(n2, cbc[2], cbc[3], ly);
  for (int j = 0; j < n2; j++) ly[j] = ly[j] * (dli[1] * dli[1]);
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n1; i++) lambdaxy[i + (long)n1 * j] = lx[i] + ly[j];
  free(lx);
  free(ly);
  /* tridmatrix, cell-centred, src/initsolver.f90:153-182 */
  for (int k = 1; k <= n3; k++) {
    a[k - 1] = dzfi[k] * dzci[k - 1];
    c[k - 1] = dzfi[k] * dzci[k];
  }
  for (int k = 0; k < n3; k++) b[k] = -(a[k] + c[k]);
  double f0 = 0.0, f1 = 0.0;
  if (cbc[4] == 'D') f0 = -1.0; else if (cbc[4] == 'N') f0 = 1.0;
  if (cbc[5] == 'D') f1 = -1.0; else if (cbc[5] == 'N') f1 = 1.0;
  b[0] = b[0] + f0 * a[0];
  b[n3 - 1] = b[n3 - 1] + f1 * c[n3 - 1];
  /* bc_rhs, src/initsolver.f90:65-74 and :185-228 */
  const double dl0 = 1.0 / dli[0], dl1 = 1.0 / dli[1];
  const double dzc0 = 1.0 / dzci[0], dzcn = 1.0 / dzci[n3];
  const double dzf1 = 1.0 / dzfi[1], dzfn = 1.0 / dzfi[n3];
  for (int ib = 0; ib < 2; ib++) {
    const double fx = bc_factor_c(cbc[0 + ib], ib, bc[0 + ib], dl0) / dl0 / dl0;
    const double fy = bc_factor_c(cbc[2 + ib], ib, bc[2 + ib], dl1) / dl1 / dl1;
    const double dlc = ib == 0 ? dzc0 : dzcn, dlf = ib == 0 ? dzf1 : dzfn;
    const double fz = bc_factor_c(cbc[4 + ib], ib, bc[4 + ib], dlc) / dlc / dlf;
    if (rhsbx) for (long q = 0; q < (long)n2 * n3; q++) rhsbx[q + (long)n2 * n3 * ib] = fx;
    if (rhsby) for (long q = 0; q < (long)n1 * n3; q++) rhsby[q + (long)n1 * n3 * ib] = fy;
    if (rhsbz) for (long q = 0; q < (long)n1 * n2; q++) rhsbz[q + (long)n1 * n2 * ib] = fz;
  }
  /* normalisation, src/fft.f90:59-139 */
  int kf, kb;
  double nx, ny;
  find_fft_c(cbc[0], cbc[1], &kf, &kb, &nx);
  find_fft_c(cbc[2], cbc[3], &kf, &kb, &ny);
  double nf = 1.0;
  nf = nf * nx * (double)(n1 + 0);
  nf = nf * ny * (double)(n2 + 0);
  *normfft = 1.0 / nf;
}

/* dgtsv_homebrewed, src/solver.f90:151-177; d is scratch of length n */
static void thomas(int n, const double *a, const double *bb, const double *c, double *p, long ps, double *d) {
  double z = 1.0 / (bb[0] + EPS);
  d[0] = c[0] * z;
  p[0] = p[0] * z;
  for (int l = 1; l < n; l++) {
    z = 1.0 / (bb[l] - a[l] * d[l - 1] + EPS);
    d[l] = c[l] * z;
    p[l * ps] = (p[l * ps] - a[l] * p[(l - 1) * ps]) * z;
  }
  for (int l = n - 2; l >= 0; l--) p[l * ps] = p[l * ps] - d[l] * p[(l + 1) * ps];
}

void or_solver(const int n[3], const char cbc[6], double normfft, const double *lambdaxy,
               const double *a, const double *b, const double *c, double *p) {
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const long N = (long)n1 * n2 * n3;
  double *W = (double *)malloc(sizeof(double) * N);
  int kfx, kbx, kfy, kby;
  double nrm;
  find_fft_c(cbc[0], cbc[1], &kfx, &kbx, &nrm);
  find_fft_c(cbc[2], cbc[3], &kfy, &kby, &nrm);
  rplan *pfx = rplan_new(kfx, n1), *pbx = rplan_new(kbx, n1);
  rplan *pfy = rplan_new(kfy, n2), *pby = rplan_new(kby, n2);
  const int periodic_z = (cbc[4] == 'P' && cbc[5] == 'P');
  /* px = p(1:n,...), src/solver.f90:37 */
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) W[IX(n, i, j, k)] = p[HX(n, i, j, k)];
#pragma omp parallel
  {
    const int mmax = 2 * ((n1 > n2 ? n1 : n2) > n3 ? (n1 > n2 ? n1 : n2) : n3) + 2;
    cplx *w = (cplx *)malloc(sizeof(cplx) * 4 * mmax);
    double *ya = (double *)malloc(sizeof(double) * (n2 + 1)), *yb = (double *)malloc(sizeof(double) * (n2 + 1));
    double *bb = (double *)malloc(sizeof(double) * (n3 + 1));
    double *d = (double *)malloc(sizeof(double) * (n3 + 1));
    double *p1 = (double *)malloc(sizeof(double) * (n3 + 1));
    double *p2 = (double *)malloc(sizeof(double) * (n3 + 1));
    /* forward x, src/solver.f90:47: lines (j,k), packed in pairs of consecutive j */
#pragma omp for schedule(static)
    for (int k = 0; k < n3; k++) r2r_lines(pfx, n2, W + (long)n1 * n2 * k, n1, w);
    /* forward y, src/solver.f90:50: lines (i,k), packed in pairs of consecutive i */
#pragma omp for schedule(static)
    for (int k = 0; k < n3; k++)
      for (int i = 0; i < n1; i += (is_quarter(kfy) ? 1 : 2)) {
        double *base = W + i + (long)n1 * n2 * k;
        const int two = !is_quarter(kfy) && (i + 1 < n1);
        for (int j = 0; j < n2; j++) { ya[j] = base[(long)n1 * j]; if (two) yb[j] = base[1 + (long)n1 * j]; }
        if (is_quarter(kfy)) r2r_quarter(pfy, ya, w); else r2r_pair(pfy, ya, two ? yb : NULL, w);
        for (int j = 0; j < n2; j++) { base[(long)n1 * j] = ya[j]; if (two) base[1 + (long)n1 * j] = yb[j]; }
      }
    /* z: gaussel / gaussel_periodic, src/solver.f90:53-59, :80-149 */
#pragma omp for schedule(static)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        const double lam = lambdaxy[i + (long)n1 * j];
        double *col = W + i + (long)n1 * j;
        const long ps = (long)n1 * n2;
        for (int l = 0; l < n3; l++) bb[l] = b[l] + lam;
        if (!periodic_z) {
          thomas(n3, a, bb, c, col, ps, d);
        } else {
          const int m = n3 - 1;
          for (int l = 0; l < m; l++) p1[l] = col[l * ps];
          thomas(m, a, bb, c, p1, 1, d);
          for (int l = 0; l < n3; l++) p2[l] = 0.0;
          p2[0] = -a[0];
          p2[m - 1] = -c[m - 1];
          thomas(m, a, bb, c, p2, 1, d);
          const double pn = (col[m * ps] - c[m] * p1[0] - a[m] * p1[m - 1]) /
                            (bb[m] + c[m] * p2[0] + a[m] * p2[m - 1] + EPS);
          col[m * ps] = pn;
          for (int l = 0; l < m; l++) col[l * ps] = p1[l] + p2[l] * pn;
        }
      }
    /* backward y and x, src/solver.f90:62-65 */
#pragma omp for schedule(static)
    for (int k = 0; k < n3; k++)
      for (int i = 0; i < n1; i += (is_quarter(kby) ? 1 : 2)) {
        double *base = W + i + (long)n1 * n2 * k;
        const int two = !is_quarter(kby) && (i + 1 < n1);
        for (int j = 0; j < n2; j++) { ya[j] = base[(long)n1 * j]; if (two) yb[j] = base[1 + (long)n1 * j]; }
        if (is_quarter(kby)) r2r_quarter(pby, ya, w); else r2r_pair(pby, ya, two ? yb : NULL, w);
        for (int j = 0; j < n2; j++) { base[(long)n1 * j] = ya[j]; if (two) base[1 + (long)n1 * j] = yb[j]; }
      }
#pragma omp for schedule(static)
    for (int k = 0; k < n3; k++) r2r_lines(pbx, n2, W + (long)n1 * n2 * k, n1, w);
    free(w); free(ya); free(yb); free(bb); free(d); free(p1); free(p2);
  }
  /* p = px*normfft, src/solver.f90:68 */
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) p[HX(n, i, j, k)] = W[IX(n, i, j, k)] * normfft;
  rplan_free(pfx); rplan_free(pbx); rplan_free(pfy); rplan_free(pby);
  free(W);
}

/* chk_poisson residual, src/debug.f90:81-114 */
double or_chk_poisson(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                      const double *fp, const double *fpp) {
  const double dxi = dli[0], dyi = dli[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
  double diffmax = 0.0;
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double val = (fpp[q + 1] - 2.0 * fpp[q] + fpp[q - 1]) * (dxi * dxi) +
                           (fpp[q + sy] - 2.0 * fpp[q] + fpp[q - sy]) * (dyi * dyi) +
                           ((fpp[q + sz] - fpp[q]) * dzci[k] - (fpp[q] - fpp[q - sz]) * dzci[k - 1]) * dzfi[k];
        diffmax = dmax(diffmax, fabs(val - fp[q]));
      }
  return diffmax;
}

/* ======================================================================
 * Projection: fillps, correc, updatep, updt_rhs_b
 * ====================================================================== */
void or_fillps(const int n[3], const double dli[3], const double *dzfi, double dti, double rho0,
               const double *u, const double *v, const double *w, double *p) {
  /* src/fillps.f90:33-50 */
  const double dtidxi = dti * dli[0], dtidyi = dti * dli[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        p[q] = rho0 * ((w[q] - w[q - sz]) * dti * dzfi[k] + (v[q] - v[q - sy]) * dtidyi + (u[q] - u[q - 1]) * dtidxi);
      }
}

void or_correc(const int n[3], const double dli[3], const double *dzci, double rho0, double dt,
               const double *p, double *u, double *v, double *w) {
  /* src/correc.f90:29-66, constant-coefficient branch (rhop = rho0) */
  const double factori = dt * dli[0], factorj = dt * dli[1], rhop = rho0;
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= n[2] + 1; k++) {
    for (int j = 0; j <= n[1] + 1; j++)
      for (int i = 0; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        u[q] = u[q] - factori * (p[q + 1] - p[q]) / rhop;
      }
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0] + 1; i++) {
        const long q = HX(n, i, j, k);
        v[q] = v[q] - factorj * (p[q + sy] - p[q]) / rhop;
      }
    if (k <= n[2])
      for (int j = 0; j <= n[1] + 1; j++)
        for (int i = 0; i <= n[0] + 1; i++) {
          const long q = HX(n, i, j, k);
          w[q] = w[q] - dt * dzci[k] * (p[q + sz] - p[q]) / rhop;
        }
  }
}

void or_updatep(const int n[3], const double *pp, double *p) {
  /* src/updatep.f90:21-23 (whole array, halos included) */
  const long N = (long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < N; q++) p[q] = p[q] + pp[q];
}

/* src/bound.f90:341-396; single device: x is always a boundary direction,
 * y and z are boundaries when not periodic in the pressure BCs */
void or_updt_rhs_b(const int n[3], const char cbc[6], const double *rhsbx, const double *rhsby,
                   const double *rhsbz, double *p) {
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const int bnd_y = !(cbc[2] == 'P' && cbc[3] == 'P'), bnd_z = !(cbc[4] == 'P' && cbc[5] == 'P');
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++) {
      p[HX(n, 1, j, k)] += rhsbx[(j - 1) + (long)n2 * (k - 1)];
      p[HX(n, n1, j, k)] += rhsbx[(j - 1) + (long)n2 * (k - 1) + (long)n2 * n3];
    }
  if (bnd_y)
    for (int k = 1; k <= n3; k++)
      for (int i = 1; i <= n1; i++) {
        p[HX(n, i, 1, k)] += rhsby[(i - 1) + (long)n1 * (k - 1)];
        p[HX(n, i, n2, k)] += rhsby[(i - 1) + (long)n1 * (k - 1) + (long)n1 * n3];
      }
  if (bnd_z)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++) {
        p[HX(n, i, j, 1)] += rhsbz[(i - 1) + (long)n1 * (j - 1)];
        p[HX(n, i, j, n3)] += rhsbz[(i - 1) + (long)n1 * (j - 1) + (long)n1 * n2];
      }
}

/* ======================================================================
 * Boundary conditions, src/bound.f90:15-293.
 * With the whole domain on one rank the reference first fills the periodic
 * halos of the non-pencil directions y and z (updthalo with the periods of
 * cbcpre, src/initmpi.f90:48-49 and src/bound.f90:398-475), then calls
 * set_bc for x (always a boundary of the x-pencil) and for y and z when they
 * are not periodic; each operation spans full planes including halos.
 * ====================================================================== */
static void plane_copy(const int *n, double *p, int dir, int dst, int src) {
  const int N[3] = {n[0] + 2, n[1] + 2, n[2] + 2};
  const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  for (int b = 0; b < N[d2]; b++)
    for (int a = 0; a < N[d1]; a++) {
      int s[3], t[3];
      s[d1] = a; s[d2] = b; s[dir] = src;
      t[d1] = a; t[d2] = b; t[dir] = dst;
      p[HX(n, t[0], t[1], t[2])] = p[HX(n, s[0], s[1], s[2])];
    }
}
static void plane_affine(const int *n, double *p, int dir, int dst, int src, double factor, double sgn, int use_src) {
  const int N[3] = {n[0] + 2, n[1] + 2, n[2] + 2};
  const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  for (int b = 0; b < N[d2]; b++)
    for (int a = 0; a < N[d1]; a++) {
      int s[3], t[3];
      s[d1] = a; s[d2] = b; s[dir] = src;
      t[d1] = a; t[d2] = b; t[dir] = dst;
      p[HX(n, t[0], t[1], t[2])] = use_src ? factor + sgn * p[HX(n, s[0], s[1], s[2])] : factor;
    }
}

/* set_bc, src/bound.f90:126-293, nh = 1 */
static void set_bc(const int *n, char ctype, int ibound, int idir, int centered, double rvalue, double dr, double *p) {
  const int nn = n[idir];
  double factor = rvalue, sgn = 0.0;
  if (ctype == 'D' && centered) { factor = 2.0 * factor; sgn = -1.0; }
  if (ctype == 'N') {
    if (ibound == 0) factor = -dr * factor; else factor = dr * factor;
    sgn = 1.0;
  }
  if (ctype == 'P') {
    plane_copy(n, p, idir, 0, nn);
    plane_copy(n, p, idir, nn + 1, 1);
    return;
  }
  if (centered) {
    if (ibound == 0) plane_affine(n, p, idir, 0, 1, factor, sgn, 1);
    else plane_affine(n, p, idir, nn + 1, nn, factor, sgn, 1);
  } else if (ctype == 'D') {
    if (ibound == 0) plane_affine(n, p, idir, 0, 0, factor, 0.0, 0);
    else { plane_copy(n, p, idir, nn + 1, nn - 1); plane_affine(n, p, idir, nn, nn, factor, 0.0, 0); }
  } else { /* non-centred 'N' */
    if (ibound == 0) plane_affine(n, p, idir, 0, 1, factor, 1.0, 1);
    else { plane_copy(n, p, idir, nn + 1, nn); plane_affine(n, p, idir, nn, nn - 1, factor, 1.0, 1); }
  }
}

void or_boundp(const int n[3], const char cbcpre[6], const char cbc[6], const double bc[6],
               const double dl[3], const double *dzc, double *p) {
  const int per_y = cbcpre[2] == 'P' && cbcpre[3] == 'P', per_z = cbcpre[4] == 'P' && cbcpre[5] == 'P';
  if (per_y) { plane_copy(n, p, 1, n[1] + 1, 1); plane_copy(n, p, 1, 0, n[1]); }
  if (per_z) { plane_copy(n, p, 2, n[2] + 1, 1); plane_copy(n, p, 2, 0, n[2]); }
  set_bc(n, cbc[0], 0, 0, 1, bc[0], dl[0], p);
  set_bc(n, cbc[1], 1, 0, 1, bc[1], dl[0], p);
  if (!per_y) {
    set_bc(n, cbc[2], 0, 1, 1, bc[2], dl[1], p);
    set_bc(n, cbc[3], 1, 1, 1, bc[3], dl[1], p);
  }
  if (!per_z) {
    set_bc(n, cbc[4], 0, 2, 1, bc[4], dzc[0], p);
    set_bc(n, cbc[5], 1, 2, 1, bc[5], dzc[n[2]], p);
  }
}

void or_bounduvw(const int n[3], const char cbcpre[6], const char cbc[18], const double bc[18], int is_correc,
                 const double dl[3], const double *dzc, const double *dzf, double *u, double *v, double *w) {
  const int per_y = cbcpre[2] == 'P' && cbcpre[3] == 'P', per_z = cbcpre[4] == 'P' && cbcpre[5] == 'P';
  double *f[3] = {u, v, w};
#define CB(b, d, c) cbc[(b) + 2 * ((d) + 3 * (c))]
#define BV(b, d, c) bc[(b) + 2 * ((d) + 3 * (c))]
  if (per_y)
    for (int c = 0; c < 3; c++) { plane_copy(n, f[c], 1, n[1] + 1, 1); plane_copy(n, f[c], 1, 0, n[1]); }
  if (per_z)
    for (int c = 0; c < 3; c++) { plane_copy(n, f[c], 2, n[2] + 1, 1); plane_copy(n, f[c], 2, 0, n[2]); }
  /* x */
  int impose = (!is_correc) || (CB(0, 0, 0) == 'P' && CB(1, 0, 0) == 'P');
  for (int ib = 0; ib < 2; ib++) {
    if (impose) set_bc(n, CB(ib, 0, 0), ib, 0, 0, BV(ib, 0, 0), dl[0], u);
    set_bc(n, CB(ib, 0, 1), ib, 0, 1, BV(ib, 0, 1), dl[0], v);
    set_bc(n, CB(ib, 0, 2), ib, 0, 1, BV(ib, 0, 2), dl[0], w);
  }
  impose = (!is_correc) || (CB(0, 1, 1) == 'P' && CB(1, 1, 1) == 'P');
  if (!per_y)
    for (int ib = 0; ib < 2; ib++) {
      set_bc(n, CB(ib, 1, 0), ib, 1, 1, BV(ib, 1, 0), dl[1], u);
      if (impose) set_bc(n, CB(ib, 1, 1), ib, 1, 0, BV(ib, 1, 1), dl[1], v);
      set_bc(n, CB(ib, 1, 2), ib, 1, 1, BV(ib, 1, 2), dl[1], w);
    }
  impose = (!is_correc) || (CB(0, 2, 2) == 'P' && CB(1, 2, 2) == 'P');
  if (!per_z)
    for (int ib = 0; ib < 2; ib++) {
      const double dc = ib == 0 ? dzc[0] : dzc[n[2]], df = ib == 0 ? dzf[0] : dzf[n[2]];
      set_bc(n, CB(ib, 2, 0), ib, 2, 1, BV(ib, 2, 0), dc, u);
      set_bc(n, CB(ib, 2, 1), ib, 2, 1, BV(ib, 2, 1), dc, v);
      if (impose) set_bc(n, CB(ib, 2, 2), ib, 2, 0, BV(ib, 2, 2), df, w);
    }
#undef CB
#undef BV
}

/* ======================================================================
 * Phase field: ACDI, src/acdi.f90:16-248 (SDF-normal sharpening)
 * ====================================================================== */
void or_acdi_transport_pf(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                          double gam, double seps, const double *u, const double *v, const double *w,
                          const double *normx, const double *normy, const double *normz, const double *phi,
                          const double *psi, double *dpsidt, double *flux_x, double *flux_y, double *flux_z) {
  const double dxi = dli[0], dyi = dli[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double dzfi_c = dzfi[k], dzci_c = dzci[k], dzci_m = dzci[k - 1];
        const double psi_ccm = psi[q - sz], psi_cmc = psi[q - sy], psi_mcc = psi[q - 1], psi_ccc = psi[q],
                     psi_pcc = psi[q + 1], psi_cpc = psi[q + sy], psi_ccp = psi[q + sz];
        const double phi_ccm = phi[q - sz], phi_cmc = phi[q - sy], phi_mcc = phi[q - 1], phi_ccc = phi[q],
                     phi_pcc = phi[q + 1], phi_cpc = phi[q + sy], phi_ccp = phi[q + sz];
        const double u_mcc = u[q - 1], u_ccc = u[q], v_cmc = v[q - sy], v_ccc = v[q], w_ccm = w[q - sz], w_ccc = w[q];
        const double nx_xm = 0.5 * (normx[q - 1] + normx[q]), nx_xp = 0.5 * (normx[q + 1] + normx[q]);
        const double nx_ym = 0.5 * (normx[q - sy] + normx[q]), nx_yp = 0.5 * (normx[q + sy] + normx[q]);
        const double nx_zm = 0.5 * (normx[q - sz] + normx[q]), nx_zp = 0.5 * (normx[q + sz] + normx[q]);
        const double ny_xm = 0.5 * (normy[q - 1] + normy[q]), ny_xp = 0.5 * (normy[q + 1] + normy[q]);
        const double ny_ym = 0.5 * (normy[q - sy] + normy[q]), ny_yp = 0.5 * (normy[q + sy] + normy[q]);
        const double ny_zm = 0.5 * (normy[q - sz] + normy[q]), ny_zp = 0.5 * (normy[q + sz] + normy[q]);
        const double nz_xm = 0.5 * (normz[q - 1] + normz[q]), nz_xp = 0.5 * (normz[q + 1] + normz[q]);
        const double nz_ym = 0.5 * (normz[q - sy] + normz[q]), nz_yp = 0.5 * (normz[q + sy] + normz[q]);
        const double nz_zm = 0.5 * (normz[q - sz] + normz[q]), nz_zp = 0.5 * (normz[q + sz] + normz[q]);
        const double upsiim = 0.5 * (psi_mcc + psi_ccc) * u_mcc, upsiip = 0.5 * (psi_pcc + psi_ccc) * u_ccc;
        const double vpsijm = 0.5 * (psi_cmc + psi_ccc) * v_cmc, vpsijp = 0.5 * (psi_cpc + psi_ccc) * v_ccc;
        const double wpsikm = 0.5 * (psi_ccm + psi_ccc) * w_ccm, wpsikp = 0.5 * (psi_ccp + psi_ccc) * w_ccc;
        const double adv = dxi * (upsiip - upsiim) + dyi * (vpsijp - vpsijm) + dzfi_c * (wpsikp - wpsikm);
        const double gs = gam * seps;
        const double dpsidxm = gs * (psi_ccc - psi_mcc) * dxi, dpsidxp = gs * (psi_pcc - psi_ccc) * dxi;
        const double dpsidym = gs * (psi_ccc - psi_cmc) * dyi, dpsidyp = gs * (psi_cpc - psi_ccc) * dyi;
        const double dpsidzm = gs * (psi_ccc - psi_ccm) * (PLANTED(PL_ACDI_DZCI_M) ? dzci_c : dzci_m),
                     dpsidzp = gs * (psi_ccp - psi_ccc) * dzci_c;
        const double rn_01 = nx_xm / (sqrt(nx_xm * nx_xm + ny_xm * ny_xm + nz_xm * nz_xm) + EPS);
        const double rn_11 = (PLANTED(PL_ACDI_NORMAL_COMP) ? ny_xp : nx_xp) /
                             (sqrt(nx_xp * nx_xp + ny_xp * ny_xp + nz_xp * nz_xp) + EPS);
        const double rn_02 = ny_ym / (sqrt(nx_ym * nx_ym + ny_ym * ny_ym + nz_ym * nz_ym) + EPS);
        const double rn_12 = ny_yp / (sqrt(nx_yp * nx_yp + ny_yp * ny_yp + nz_yp * nz_yp) + EPS);
        const double rn_03 = nz_zm / (sqrt(nx_zm * nx_zm + ny_zm * ny_zm + nz_zm * nz_zm) + EPS);
        const double rn_13 = nz_zp / (sqrt(nx_zp * nx_zp + ny_zp * ny_zp + nz_zp * nz_zp) + EPS);
        const double g4 = (PLANTED(PL_ACDI_QUARTER) ? 0.5 : 0.25) * gam;
        const double ta = PLANTED(PL_ACDI_TANH_ARG) ? 0.5 : 0.25;
        double t;
        t = om_tanh(ta * (phi_ccc + phi_mcc) / seps); const double sharpxm = g4 * ((1.0 - t * t) * rn_01);
        t = om_tanh(ta * (phi_pcc + phi_ccc) / seps); const double sharpxp = g4 * ((1.0 - t * t) * rn_11);
        t = om_tanh(ta * (phi_ccc + phi_cmc) / seps); const double sharpym = g4 * ((1.0 - t * t) * rn_02);
        t = om_tanh(ta * (phi_cpc + phi_ccc) / seps); const double sharpyp = g4 * ((1.0 - t * t) * rn_12);
        t = om_tanh(ta * (phi_ccc + phi_ccm) / seps); const double sharpzm = g4 * ((1.0 - t * t) * rn_03);
        t = om_tanh(ta * (phi_ccp + phi_ccc) / seps); const double sharpzp = g4 * ((1.0 - t * t) * rn_13);
        const double sg = PLANTED(PL_ACDI_SHARP_SIGN) ? -1.0 : 1.0;
        const double rglrxm = dpsidxm - sg * sharpxm, rglrxp = dpsidxp - sg * sharpxp;
        const double rglrym = dpsidym - sg * sharpym, rglryp = dpsidyp - sg * sharpyp;
        const double rglrzm = dpsidzm - sg * sharpzm, rglrzp = dpsidzp - sg * sharpzp;
        const double rglr = (rglrxp - rglrxm) * dxi + (rglryp - rglrym) * dyi + (rglrzp - rglrzm) * dzfi_c;
        dpsidt[IX(n, i, j, k)] = -adv + rglr;
        flux_x[q] = upsiip - rglrxp;
        flux_y[q] = vpsijp - rglryp;
        flux_z[q] = wpsikp - rglrzp;
      }
}

void or_acdi_cmpt_phi(const int n[3], double seps, const double *psi, double *phi) {
  /* src/acdi.f90:239-247, halo-inclusive */
  const long N = (long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < N; q++) phi[q] = seps * om_log((psi[q] + EPS) / (1.0 - psi[q] + EPS));
}

double or_acdi_set_gamma(const int n[3], double gam_factor, const double *u, const double *v, const double *w) {
  /* src/acdi.f90:44-61 */
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
  double velmax = 0.0;
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double uc = 0.5 * (u[q] + u[q - 1]), vc = 0.5 * (v[q] + v[q - sy]), wc = 0.5 * (w[q] + w[q - sz]);
        velmax = dmax(velmax, uc * uc + vc * vc + wc * wc);
      }
  return sqrt(velmax) * gam_factor;
}

double or_acdi_set_epsilon(const double dl[3], const double *dzfi, int n3, double seps_factor) {
  /* src/acdi.f90:27-29; maxval(1./dzfi) spans 0..n3+1 */
  double dlmax = dmax(dl[0], dl[1]);
  double m = 1.0 / dzfi[0];
  for (int k = 1; k <= n3 + 1; k++) m = dmax(m, 1.0 / dzfi[k]);
  dlmax = dmax(dlmax, m);
  return dlmax * seps_factor;
}

/* THINC/QQ midpoint transport, src/vof_thinc_qq.f90:66-132 (SDF normals) */
void or_vof_thinc_transport_psi(const int n[3], const double dli[3], const double *dzfi, double beta,
                                const double *u, const double *v, const double *w, const double *normx,
                                const double *normy, const double *normz, const double *phi,
                                double *dpsidt, double *flux_x, double *flux_y, double *flux_z) {
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= n[2]; k++)
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double vel_x = u[q], vel_y = v[q], vel_z = w[q];
        long qx, qy, qz;
        double xx, yy, zz;
        if (vel_x > 0.0) {
          qx = PLANTED(PL_THINC_UPWIND) ? q + 1 : q;
          xx = PLANTED(PL_THINC_OFFSET) ? -0.5 : 0.5;
        } else {
          qx = q + 1;
          xx = -0.5;
        }
        if (vel_y > 0.0) { qy = q; yy = 0.5; } else { qy = q + sy; yy = -0.5; }
        if (vel_z > 0.0) { qz = q; zz = 0.5; } else { qz = q + sz; zz = -0.5; }
        const double lnx = PLANTED(PL_THINC_NORMAL_COMP) ? normy[qx] : normx[qx], lny = normy[qy], lnz = normz[qz];
        const double d_x = phi[qx], d_y = phi[qy], d_z = phi[qz];
        const double lfx = 1.0 / (1.0 + om_exp(-2.0 * beta * (lnx * xx + d_x)));
        const double lfy = 1.0 / (1.0 + om_exp(-2.0 * beta * (lny * yy + d_y)));
        const double lfz = 1.0 / (1.0 + om_exp(-2.0 * beta * (lnz * zz + d_z)));
        flux_x[q] = vel_x * lfx;
        flux_y[q] = vel_y * lfy;
        flux_z[q] = vel_z * lfz;
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        dpsidt[IX(n, i, j, k)] = -(dli[0] * (flux_x[q] - flux_x[q - 1]) + dli[1] * (flux_y[q] - flux_y[q - sy]) +
                                   dzfi[k] * (flux_z[q] - flux_z[q - sz]));
      }
}

void or_vof_thinc_cmpt_phi(const int n[3], double beta, const double *psi, double *phi) {
  /* src/vof_thinc_qq.f90:148-156 */
  const long N = (long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < N; q++) {
    const double pa = dmin(dmax(psi[q], EPS), 1.0 - EPS);
    phi[q] = 0.5 / beta * om_log(pa / (1.0 - pa));
  }
}

void or_clip_field(const int n[3], double rmin, double rmax, double *p) {
  /* src/two_fluid.f90:54-57, halo-inclusive */
  const long N = (long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < N; q++) p[q] = dmin(dmax(rmin, p[q]), rmax);
}

/* Youngs normals and curvature, src/two_fluid.f90:81-174 */
void or_cmpt_norm_curv_youngs(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                              const double *psi, double *normx, double *normy, double *normz, double *kappa) {
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
#define P_(a, b, c) psi[q + (a) + (b) * sy + (c) * sz]
        const double mmm = P_(-1, -1, -1), mcm = P_(-1, 0, -1), mpm = P_(-1, 1, -1), cmm = P_(0, -1, -1),
                     ccm = P_(0, 0, -1), cpm = P_(0, 1, -1), pmm = P_(1, -1, -1), pcm = P_(1, 0, -1),
                     ppm = P_(1, 1, -1);
        const double mmc = P_(-1, -1, 0), mcc = P_(-1, 0, 0), mpc = P_(-1, 1, 0), cmc = P_(0, -1, 0),
                     ccc = P_(0, 0, 0), cpc = P_(0, 1, 0), pmc = P_(1, -1, 0), pcc = P_(1, 0, 0),
                     ppc = P_(1, 1, 0);
        const double mmp = P_(-1, -1, 1), mcp = P_(-1, 0, 1), mpp = P_(-1, 1, 1), cmp = P_(0, -1, 1),
                     ccp = P_(0, 0, 1), cpp = P_(0, 1, 1), pmp = P_(1, -1, 1), pcp = P_(1, 0, 1),
                     ppp = P_(1, 1, 1);
#undef P_
        const double dxi = dli[0], dyi = dli[1], zc0 = dzci[k], zcm = dzci[k - 1];
        double mx[8], my[8], mz[8];
        mx[0] = 0.25 * ((pcc + ppc + pcp + ppp) - (ccc + cpc + ccp + cpp)) * dxi;
        mx[1] = 0.25 * ((pcc + pmc + pcp + pmp) - (ccc + cmc + ccp + cmp)) * dxi;
        mx[2] = 0.25 * ((pcc + ppc + pcm + ppm) - (ccc + cpc + ccm + cpm)) * dxi;
        mx[3] = 0.25 * ((pcc + pmc + pcm + pmm) - (ccc + cmc + ccm + cmm)) * dxi;
        mx[4] = 0.25 * ((ccc + cpc + ccp + cpp) - (mcc + mpc + mcp + mpp)) * dxi;
        mx[5] = 0.25 * ((ccc + cmc + ccp + cmp) - (mcc + mmc + mcp + mmp)) * dxi;
        mx[6] = 0.25 * ((ccc + cpc + ccm + cpm) - (mcc + mpc + mcm + mpm)) * dxi;
        mx[7] = 0.25 * ((ccc + cmc + ccm + cmm) - (mcc + mmc + mcm + mmm)) * dxi;
        my[0] = 0.25 * ((cpc + ppc + cpp + ppp) - (ccc + pcc + ccp + pcp)) * dyi;
        my[1] = 0.25 * ((ccc + pcc + ccp + pcp) - (cmc + pmc + cmp + pmp)) * dyi;
        my[2] = 0.25 * ((cpc + ppc + cpm + ppm) - (ccc + pcc + ccm + pcm)) * dyi;
        my[3] = 0.25 * ((ccc + pcc + ccm + pcm) - (cmc + pmc + cmm + pmm)) * dyi;
        my[4] = 0.25 * ((cpc + mpc + cpp + mpp) - (ccc + mcc + ccp + mcp)) * dyi;
        my[5] = 0.25 * ((ccc + mcc + ccp + mcp) - (cmc + mmc + cmp + mmp)) * dyi;
        my[6] = 0.25 * ((cpc + mpc + cpm + mpm) - (ccc + mcc + ccm + mcm)) * dyi;
        my[7] = 0.25 * ((ccc + mcc + ccm + mcm) - (cmc + mmc + cmm + mmm)) * dyi;
        mz[0] = 0.25 * ((ccp + pcp + cpp + ppp) - (ccc + pcc + cpc + ppc)) * zc0;
        mz[1] = 0.25 * ((ccp + pcp + cmp + pmp) - (ccc + pcc + cmc + pmc)) * zc0;
        mz[2] = 0.25 * ((ccc + pcc + cpc + ppc) - (ccm + pcm + cpm + ppm)) * zcm;
        mz[3] = 0.25 * ((ccc + pcc + cmc + pmc) - (ccm + pcm + cmm + pmm)) * zcm;
        mz[4] = 0.25 * ((ccp + mcp + cpp + mpp) - (ccc + mcc + cpc + mpc)) * zc0;
        mz[5] = 0.25 * ((ccp + mcp + cmp + mmp) - (ccc + mcc + cmc + mmc)) * zc0;
        mz[6] = 0.25 * ((ccc + mcc + cpc + mpc) - (ccm + mcm + cpm + mpm)) * zcm;
        mz[7] = 0.25 * ((ccc + mcc + cmc + mmc) - (ccm + mcm + cmm + mmm)) * zcm;
        for (int m = 0; m < 8; m++) {
          const double nrm = sqrt(mx[m] * mx[m] + my[m] * my[m] + mz[m] * mz[m]) + EPS;
          mx[m] = mx[m] / nrm; my[m] = my[m] / nrm; mz[m] = mz[m] / nrm;
        }
        const double nx = .125 * (mx[0] + mx[1] + mx[2] + mx[3] + mx[4] + mx[5] + mx[6] + mx[7]);
        const double ny = .125 * (my[0] + my[1] + my[2] + my[3] + my[4] + my[5] + my[6] + my[7]);
        const double nz = .125 * (mz[0] + mz[1] + mz[2] + mz[3] + mz[4] + mz[5] + mz[6] + mz[7]);
        const double nrm = sqrt(nx * nx + ny * ny + nz * nz) + EPS;
        normx[q] = nx / nrm;
        normy[q] = ny / nrm;
        normz[q] = nz / nrm;
        kappa[q] = -(0.25 * ((mx[0] + mx[1] + mx[2] + mx[3]) - (mx[4] + mx[5] + mx[6] + mx[7])) * dxi +
                     0.25 * ((my[0] + my[2] + my[4] + my[6]) - (my[1] + my[3] + my[5] + my[7])) * dyi +
                     0.25 * ((mz[0] + mz[1] + mz[4] + mz[5]) - (mz[2] + mz[3] + mz[6] + mz[7])) * dzfi[k]);
      }
}

/* ======================================================================
 * Momentum: mom_xyz_ad (src/mom.f90:505-851, non-split diffusion) and
 * mom_xyz_oth (src/mom.f90:853-1039, constant-coefficient pressure split,
 * Brackbill normalisation, no scalar buoyancy)
 * ====================================================================== */
void or_mom_xyz_ad(const int n[3], const double dli[3], const double *dzci, const double *dzfi,
                   const double rho12[2], const double mu12[2], const double *psiflx_x,
                   const double *psiflx_y, const double *psiflx_z, const double *u, const double *v,
                   const double *w, const double *psi, double *dudt, double *dvdt, double *dwdt) {
  const double rho = rho12[1], drho = rho12[0] - rho12[1], mu = mu12[1], dmu = mu12[0] - mu12[1];
  const double dxi = dli[0], dyi = dli[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
#define S_(f, a, b, c) f[q + (a) + (b) * sy + (c) * sz]
        const double u_ccm = S_(u, 0, 0, -1), u_pcm = S_(u, 1, 0, -1), u_cpm = S_(u, 0, 1, -1), u_cmc = S_(u, 0, -1, 0),
                     u_pmc = S_(u, 1, -1, 0), u_mcc = S_(u, -1, 0, 0), u_ccc = S_(u, 0, 0, 0), u_pcc = S_(u, 1, 0, 0),
                     u_mpc = S_(u, -1, 1, 0), u_cpc = S_(u, 0, 1, 0), u_cmp = S_(u, 0, -1, 1), u_mcp = S_(u, -1, 0, 1),
                     u_ccp = S_(u, 0, 0, 1);
        const double v_ccm = S_(v, 0, 0, -1), v_pcm = S_(v, 1, 0, -1), v_cpm = S_(v, 0, 1, -1), v_cmc = S_(v, 0, -1, 0),
                     v_pmc = S_(v, 1, -1, 0), v_mcc = S_(v, -1, 0, 0), v_ccc = S_(v, 0, 0, 0), v_pcc = S_(v, 1, 0, 0),
                     v_mpc = S_(v, -1, 1, 0), v_cpc = S_(v, 0, 1, 0), v_cmp = S_(v, 0, -1, 1), v_mcp = S_(v, -1, 0, 1),
                     v_ccp = S_(v, 0, 0, 1);
        const double w_ccm = S_(w, 0, 0, -1), w_pcm = S_(w, 1, 0, -1), w_cpm = S_(w, 0, 1, -1), w_cmc = S_(w, 0, -1, 0),
                     w_pmc = S_(w, 1, -1, 0), w_mcc = S_(w, -1, 0, 0), w_ccc = S_(w, 0, 0, 0), w_pcc = S_(w, 1, 0, 0),
                     w_mpc = S_(w, -1, 1, 0), w_cpc = S_(w, 0, 1, 0), w_cmp = S_(w, 0, -1, 1), w_mcp = S_(w, -1, 0, 1),
                     w_ccp = S_(w, 0, 0, 1);
        const double c_ccm = S_(psi, 0, 0, -1), c_pcm = S_(psi, 1, 0, -1), c_cpm = S_(psi, 0, 1, -1),
                     c_cmc = S_(psi, 0, -1, 0), c_pmc = S_(psi, 1, -1, 0), c_mcc = S_(psi, -1, 0, 0),
                     c_ccc = S_(psi, 0, 0, 0), c_pcc = S_(psi, 1, 0, 0), c_mpc = S_(psi, -1, 1, 0),
                     c_cpc = S_(psi, 0, 1, 0), c_cmp = S_(psi, 0, -1, 1), c_mcp = S_(psi, -1, 0, 1),
                     c_ccp = S_(psi, 0, 0, 1), c_cpp = S_(psi, 0, 1, 1), c_ppc = S_(psi, 1, 1, 0),
                     c_pcp = S_(psi, 1, 0, 1);
        (void)v_mcp; (void)w_mpc; (void)u_pcm; (void)u_cpm; (void)v_pcm; (void)v_cpm; (void)u_pmc; (void)w_pmc;
        (void)w_mcp; (void)u_cmp; (void)v_mpc;
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
        /* advection, src/mom.f90:646-668 */
        const double uuip = 0.25 * (flx_x_pcc + flx_x_ccc) * (u_ccc + u_pcc);
        const double uuim = 0.25 * (flx_x_mcc + flx_x_ccc) * (u_ccc + u_mcc);
        const double vujp = 0.25 * (flx_y_pcc + flx_y_ccc) * (u_ccc + u_cpc);
        const double vujm = 0.25 * (flx_y_pmc + flx_y_cmc) * (u_ccc + u_cmc);
        const double wukp = 0.25 * (flx_z_pcc + flx_z_ccc) * (u_ccc + u_ccp);
        const double wukm = 0.25 * (flx_z_pcm + flx_z_ccm) * (u_ccc + u_ccm);
        double dudt_aux = dxi * (-uuip + uuim) + dyi * (-vujp + vujm) + dzfi_c * (-wukp + wukm);
        const double uvip = 0.25 * (flx_x_ccc + flx_x_cpc) * (v_ccc + v_pcc);
        const double uvim = 0.25 * (flx_x_mcc + flx_x_mpc) * (v_ccc + v_mcc);
        const double vvjp = 0.25 * (flx_y_ccc + flx_y_cpc) * (v_ccc + v_cpc);
        const double vvjm = 0.25 * (flx_y_ccc + flx_y_cmc) * (v_ccc + v_cmc);
        const double wvkp = 0.25 * (flx_z_ccc + flx_z_cpc) * (v_ccc + v_ccp);
        const double wvkm = 0.25 * (flx_z_ccm + flx_z_cpm) * (v_ccc + v_ccm);
        double dvdt_aux = dxi * (-uvip + uvim) + dyi * (-vvjp + vvjm) + dzfi_c * (-wvkp + wvkm);
        const double uwip = 0.25 * (flx_x_ccc + flx_x_ccp) * (w_ccc + w_pcc);
        const double uwim = 0.25 * (flx_x_mcc + flx_x_mcp) * (w_ccc + w_mcc);
        const double vwjp = 0.25 * (flx_y_ccc + flx_y_ccp) * (w_ccc + w_cpc);
        const double vwjm = 0.25 * (flx_y_cmc + flx_y_cmp) * (w_ccc + w_cmc);
        const double wwkp = 0.25 * (flx_z_ccc + flx_z_ccp) * (w_ccc + w_ccp);
        const double wwkm = 0.25 * (flx_z_ccc + flx_z_ccm) * (w_ccc + w_ccm);
        double dwdt_aux = dxi * (-uwip + uwim) + dyi * (-vwjp + vwjm) + dzci_c * (-wwkp + wwkm);
        /* diffusion, src/mom.f90:784-842 */
        {
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
        {
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
        {
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
        const long r = IX(n, i, j, k);
        dudt[r] = dudt_aux;
        dvdt[r] = dvdt_aux;
        dwdt[r] = dwdt_aux;
      }
}

void or_mom_xyz_oth(const int n[3], const double dli[3], const double *dzci, const double *dzfi, double dt_r,
                    const double rho12[2], const double bforce[3], const double gacc[3], double sigma,
                    double rho0, const double rho_av[3], const double *p, const double *psi,
                    const double *kappa, const double *pn, const double *po,
                    double *dudt, double *dvdt, double *dwdt) {
  (void)dzfi;
  const double rho = rho12[1], drho = rho12[0] - rho12[1];
  const double dxi = dli[0], dyi = dli[1];
  const double surf_factor = sigma * (PLANTED(PL_OTH_SURF_FACTOR) ? 1.0 : 2.0) / (rho12[0] + rho12[1]);
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double dzci_c = dzci[k];
        const double p_ccc = p[q], p_pcc = p[q + 1], p_cpc = p[q + sy], p_ccp = p[q + sz];
        const double *pe1 = PLANTED(PL_OTH_EXTRAP) ? po : pn, *pe2 = PLANTED(PL_OTH_EXTRAP) ? pn : po;
        const double q_ccc = (1.0 + dt_r) * pe1[q] - dt_r * pe2[q];
        const double q_pcc = (1.0 + dt_r) * pe1[q + 1] - dt_r * pe2[q + 1];
        const double q_cpc = (1.0 + dt_r) * pe1[q + sy] - dt_r * pe2[q + sy];
        const double q_ccp = (1.0 + dt_r) * pe1[q + sz] - dt_r * pe2[q + sz];
        const double k_ccc = kappa[q], k_pcc = kappa[q + 1], k_cpc = kappa[q + sy], k_ccp = kappa[q + sz];
        const double c_ccc = psi[q], c_pcc = psi[q + 1], c_cpc = psi[q + sy], c_ccp = psi[q + sz];
        const double psixp = 0.5 * (c_pcc + c_ccc), psiyp = 0.5 * (c_cpc + c_ccc), psizp = 0.5 * (c_ccp + c_ccc);
        const double rhoxp = rho + drho * psixp, rhoyp = rho + drho * psiyp, rhozp = rho + drho * psizp;
        const double dpdx = (p_pcc - p_ccc) * dxi, dpdy = (p_cpc - p_ccc) * dyi, dpdz = (p_ccp - p_ccc) * dzci_c;
        double dudt_aux = bforce[0] + (rhoxp - rho_av[0]) * gacc[0];
        double dvdt_aux = bforce[1] + (rhoyp - rho_av[1]) * gacc[1];
        double dwdt_aux = bforce[2] + (rhozp - rho_av[2]) * gacc[2];
        double skx = 0.5 * (k_pcc + k_ccc) * surf_factor, sky = 0.5 * (k_cpc + k_ccc) * surf_factor,
               skz = 0.5 * (k_ccp + k_ccc) * surf_factor;
        if (PLANTED(PL_OTH_KAPPA_FACE)) skx = k_ccc * surf_factor;
        if (!PLANTED(PL_OTH_BRACKBILL)) { skx = skx * rhoxp; sky = sky * rhoyp; skz = skz * rhozp; }
        const double surfx = skx * (c_pcc - c_ccc) * dxi, surfy = sky * (c_cpc - c_ccc) * dyi,
                     surfz = skz * (c_ccp - c_ccc) * dzci_c;
        dudt_aux = dudt_aux + surfx;
        dvdt_aux = dvdt_aux + surfy;
        dwdt_aux = dwdt_aux + surfz;
        const double dpdx_e = (q_pcc - q_ccc) * dxi, dpdy_e = (q_cpc - q_ccc) * dyi, dpdz_e = (q_ccp - q_ccc) * dzci_c;
        dudt_aux = dudt_aux - (dpdx_e + (PLANTED(PL_OTH_SPLIT) ? rho0 / rhoxp : rhoxp / rho0) * (dpdx - dpdx_e));
        dvdt_aux = dvdt_aux - (dpdy_e + rhoyp / rho0 * (dpdy - dpdy_e));
        dwdt_aux = dwdt_aux - (dpdz_e + rhozp / rho0 * (dpdz - dpdz_e));
        const long r = IX(n, i, j, k);
        dudt[r] = dudt_aux;
        dvdt[r] = dvdt_aux;
        dwdt[r] = dwdt_aux;
      }
}

/* AB2 momentum update with the density rescale, src/rk.f90:85-101 */
void or_rk_mom_ab2(const int n[3], double factor1, double factor2, const double rho12[2], const double *psio,
                   const double *psi, const double *dudt, const double *dvdt, const double *dwdt,
                   const double *dudto, const double *dvdto, const double *dwdto,
                   double *u, double *v, double *w) {
  const double rho = rho12[1], drho = rho12[0] - rho12[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k), r = IX(n, i, j, k);
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
}

/* pressure/surface-tension/gravity increment, src/rk.f90:118-130 */
void or_rk_mom_oth(const int n[3], double factor12, const double rho12[2], const double *psi,
                   const double *dudt, const double *dvdt, const double *dwdt, double *u, double *v, double *w) {
  const double rho = rho12[1], drho = rho12[0] - rho12[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k), r = IX(n, i, j, k);
        const double rhox_p = rho + drho * 0.5 * (psi[q] + psi[q + 1]);
        const double rhoy_p = rho + drho * 0.5 * (psi[q] + psi[q + sy]);
        const double rhoz_p = rho + drho * 0.5 * (psi[q] + psi[q + sz]);
        u[q] = u[q] + factor12 * dudt[r] / rhox_p;
        v[q] = v[q] + factor12 * dvdt[r] / rhoy_p;
        w[q] = w[q] + factor12 * dwdt[r] / rhoz_p;
      }
}

/* phase-field AB2 update, src/rk.f90:240-247 */
void or_rk_psi_ab2(const int n[3], double factor1, double factor2, const double *dpsidt,
                   const double *dpsidto, double *psi) {
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k), r = IX(n, i, j, k);
        psi[q] = psi[q] + factor1 * dpsidt[r] + factor2 * dpsidto[r];
      }
}

/* Taylor-Green large-scale forcing, src/forcing.f90:48-63 (k0=2) */
void or_lscale_forcing_tg(const int n[3], double alpha, double dt, const double l[3], const double dl[3],
                          const double *zc, const double *zf, double *u, double *v) {
  const double k0 = 2.0;
  const double f0 = alpha * ((2.0 * PI / l[2]) * (2.0 * PI / l[2]));
  const double factor = dt;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double xxc = (double)((float)i - 0.5f) * dl[0] / l[0];
        const double xxf = xxc + 0.5 * dl[0] / l[0];
        const double yyc = (double)((float)j - 0.5f) * dl[1] / l[1];
        const double yyf = yyc + 0.5 * dl[1] / l[1];
        const double zzc = zc[k] / l[2];
        (void)zf;
        u[q] = u[q] + factor * (f0 * om_sin(k0 * xxf) * om_cos(k0 * yyc) * om_cos(k0 * zzc));
        v[q] = v[q] - factor * (f0 * om_cos(k0 * xxc) * om_sin(k0 * yyf) * om_cos(k0 * zzc));
      }
}

/* bulk_mean_12_stag, src/utils.f90:68-136; summed per k-plane then over k
 * so that the result does not depend on the thread count */
double or_bulk_mean_12_stag(const int n[3], int idir, const double *grid_vol_ratio, const double *psi,
                            const double p12[2]) {
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
  const long off = idir == 1 ? 1 : (idir == 2 ? sy : sz);
  double *part = (double *)calloc(n[2] + 1, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++) {
    double s = 0.0;
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double pl = 0.5 * (psi[q + off] + psi[q]);
        s = s + (pl * p12[0] + (1.0 - pl) * p12[1]) * grid_vol_ratio[k];
      }
    part[k] = s;
  }
  double mean = 0.0;
  for (int k = 1; k <= n[2]; k++) mean += part[k];
  free(part);
  return mean;
}

/* ======================================================================
 * Checks, src/chkdiv.f90:15-49 and src/chkdt.f90:15-94
 * ====================================================================== */
void or_chkdiv(const int n[3], const double dli[3], const double *dzfi, const double *u, const double *v,
               const double *w, double out[2]) {
  const double dxi = dli[0], dyi = dli[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
  double *part = (double *)calloc(n[2] + 1, sizeof(double)), *pmax = (double *)calloc(n[2] + 1, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= n[2]; k++) {
    double s = 0.0, m = 0.0;
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
        const double div = (w[q] - w[q - sz]) * dzfi[k] + (v[q] - v[q - sy]) * dyi + (u[q] - u[q - 1]) * dxi;
        m = dmax(m, fabs(div));
        s = s + div;
      }
    part[k] = s;
    pmax[k] = m;
  }
  double divtot = 0.0, divmax = 0.0;
  for (int k = 1; k <= n[2]; k++) { divtot += part[k]; divmax = dmax(divmax, pmax[k]); }
  free(part);
  free(pmax);
  out[0] = divtot;
  out[1] = divmax;
}

double or_chkdt(const int n[3], const double dl[3], const double *dzci, const double *dzfi, int is_solve_ns,
                int is_track_interface, int vof, const double mu12[2], const double rho12[2], double sigma,
                const double gacc[3], const double *u, const double *v, const double *w, double gam, double seps) {
  const double dxi = 1.0 / dl[0], dyi = 1.0 / dl[1];
  const long sy = n[0] + 2, sz = (long)(n[0] + 2) * (n[1] + 2);
  double dlmin = dmin(dl[0], dl[1]);
  double m = 1.0 / dzfi[0];
  for (int k = 1; k <= n[2] + 1; k++) m = dmin(m, 1.0 / dzfi[k]);
  if (!PLANTED(PL_DT_DLMIN)) dlmin = dmin(dlmin, m);
  double dti = 0.0;
  for (int k = 1; k <= n[2]; k++)
    for (int j = 1; j <= n[1]; j++)
      for (int i = 1; i <= n[0]; i++) {
        const long q = HX(n, i, j, k);
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
        dti = dmax(dti, dmax(dtix, dmax(dtiy, dtiz)));
      }
  double dtipsi;
  if (!vof) {
    dtipsi = gam * seps / (dlmin * dlmin);
    if (dtipsi == 0.0) dtipsi = 1.0;
  } else {
    dtipsi = dti;
  }
  double dtmax;
  if (!is_solve_ns) {
    dtmax = dmin(1.0 / dti, 1.0 / dtipsi);
  } else {
    const double dtiv = dmax(mu12[0] / rho12[0], mu12[1] / rho12[1]) * (PLANTED(PL_DT_VISC) ? 1.0 : 2.0) *
                        (3.0 / (dlmin * dlmin));
    const double rmin = PLANTED(PL_DT_CAP_RHO) ? dmax(rho12[0], rho12[1]) : dmin(rho12[0], rho12[1]);
    const double dtik = sqrt(sigma / (rmin) / (dlmin * dlmin * dlmin));
    const double dtig = dmax(fabs(gacc[0]), dmax(fabs(gacc[1]), fabs(gacc[2]))) /
                        (PLANTED(PL_DT_GRAV) ? dlmin * dlmin : dlmin);
    dti = (PLANTED(PL_DT_COMBINE) ? 1.0 : 0.5) *
          (dti + dtiv + sqrt((dti + dtiv) * (dti + dtiv) + 4.0 * (dtig * dtig + dtik * dtik)));
    if (dti == 0.0) dti = 1.0;
    dtmax = 1.0 / dti;
    if (is_track_interface) dtmax = dmin(dtmax, 1.0 / dtipsi);
  }
  return dtmax;
}

/* ======================================================================
 * Initial phase field, src/two_fluid.f90:247-437 and :503-513
 * ====================================================================== */
static double smooth_step_tanh(double r, double eps) {
  return .5 * (1.0 + om_tanh(r / ((PLANTED(PL_INIT_WIDTH) ? 1.0 : 2.0) * eps)));
}

void or_init2fl(const or_params *prm, double seps, const double *zc, double *psi) {
  const int *n = prm->ng;
  const long N = (long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
  double dl[3];
  for (int d = 0; d < 3; d++) dl[d] = prm->l[d] / (double)(float)n[d];
  for (long q = 0; q < N; q++) psi[q] = 1.0;
  int is_sphere = 0, is_swap = 0, is_dim[3] = {1, 1, 1};
  switch (prm->inipsi) {
  case 0: break;
  case 1: for (long q = 0; q < N; q++) psi[q] = 0.0; break;
  case 2: is_sphere = 1; break;
  case 3: is_sphere = 1; is_dim[1] = 0; break;
  case 4: is_sphere = 1; is_dim[0] = 0; is_dim[1] = 0; break;
  case 5: is_sphere = 1; is_swap = 1; break;
  case 6: is_sphere = 1; is_swap = 1; is_dim[1] = 0; break;
  case 7: is_sphere = 1; is_swap = 1; is_dim[0] = 0; is_dim[1] = 0; break;
  case 8:
    for (int k = 1; k <= n[2]; k++)
      for (int j = 1; j <= n[1]; j++)
        for (int i = 1; i <= n[0]; i++)
          psi[HX(n, i, j, k)] = smooth_step_tanh(zc[PLANTED(PL_INIT_FLM_INDEX) ? k - 1 : k] - prm->l[2] / 2.0, seps);
    break;
  default: break;
  }
  if (is_sphere) {
    int nsph = prm->nspheres;
    double sph_default[4];
    const double *sph = prm->spheres;
    if (nsph == 0) {
      double lmin = 1e300;
      for (int d = 0; d < 3; d++) if (is_dim[d]) lmin = dmin(lmin, prm->l[d]);
      sph_default[0] = prm->l[0] / 2; sph_default[1] = prm->l[1] / 2; sph_default[2] = prm->l[2] / 2;
      sph_default[3] = 0.25 * lmin;
      sph = sph_default;
      nsph = 1;
    }
    int iperiod[3];
    for (int d = 0; d < 3; d++)
      iperiod[d] = (prm->cbcpsi[2 * d] == 'P' && prm->cbcpsi[2 * d + 1] == 'P') && !PLANTED(PL_INIT_PERIODIC);
    double lmax = 0.0;
    for (int d = 0; d < 3; d++) lmax = dmax(lmax, prm->l[d] * is_dim[d]);
    for (int s = 0; s < nsph; s++) {
      const double xxc = sph[4 * s + 0], yyc = sph[4 * s + 1], zzc = sph[4 * s + 2], r = sph[4 * s + 3];
#pragma omp parallel for schedule(static)
      for (int k = 1; k <= n[2]; k++) {
        const double z = zc[k];
        for (int j = 1; j <= n[1]; j++) {
          const double y = (double)((float)j - 0.5f) * dl[1];
          for (int i = 1; i <= n[0]; i++) {
            const double x = (double)((float)i - 0.5f) * dl[0];
            double sdistmin = lmax * 2.0;
            for (int kk = -1; kk <= 1; kk++)
              for (int jj = -1; jj <= 1; jj++)
                for (int ii = -1; ii <= 1; ii++) {
                  const double ax = x + (double)(ii * iperiod[0]) * prm->l[0] - xxc;
                  const double ay = y + (double)(jj * iperiod[1]) * prm->l[1] - yyc;
                  const double az = z + (double)(kk * iperiod[2]) * prm->l[2] - zzc;
                  const double sd = sqrt(ax * ax * is_dim[0] + ay * ay * is_dim[1] + az * az * is_dim[2]) - r;
                  if (fabs(sd) <= sdistmin) sdistmin = sd;
                }
            const long q = HX(n, i, j, k);
            psi[q] = dmin(psi[q], smooth_step_tanh(sdistmin, seps));
          }
        }
      }
    }
  }
  if (is_swap)
    for (long q = 0; q < N; q++) psi[q] = 1.0 - psi[q];
}

/* ======================================================================
 * A whole case: the time loop of src/main.f90:414-562 with the AB2 state
 * of src/rk.f90 held per case instead of in `save` variables.
 * ====================================================================== */
struct or_case {
  or_params prm;
  int n[3];
  double dl[3], dli[3];
  double *dzc, *dzf, *zc, *zf, *dzci, *dzfi, *grid_vol_ratio_c, *grid_vol_ratio_f;
  double *lambdaxy, *a, *b, *c, *rhsbx, *rhsby, *rhsbz, normfft;
  double *f[19];
  double *dpsidt, *dudt, *dvdt, *dwdt, *pp, *psio;
  double sc[12];
  char cbccur[6];
  double bccur[6];
};

or_case *or_case_create(const or_params *prm) {
  or_case *c = (or_case *)calloc(1, sizeof(or_case));
  c->prm = *prm;
  for (int d = 0; d < 3; d++) {
    c->n[d] = prm->ng[d];
    c->dl[d] = prm->l[d] / (double)(float)prm->ng[d];
    c->dli[d] = 1.0 / c->dl[d];
  }
  const int n3 = c->n[2];
  double **g[8] = {&c->dzc, &c->dzf, &c->zc, &c->zf, &c->dzci, &c->dzfi, &c->grid_vol_ratio_c, &c->grid_vol_ratio_f};
  for (int q = 0; q < 8; q++) *g[q] = (double *)calloc(n3 + 2, sizeof(double));
  or_initgrid(prm->gtype, n3, prm->gr, prm->l[2], c->dzc, c->dzf, c->zc, c->zf);
  for (int k = 0; k <= n3 + 1; k++) {
    c->dzci[k] = 1.0 / c->dzc[k];
    c->dzfi[k] = 1.0 / c->dzf[k];
    c->grid_vol_ratio_c[k] = c->dl[0] * c->dl[1] * c->dzc[k] / (prm->l[0] * prm->l[1] * prm->l[2]);
    c->grid_vol_ratio_f[k] = c->dl[0] * c->dl[1] * c->dzf[k] / (prm->l[0] * prm->l[1] * prm->l[2]);
  }
  const long NH = (long)(c->n[0] + 2) * (c->n[1] + 2) * (c->n[2] + 2);
  const long NI = (long)c->n[0] * c->n[1] * c->n[2];
  for (int q = 0; q < 15; q++) c->f[q] = (double *)calloc(NH, sizeof(double));
  for (int q = 15; q < 19; q++) c->f[q] = (double *)calloc(NI, sizeof(double));
  c->dpsidt = (double *)calloc(NI, sizeof(double));
  c->dudt = (double *)calloc(NI, sizeof(double));
  c->dvdt = (double *)calloc(NI, sizeof(double));
  c->dwdt = (double *)calloc(NI, sizeof(double));
  c->pp = (double *)calloc(NH, sizeof(double));
  c->psio = (double *)calloc(NH, sizeof(double));
  c->lambdaxy = (double *)calloc((long)c->n[0] * c->n[1], sizeof(double));
  c->a = (double *)calloc(n3, sizeof(double));
  c->b = (double *)calloc(n3, sizeof(double));
  c->c = (double *)calloc(n3, sizeof(double));
  c->rhsbx = (double *)calloc((long)c->n[1] * n3 * 2, sizeof(double));
  c->rhsby = (double *)calloc((long)c->n[0] * n3 * 2, sizeof(double));
  c->rhsbz = (double *)calloc((long)c->n[0] * c->n[1] * 2, sizeof(double));
  or_initsolver(c->n, c->dli, c->dzci, c->dzfi, prm->cbcpre, prm->bcpre, c->lambdaxy, c->a, c->b, c->c, &c->normfft,
                c->rhsbx, c->rhsby, c->rhsbz);
  /* cbccur = cbcpsi, bccur = bcpre, src/param.f90:172-173 */
  memcpy(c->cbccur, prm->cbcpsi, 6);
  memcpy(c->bccur, prm->bcpre, sizeof(double) * 6);
  /* rho0 = minval(rho12) under constant-coefficient Poisson, src/param.f90:174-177 */
  c->sc[5] = dmin(prm->rho12[0], prm->rho12[1]);
  c->sc[10] = 1.0;
  return c;
}

void or_case_destroy(or_case *c) {
  if (!c) return;
  free(c->dzc); free(c->dzf); free(c->zc); free(c->zf); free(c->dzci); free(c->dzfi);
  free(c->grid_vol_ratio_c); free(c->grid_vol_ratio_f);
  for (int q = 0; q < 19; q++) free(c->f[q]);
  free(c->dpsidt); free(c->dudt); free(c->dvdt); free(c->dwdt); free(c->pp); free(c->psio);
  free(c->lambdaxy); free(c->a); free(c->b); free(c->c); free(c->rhsbx); free(c->rhsby); free(c->rhsbz);
  free(c);
}

double *or_case_field(or_case *c, int id) { return (id >= 0 && id < 19) ? c->f[id] : NULL; }
double *or_case_scalars(or_case *c) { return c->sc; }

static void phase_bcs_and_geometry(or_case *c) {
  const or_params *P = &c->prm;
  double **f = c->f;
  if (!P->vof) or_acdi_cmpt_phi(c->n, c->sc[4], f[6], f[7]);
  else or_vof_thinc_cmpt_phi(c->n, 1.0 / (2.0 * P->psi_thickness_factor), f[6], f[7]);
  or_cmpt_norm_curv_youngs(c->n, c->dli, c->dzci, c->dzfi, f[7], f[9], f[10], f[11], f[8]);
  or_boundp(c->n, P->cbcpre, c->cbccur, c->bccur, c->dl, c->dzc, f[8]);
  or_boundp(c->n, P->cbcpre, P->cbcnor + 0, P->bcnor + 0, c->dl, c->dzc, f[9]);
  or_boundp(c->n, P->cbcpre, P->cbcnor + 6, P->bcnor + 6, c->dl, c->dzc, f[10]);
  or_boundp(c->n, P->cbcpre, P->cbcnor + 12, P->bcnor + 12, c->dl, c->dzc, f[11]);
}

static double new_dt(or_case *c, int *kill) {
  const or_params *P = &c->prm;
  double **f = c->f;
  const double dt_cfl = or_chkdt(c->n, c->dl, c->dzci, c->dzfi, P->is_solve_ns, P->is_track_interface, P->vof,
                                 P->mu12, P->rho12, P->sigma, P->gacc, f[0], f[1], f[2], c->sc[3], c->sc[4]);
  c->sc[8] = dt_cfl;
  double dt = dmin(P->cfl * dt_cfl, P->dtmax);
  if (P->dt_f > 0.0) {
    /* in the loop (src/main.f90:534-541) a fixed dt above the stable one stops
     * the run and keeps the stable dt; at start-up (:397-403) it only warns */
    if (P->dt_f > dt && kill)
      *kill = 1;
    else
      dt = P->dt_f;
  }
  return dt;
}

void or_case_init_fields(or_case *c) {
  const or_params *P = &c->prm;
  double **f = c->f;
  const long NH = (long)(c->n[0] + 2) * (c->n[1] + 2) * (c->n[2] + 2);
  /* seps, src/main.f90:301 (ACDI) or thickness for THINC */
  c->sc[4] = or_acdi_set_epsilon(c->dl, c->dzfi, c->n[2], P->psi_thickness_factor);
  /* initflow 'zer', src/initflow.f90:69-70 */
  for (int q = 0; q < 4; q++) memset(f[q], 0, sizeof(double) * NH);
  or_init2fl(P, c->sc[4], c->zc, f[6]);
  c->sc[0] = 0.0;
  c->sc[9] = 0.0;
  or_case_finish_init(c);
}

/* the initialisation after the fields are set (fresh start or restart,
 * src/main.f90:333-407): BCs, pn = po = p, phi/normals, gamma, dt */
void or_case_finish_init(or_case *c) {
  const or_params *P = &c->prm;
  double **f = c->f;
  const long NH = (long)(c->n[0] + 2) * (c->n[1] + 2) * (c->n[2] + 2);
  c->sc[4] = or_acdi_set_epsilon(c->dl, c->dzfi, c->n[2], P->psi_thickness_factor);
  or_bounduvw(c->n, P->cbcpre, P->cbcvel, P->bcvel, 0, c->dl, c->dzc, c->dzf, f[0], f[1], f[2]);
  or_boundp(c->n, P->cbcpre, P->cbcpre, P->bcpre, c->dl, c->dzc, f[3]);
  memcpy(f[4], f[3], sizeof(double) * NH);
  memcpy(f[5], f[4], sizeof(double) * NH);
  or_boundp(c->n, P->cbcpre, P->cbcpsi, P->bcpsi, c->dl, c->dzc, f[6]);
  phase_bcs_and_geometry(c);
  if (!P->vof) {
    const double g = or_acdi_set_gamma(c->n, P->acdi_gam_factor, f[0], f[1], f[2]);
    c->sc[3] = dmax(g, P->acdi_gam_min);
  }
  const double dt = new_dt(c, NULL);
  c->sc[1] = dt;
  c->sc[2] = dt;
  c->sc[10] = 1.0;
}

int or_case_step(or_case *c, int nsteps) {
  const or_params *P = &c->prm;
  double **f = c->f;
  const long NH = (long)(c->n[0] + 2) * (c->n[1] + 2) * (c->n[2] + 2);
  const long NI = (long)c->n[0] * c->n[1] * c->n[2];
  const double rho0 = c->sc[5];
  int kill = 0;
  for (int s = 0; s < nsteps && !kill; s++) {
    const double dt = c->sc[1], dto = c->sc[2];
    const int is_first = c->sc[10] != 0.0;
    c->sc[9] += 1.0;
    c->sc[0] += dt;
    const double dt_r = dt / dto;
    const double tm1 = 1.0 + 0.5 * dt_r, tm2 = -0.5 * dt_r;
    const double factor1 = tm1 * dt, factor2 = tm2 * dt, factor12 = factor1 + factor2;
    /* psio = psi, src/main.f90:430-432 */
    memcpy(c->psio, f[6], sizeof(double) * NH);
    if (P->is_track_interface) {
      /* rk_2fl, src/rk.f90:193-254 */
      if (!P->vof)
        or_acdi_transport_pf(c->n, c->dli, c->dzci, c->dzfi, c->sc[3], c->sc[4], f[0], f[1], f[2], f[9], f[10],
                             f[11], f[7], f[6], c->dpsidt, f[12], f[13], f[14]);
      else
        or_vof_thinc_transport_psi(c->n, c->dli, c->dzfi, 1.0 / (2.0 * P->psi_thickness_factor), f[0], f[1], f[2],
                                   f[9], f[10], f[11], f[7], c->dpsidt, f[12], f[13], f[14]);
      if (is_first) memcpy(f[15], c->dpsidt, sizeof(double) * NI);
      or_rk_psi_ab2(c->n, factor1, factor2, c->dpsidt, f[15], f[6]);
      or_clip_field(c->n, 0.0, 1.0, f[6]);
      memcpy(f[15], c->dpsidt, sizeof(double) * NI); /* swap: the current RHS becomes the old one */
      or_bounduvw(c->n, P->cbcpre, P->cbcvel, P->bcvel, 0, c->dl, c->dzc, c->dzf, f[12], f[13], f[14]);
      or_boundp(c->n, P->cbcpre, P->cbcpsi, P->bcpsi, c->dl, c->dzc, f[6]);
      phase_bcs_and_geometry(c);
    } else {
      memset(f[12], 0, sizeof(double) * NH);
      memset(f[13], 0, sizeof(double) * NH);
      memset(f[14], 0, sizeof(double) * NH);
    }
    if (P->is_solve_ns) {
      double rho_av[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < 3; d++)
        if (fabs(P->gacc[d]) > 0.0 && P->cbcpre[2 * d] == 'P' && P->cbcpre[2 * d + 1] == 'P')
          rho_av[d] = or_bulk_mean_12_stag(c->n, d + 1, d < 2 ? c->grid_vol_ratio_c : c->grid_vol_ratio_f, f[6],
                                           P->rho12);
      /* rk, src/rk.f90:16-131 */
      or_mom_xyz_ad(c->n, c->dli, c->dzci, c->dzfi, P->rho12, P->mu12, f[12], f[13], f[14], f[0], f[1], f[2], f[6],
                    c->dudt, c->dvdt, c->dwdt);
      if (is_first) {
        memcpy(f[16], c->dudt, sizeof(double) * NI);
        memcpy(f[17], c->dvdt, sizeof(double) * NI);
        memcpy(f[18], c->dwdt, sizeof(double) * NI);
      }
      or_rk_mom_ab2(c->n, factor1, factor2, P->rho12, c->psio, f[6], c->dudt, c->dvdt, c->dwdt, f[16], f[17], f[18],
                    f[0], f[1], f[2]);
      memcpy(f[16], c->dudt, sizeof(double) * NI);
      memcpy(f[17], c->dvdt, sizeof(double) * NI);
      memcpy(f[18], c->dwdt, sizeof(double) * NI);
      or_mom_xyz_oth(c->n, c->dli, c->dzci, c->dzfi, dt_r, P->rho12, P->bforce, P->gacc, P->sigma, rho0, rho_av,
                     f[3], f[6], f[8], f[4], f[5], c->dudt, c->dvdt, c->dwdt);
      or_rk_mom_oth(c->n, factor12, P->rho12, f[6], c->dudt, c->dvdt, c->dwdt, f[0], f[1], f[2]);
      if (P->is_forced_hit) or_lscale_forcing_tg(c->n, 0.5, dt, P->l, c->dl, c->zc, c->zf, f[0], f[1]);
      or_bounduvw(c->n, P->cbcpre, P->cbcvel, P->bcvel, 0, c->dl, c->dzc, c->dzf, f[0], f[1], f[2]);
      memcpy(c->pp, f[3], sizeof(double) * NH);
      or_fillps(c->n, c->dli, c->dzfi, 1.0 / dt, rho0, f[0], f[1], f[2], f[3]);
      or_updt_rhs_b(c->n, P->cbcpre, c->rhsbx, c->rhsby, c->rhsbz, f[3]);
      or_solver(c->n, P->cbcpre, c->normfft, c->lambdaxy, c->a, c->b, c->c, f[3]);
      or_boundp(c->n, P->cbcpre, P->cbcpre, P->bcpre, c->dl, c->dzc, f[3]);
      or_correc(c->n, c->dli, c->dzci, rho0, dt, f[3], f[0], f[1], f[2]);
      or_bounduvw(c->n, P->cbcpre, P->cbcvel, P->bcvel, 1, c->dl, c->dzc, c->dzf, f[0], f[1], f[2]);
      or_updatep(c->n, c->pp, f[3]);
      or_boundp(c->n, P->cbcpre, P->cbcpre, P->bcpre, c->dl, c->dzc, f[3]);
    }
    /* po = pn; pn = p, src/main.f90:505-508 */
    memcpy(f[5], f[4], sizeof(double) * NH);
    memcpy(f[4], f[3], sizeof(double) * NH);
    c->sc[10] = 0.0;
    c->sc[2] = dt; /* dto = dt */
    if (!P->vof) {
      const double g = or_acdi_set_gamma(c->n, P->acdi_gam_factor, f[0], f[1], f[2]);
      c->sc[3] = dmax(g, P->acdi_gam_min);
    }
    const int istep = (int)c->sc[9];
    if (P->icheck > 0 && istep % P->icheck == 0) {
      double dtn = new_dt(c, &kill);
      if (c->sc[8] < 2.220446049250313e-16 * 1e7) kill = 1; /* small = eps*10**(precision/2), precision(1._dp)=15 */
      c->sc[1] = dtn;
      double dv[2];
      or_chkdiv(c->n, c->dli, c->dzfi, f[0], f[1], f[2], dv);
      c->sc[6] = dv[0];
      c->sc[7] = dv[1];
      if (dv[1] > 2.220446049250313e-16 * 1e7 || isnan(dv[0])) kill = 1;
    }
  }
  return kill;
}

/* evaluation hook for tests/test_oracle_math.py: fn 0 exp,1 expm1,2 log,3 tanh,4 sin,5 cos */
void or_math_eval(int fn, long n, const double *x, double *y) {
  for (long q = 0; q < n; q++) {
    switch (fn) {
    case 0: y[q] = om_exp(x[q]); break;
    case 1: y[q] = om_expm1(x[q]); break;
    case 2: y[q] = om_log(x[q]); break;
    case 3: y[q] = om_tanh(x[q]); break;
    case 4: y[q] = om_sin(x[q]); break;
    default: y[q] = om_cos(x[q]); break;
    }
  }
}
