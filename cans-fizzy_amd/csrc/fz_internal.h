// Internal declarations of libfizzy (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/fizzy.h"

#define FZ_EPS 2.220446049250313080847e-16  // epsilon(1._rp)
#define FZ_GAMMA_SLOTS 256                   // atomic-max slots of the fused gamma reduction
#define FZ_PI 3.141592653589793115998        // acos(-1._rp)

int fz_set_error(int code, const char *fmt, ...);

#define FZ_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return fz_set_error(FZ_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)
#define FZ_LAUNCHED() FZ_HIP(hipGetLastError())
#define FZ_TRY(call)          \
  do {                        \
    int r_ = (call);          \
    if (r_ != FZ_OK) return r_; \
  } while (0)

// Geometry of one field in HBM: the reference layout, column-major with a
// one-cell halo; (i,j,k) lives at i + sy*j + sz*k.
struct fz_dims {
  int n1, n2, n3;
  long sy, sz;
  int tile_y;  // rows of blocks per z-marching pencil in the stencil block order (0: plain order)
};

// Batched line transform plan along one axis (r2r kinds as FFTW's).
struct fz_fftplan {
  int n = 0;         // real line length
  int m = 0;         // complex FFT length (n, or 2n for the quarter-wave kinds)
  int kind_f = 0, kind_b = 0;
  int nrad = 0;
  int rad[24] = {0};
  double2 *tw = nullptr;  // exp(-2 pi i j / m), j < m
  double2 *q = nullptr;   // (cos, sin)(pi k / 2n), k <= n
  double2 *t2 = nullptr;  // (cos, sin)(pi (2k+1) / 4n), k < n
  float2 *tw32 = nullptr, *q32 = nullptr, *t232 = nullptr;  // the same tables rounded to FP32
};

struct fz_poisson {
  fz_fftplan px, py;
  double normfft = 1.0;
  double *lambdaxy = nullptr;  // device n1*n2
  double *a = nullptr, *b = nullptr, *c = nullptr;  // device n3
  double *rhsbx = nullptr, *rhsby = nullptr, *rhsbz = nullptr;  // device
  double *work = nullptr;   // spectral buffer, n1*n2*n3
  double *piv = nullptr;    // Thomas pivots, n1*n2*n3
  double *aux = nullptr;    // periodic z: gaussel_periodic's second solve, made at set-up, n1*n2*n3
  double *zpre = nullptr;   // non-periodic z with precomputed pivots: 1/pivot and d of every row, n1*n2*n3
  double *dpre = nullptr;   // each
  int periodic_z = 0;
  int rhsb_zero = 0;  // all boundary right-hand sides vanish (homogeneous pressure BCs)
  long null_col = -1;  // spectral column of the operator's null mode, or -1
  float *work32 = nullptr;              // FP32 spectral buffer of the mixed-precision solve
  double *bcopy = nullptr;              // right-hand side kept for the residuals (halo'd layout)
  unsigned long long *rslots = nullptr; // max |residual| slots (FP64 bits)
  void **batch_ptrs = nullptr;          // device pointer arrays of fz_solver_batched
  size_t batch_cap = 0;
  char cbc[6];
};

// per-launch timing with HIP events on the context's stream (off by default)
struct fz_prof_rec {
  const char *name;
  hipEvent_t a, b;
};

struct fz_ctx {
  bool prof_on = false;
  std::vector<fz_prof_rec> prof_recs;
  std::vector<hipEvent_t> prof_pool;
  fz_dims d;
  long nh = 0, ni = 0;  // element counts of a halo'd and of an interior array
  double l[3], dl[3], dli[3];
  std::vector<double> h_dzc, h_dzf, h_zc, h_zf, h_dzci, h_dzfi;
  double *dzc = nullptr, *dzf = nullptr, *zc = nullptr, *zf = nullptr, *dzci = nullptr, *dzfi = nullptr;
  char cbcpre[6];
  int per[3];
  hipStream_t st = nullptr;
  bool own_stream = false;
  fz_poisson *pois = nullptr;
  double *red = nullptr;      // per-block partial results of reductions
  size_t red_cap = 0;
  double *h_small = nullptr;  // pinned host landing zone for reduction results
  double *scratch = nullptr;  // device address of this GPU's static reduction scratch (fz_scratch_setup)
  double *ftab = nullptr;     // the forcing's sine/cosine tables by coordinate, made at its first call
  bool youngs_fast_grid = false;  // dxi, dyi and every dzci in [2^-20, 2^20] (k_youngs' range argument)
  std::map<std::string, struct fz_bcmap *> bcmaps;  // composed boundary maps by signature
  // launch parameters (defaults, environment overrides, or fz_case_autotune)
  int mom_variant = 0;  // momentum kernel: 0 the default (8), 1 staged
                        // loads, 2 plain, 4 or 5 waves per EU, 8 stencil fields in LDS (k_mom_lds)
  int tk64 = 16;        // rows per chunk of the FP64 Thomas kernel: 8, 16 or 32
  int tk32 = -8;        // the same for the FP32 one; negative: two columns per lane, -8 or -16 (the default,
                        // -8, is the tuner's choice at 256x256x512 and 256^3)
  int acdi_variant = 0;  // ACDI step: 0 the default (3), 1 fields staged in LDS, 2 one cell per thread, 3 each face once
  int youngs_variant = 0;  // Youngs kernel: 0 the default (FP32 1, FP64 3), 1 marching in z, 2 one cell per thread,
                           // 3 marching without range checks when the range flag allows, else one cell per thread
                         // (k_acdi_lds), 2 one cell per thread
  int correc_kz = 1;    // correction step: 0 one cell per thread, 1 dp, u, v staged in LDS (k_correc_lds),
                        // 8 or 16 cells per thread marching in z (k_correc_step_mz)
  int correc_my = 4;    // rows of cells per block of the LDS correction kernel: 4, 8 or 16
  int p2_xdiv = 2;      // x-line transforms run 1 or 2 times fewer lines per block than y lines
  int thomas_pre = 0;   // 1: Thomas solves with the pivots made once at set-up (k_thomas_pre); set at
                        // creation (1 unless thomas_rc), cleared for cases whose solves are FP32
  int thomas_rc = 0;    // 1: Thomas kernels without a pivot array (the elimination re-run per chunk);
                        // set at creation from the column count
};

// brackets one launch with events when profiling is on
struct fz_prof_scope {
  fz_ctx *c;
  const char *name;
  hipEvent_t a = nullptr;
  fz_prof_scope(fz_ctx *c_, const char *n);
  ~fz_prof_scope();
};
#define FZ_LAUNCH(c, name, kernel, grid, block, shmem, ...)                              \
  do {                                                                                   \
    fz_prof_scope fz_ps_((c), name);                                                     \
    hipLaunchKernelGGL(kernel, grid, block, shmem, (c)->st, __VA_ARGS__);                \
  } while (0)

// Workgroups are dealt round-robin to the 8 XCDs in dispatch order, each XCD
// with its own L2. Remapping the linear block id so that XCD x runs one
// contiguous range of the logical ids keeps the neighbouring rows and planes
// of a stencil in one L2. With nb = 8q + r, XCD x receives q + (x < r)
// blocks, and its range starts at x*q + min(x, r): a bijection of [0, nb).
__device__ inline unsigned fz_xcd_block(unsigned P, unsigned nb) {
  const unsigned x = P & 7u, q = nb >> 3, r = nb & 7u;
  return x * q + (x < r ? x : r) + (P >> 3);
}

// Block coordinates of a 3-D stencil grid (x blocks, y blocks, one z plane
// per block) in pencil order: after the XCD remap, consecutive blocks cover
// one pencil of ty rows of blocks (full x extent) and march it through z, so
// the k-1/k/k+1 planes an XCD's resident blocks read stay in its L2.
// A batched launch (fz_case_batch) stacks the grids of its cases in z: gz
// blocks per case, the case in cs (gz = gridDim.z and cs = 0 otherwise).
__device__ inline void fz_stencil_block_b(unsigned ty, unsigned gz, unsigned &cs, unsigned &bx, unsigned &by,
                                          unsigned &bz) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned nb = gx * gy * gridDim.z;
  unsigned L = fz_xcd_block(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), nb);
  cs = L / (gx * gy * gz);
  L -= cs * (gx * gy * gz);
  if (ty == 0 || gy % ty != 0) {
    bz = L / (gx * gy);
    by = (L - bz * gx * gy) / gx;
    bx = L - bz * gx * gy - by * gx;
    return;
  }
  const unsigned per_plane = gx * ty;      // blocks of one pencil in one plane
  const unsigned pencil = L / (per_plane * gz);
  const unsigned rem = L - pencil * per_plane * gz;
  bz = rem / per_plane;
  const unsigned in = rem - bz * per_plane;
  by = pencil * ty + in / gx;
  bx = in - (in / gx) * gx;
}
__device__ inline void fz_stencil_block(unsigned ty, unsigned &bx, unsigned &by, unsigned &bz) {
  unsigned cs;
  fz_stencil_block_b(ty, gridDim.z, cs, bx, by, bz);
}

// One case of a batched step (fz_case_batch): the buffer of every role the
// step's kernels read or write, in the two pointer states the case
// alternates between from step to step (the rotations psio <-> psi and
// p <-> po), and the case's own scalars. A batched kernel takes the case of
// its block, reads the state s = par[case] and uses at[s][role].
enum {
  BR_U, BR_V, BR_W, BR_UN, BR_VN, BR_WN, BR_P, BR_PO, BR_DP, BR_PSI_OLD, BR_PSI_NEW, BR_PHI, BR_NX, BR_NY,
  BR_NZ, BR_KAPPA, BR_FX, BR_FY, BR_FZ, BR_DPSIO, BR_DUO, BR_DVO, BR_DWO, BR_N
};
struct fz_bcase {
  void *at[2][BR_N];
  const double *stc;              // dt and the AB2 factors (FZ_STC_*)
  double *dsc;                    // gamma in dsc[0]
  unsigned long long *gslots;     // gamma slots of the correction step
  double rho, drho, mu, dmu, surf_factor, rho0, seps, gam_factor, gam_min;
  double bforce[3], gacc[3];
};
// the batch of a launch: null bt for a single case
struct fz_bref {
  const fz_bcase *bt;
  const int *par;
  unsigned gz;  // grid z blocks of one case
};
__device__ inline void *fz_brole(const fz_bref &b, unsigned cs, int role) { return b.bt[cs].at[b.par[cs]][role]; }
static const fz_bref kNoBRef = {nullptr, nullptr, 0};

// ---- launch helpers ----
static inline dim3 fz_block3() { return dim3(64, 4, 1); }
static inline dim3 fz_grid_interior(const fz_dims &d) {
  return dim3((d.n1 + 63) / 64, (d.n2 + 3) / 4, d.n3);
}
static inline dim3 fz_grid_halo(const fz_dims &d) {
  return dim3((d.n1 + 2 + 63) / 64, (d.n2 + 2 + 3) / 4, d.n3 + 2);
}

// internal entry points shared between translation units
int fz_reduce_ensure(fz_ctx *ctx, size_t nblocks);
int fz_scratch_setup(fz_ctx *ctx);
int fz_batch_solve(fz_ctx *c0, void **dptrs, int n);
// why cs cannot join a batched step led by first (null: it can)
const char *fz_case_batch_refusal(const fz_case *first, const fz_case *cs);
void fz_bcmaps_free(fz_ctx *ctx);
int fz_poisson_destroy(fz_poisson *p);
int fz_solver_dev(fz_ctx *ctx, double *p);
// a case's per-step scalars in device memory (kernels read them when given
// a pointer, so a captured step graph serves any dt)
enum { FZ_STC_DT = 0, FZ_STC_DTI, FZ_STC_F1, FZ_STC_F2, FZ_STC_F12, FZ_STC_DTR, FZ_STC_N };
int fz_projection_rhs_solve(fz_ctx *ctx, double dti, double rho0, const double *u, const double *v, const double *w,
                            double *p, const double *stc = nullptr);
int fz_projection_rhs_solve_mixed(fz_ctx *ctx, double dti, double rho0, const double *u, const double *v,
                                  const double *w, double *p, int nref, const double *stc = nullptr);
int fz_projection_device_scalars_ok(fz_ctx *ctx, int mixed);
const fz_params *fz_case_prm(const fz_case *cs);
// one boundary update of a group: kind 0 boundp (f[0]), 1 bounduvw (f[0..2]);
// xdone: the producing kernel wrote the periodic x images of the interior
// rows (the halo cells 0 and n1+1), which the update then leaves alone
struct FzBcItem {
  int kind;
  const char *cbc;
  const double *bc;
  int is_correc;
  void *f[3];
  int xdone;
};
template <typename T>
int fz_bounds_group(fz_ctx *c, int n, const FzBcItem *items);
void fz_case_drop_graphs(fz_case *cs);
// while on, a step that would capture its graph runs its launches directly
// (the server sets it while its set-up thread makes HIP calls)
void fz_case_block_capture(fz_case *cs, bool on);
int fz_acdi_gamma_dev(fz_ctx *ctx, double gam_factor, double gam_min, const double *u, const double *v,
                      const double *w, double *d_gam);
int fz_acdi_transport_dev(fz_ctx *ctx, const double *d_gam, double gam, double seps, const double *u,
                          const double *v, const double *w, const double *normx, const double *normy,
                          const double *normz, const double *phi, const double *psi, double *dpsidt,
                          double *flux_x, double *flux_y, double *flux_z);
