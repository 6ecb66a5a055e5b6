// Autotuner of a case's launch parameters. No reference counterpart.
//
// The parameters change the order and shape of launches, never the
// arithmetic, so every choice gives the same fields bit for bit:
//  - tile_y: rows of blocks per z-marching pencil in the stencil block order;
//  - the variant of the fused momentum kernel (occupancy, staged loads,
//    stencil fields staged in LDS);
//  - rows per chunk of the Thomas kernel (FP64, or FP32 in mixed cases,
//    where two columns per lane are a candidate too);
//  - for a non-periodic z, the Thomas kernel without a pivot array (it
//    reads a column twice instead of storing and reloading pivots).
// A scratch case with the same parameters (dt checks off, so every trial
// times the same work) is stepped once per candidate to warm up and then
// twice `trial_steps` times between HIP events on its stream (the faster
// of the two counts); the parameters are
// chosen one after another (coordinate descent), each keeping the best time
// so far. The result is applied to the case's context and cached by grid,
// boundary conditions, precision and the precision of the projection (a
// mixed case with an FP64 projection tunes the FP64 Thomas kernel), so cases
// of the same shape reuse it.
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "fz_internal.h"

namespace {
struct Tuned {
  int tile_y, mom_variant, tk, rc;
  float ms;
};
std::map<std::string, Tuned> g_tuned;

// proj32: the projection solves in FP32 (mixed precision with refinements)
std::string shape_key(const fz_params &p, int precision, bool proj32) {
  char b[160];
  snprintf(b, sizeof(b), "%d,%d,%d|%.6s|%d|%d|%d", p.ng[0], p.ng[1], p.ng[2], p.cbcpre, precision, p.vof,
           proj32 ? 1 : 0);
  return b;
}

void apply(fz_ctx *c, const Tuned &t, bool proj32) {
  c->d.tile_y = t.tile_y;
  c->mom_variant = t.mom_variant;
  if (proj32) c->tk32 = t.tk;
  else c->tk64 = t.tk;
  c->thomas_rc = t.rc;
}

// ms per step of the scratch case under t: the better of two timed runs
int trial(fz_case *sc, fz_ctx *c, const Tuned &t, bool proj32, int steps, hipEvent_t a, hipEvent_t b, float *ms) {
  apply(c, t, proj32);
  int killed = 0;
  FZ_TRY(fz_case_step(sc, 1, &killed));
  float best = 0.0f;
  for (int rep = 0; rep < 2; rep++) {
    FZ_HIP(hipEventRecord(a, c->st));
    FZ_TRY(fz_case_step(sc, steps, &killed));
    FZ_HIP(hipEventRecord(b, c->st));
    FZ_HIP(hipEventSynchronize(b));
    float e = 0.0f;
    FZ_HIP(hipEventElapsedTime(&e, a, b));
    if (rep == 0 || e < best) best = e;
  }
  *ms = best / (float)steps;
  return FZ_OK;
}
}  // namespace

extern "C" int fz_case_autotune(fz_case *cs, int trial_steps, char *report, size_t cap) {
  if (!cs || trial_steps < 1) return fz_set_error(FZ_ERR_ARG, "fz_case_autotune: bad argument");
  int precision = FZ_PREC_FP64, nref = 0;
  FZ_TRY(fz_case_precision(cs, &precision, &nref));
  // the Thomas kernel the step runs: FP32 columns only when a mixed case
  // refines an FP32 solve (nref >= 0); nref = -1 keeps the FP64 projection
  const bool proj32 = precision == FZ_PREC_MIXED && nref >= 0;
  fz_ctx *ctx = nullptr;
  FZ_TRY(fz_case_ctx(cs, &ctx));
  fz_params prm = *fz_case_prm(cs);
  const std::string key = shape_key(prm, precision, proj32);
  auto it = g_tuned.find(key);
  Tuned best;
  bool cached = it != g_tuned.end();
  if (cached) {
    best = it->second;
  } else {
    prm.icheck = 0;  // the same work in every trial
    fz_case *sc = nullptr;
    FZ_TRY(fz_case_create_ex(&sc, &prm, nullptr, precision, nref));
    fz_ctx *c = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = fz_case_ctx(sc, &c);
    if (rc == FZ_OK) rc = fz_case_init(sc);
    if (rc == FZ_OK && hipEventCreate(&a) != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "hipEventCreate");
    if (rc == FZ_OK && hipEventCreate(&b) != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "hipEventCreate");
    best = {c ? c->d.tile_y : 8, c ? c->mom_variant : 0, proj32 ? (c ? c->tk32 : -8) : (c ? c->tk64 : 16),
            c ? c->thomas_rc : 0, 0.0f};
    if (rc == FZ_OK) rc = trial(sc, c, best, proj32, trial_steps, a, b, &best.ms);
    const int tiles[] = {0, 2, 4, 8, 16};
    const int moms[] = {1, 2, 4, 5, 8};
    const int tks[] = {8, 16, 32, -8, -16};  // negative: two FP32 columns per lane (FP32 solves)
    const int rcs[] = {0, 1};
    const bool periodic_z = prm.cbcpre[4] == 'P';
    for (int dim = 0; dim < 4 && rc == FZ_OK; dim++) {
      const int *cand = dim == 0 ? tiles : (dim == 1 ? moms : (dim == 2 ? tks : rcs));
      int ncand = dim == 0 ? 5 : (dim == 1 ? 5 : (dim == 2 ? (proj32 ? 5 : 3) : 2));
      if (dim == 3 && (periodic_z || best.tk < 0)) ncand = 0;  // no such kernel there
      for (int q = 0; q < ncand && rc == FZ_OK; q++) {
        Tuned t = best;
        int &field = dim == 0 ? t.tile_y : (dim == 1 ? t.mom_variant : (dim == 2 ? t.tk : t.rc));
        if (field == cand[q]) continue;
        field = cand[q];
        rc = trial(sc, c, t, proj32, trial_steps, a, b, &t.ms);
        if (rc == FZ_OK && t.ms < best.ms) best = t;
      }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    fz_case_destroy(sc);
    if (rc != FZ_OK) return rc;
    g_tuned[key] = best;
  }
  FZ_HIP(hipStreamSynchronize(ctx->st));
  apply(ctx, best, proj32);
  fz_case_drop_graphs(cs);  // captured launches carry the old parameters
  if (report && cap > 0)
    snprintf(report, cap, "tile_y=%d mom_variant=%d thomas_rows=%d%s ms_per_step=%.4f%s", best.tile_y,
             best.mom_variant, best.tk < 0 ? -best.tk : best.tk, best.tk < 0 ? "x2" : (best.rc ? "rc" : ""), best.ms,
             cached ? " (cached)" : "");
  return FZ_OK;
}
