// Serving: many independent cases on one device, admitted by HBM footprint
// and stepped together (continuous batching). No reference counterpart: the
// reference runs one case per job. The footprint model follows the idea of
// device_memory_footprint (src/utils.f90:155-219) for this library's own
// buffers.
//
// Each admitted case owns its context, stream and every buffer, and its
// kernels are deterministic, so a case's fields are the same whether it runs
// alone or in a mix. A scheduling round admits queued requests in arrival
// order while the footprint budget allows (head-of-line: a request that
// does not fit waits, and the ones behind it wait too) and hands them to a
// set-up thread, which creates and initialises each case (allocation, the
// solver's set-up, the initial fields) on the case's own stream while the
// rounds go on stepping the others. The round then takes the cases that are
// ready (it waits for one only when nothing else is running),
// enqueues one time step of every active case on its own stream (the steps
// of different cases overlap on the device), then waits for each case's
// step and records its duration from HIP events on that case's stream. A
// case that has done its steps leaves the active set; its memory is returned
// when the caller releases it.
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "fz_internal.h"

// Device bytes of `count` buffers of `b` bytes each: allocations of 2 MiB
// or more are made in 2 MiB units; smaller ones are packed into 2 MiB pages
// (as many whole buffers, 4 KiB aligned, as fit in one page).
static double alloc_bytes(double b, double count) {
  const double page = 2.0 * 1024 * 1024;
  if (b >= page) return count * page * std::ceil(b / page);
  const double aligned = 4096.0 * std::ceil(b / 4096.0);
  const double per_page = std::floor(page / aligned);
  return page * std::ceil(count / per_page);
}

// bytes of device memory fz_case_create_ex + fz_case_init allocate for prm
// nref: the refinements of a mixed case (-1: its projection is FP64)
static int footprint(const fz_params *prm, int precision, int nref, size_t *bytes) {
  if (!prm || !bytes) return fz_set_error(FZ_ERR_ARG, "fz_case_footprint: null argument");
  const double n1 = prm->ng[0], n2 = prm->ng[1], n3 = prm->ng[2];
  const double nh = (n1 + 2) * (n2 + 2) * (n3 + 2), ni = n1 * n2 * n3;
  const bool periodic_z = prm->cbcpre[4] == 'P' && prm->cbcpre[5] == 'P';
  double b = 0.0;
  // solves of a non-periodic z below 512x512 columns (FP64, or FP32 in pairs
  // of columns) keep pivots made at set-up: two more N-element arrays (the
  // context's default)
  (void)nref;
  if (!periodic_z && n1 * n2 < 262144.0) b += alloc_bytes(8.0 * ni, 2);
  if (precision == FZ_PREC_MIXED) {
    b += alloc_bytes(8.0 * nh, 5 + 4);  // u, v, w, p, po + dp, three spare velocities
    b += alloc_bytes(4.0 * nh, 9 + 1);  // psi, phi, kappa, normals, fluxes + psio (FP32)
    b += alloc_bytes(4.0 * ni, 4 + 1);  // four AB2 histories + the FP32 spectral buffer
    b += alloc_bytes(8.0 * ni, periodic_z ? 3 : 2);  // FP64 Poisson work array and pivots (+ periodic z)
  } else {
    b += alloc_bytes(8.0 * nh, 14 + 5);  // case fields (pn aliases p) + psio, dp, three spare velocities
    // four AB2 histories, the Poisson work array and Thomas pivots (+ the
    // periodic-z second solve)
    b += alloc_bytes(8.0 * ni, 4 + (periodic_z ? 3 : 2));
  }
  b += 8.0 * (n1 * n2 + 3 * n3 + 2 * (n2 * n3 + n1 * n3 + n1 * n2));  // eigenvalues, tridiagonal, boundary terms
  const double planes = 2.0 * ((n1 + 2) * (n2 + 2) + (n1 + 2) * (n3 + 2) + (n2 + 2) * (n3 + 2));
  b += 9.0 * planes * 12;    // composed boundary maps (about 12 signatures per case)
  b += 8.0 * ni / 64.0 * 2;  // reduction partials
  b += 4.0 * 1024 * 1024;    // grids, plans, scalars, allocation rounding
  *bytes = (size_t)b;
  return FZ_OK;
}

extern "C" int fz_case_footprint_ex(const fz_params *prm, int precision, size_t *bytes) {
  return footprint(prm, precision, 1, bytes);  // a mixed case's default refinement count
}

extern "C" int fz_case_footprint(const fz_params *prm, size_t *bytes) {
  return fz_case_footprint_ex(prm, FZ_PREC_FP64, bytes);
}

namespace {
struct Request {
  int id;
  fz_params prm;
  int nsteps;
  int precision, nref;
  size_t bytes;
};
// a round gives each case up to max_batch steps (1, or FZ_SERVE_BATCH in
// [1, FZ_SERVE_MAX_BATCH]), about as many as fit in the slowest running
// case's step time. On the serving mix one step per round gave the best
// throughput and p99 (8 steps: the lower p50)
#define FZ_SERVE_MAX_BATCH 16
struct Running {
  int id;
  fz_case *cs;
  fz_ctx *ctx;
  int left;
  size_t bytes;
  hipEvent_t ev[2 * FZ_SERVE_MAX_BATCH];  // start and end of each step of the round
  int nq;                                 // steps enqueued in this round
  float last_ms;                          // duration of its latest step
  bool killed;
  int rc;           // the first error of this case in the round (FZ_OK if none)
  std::string err;
};
struct Finished {
  fz_case *cs;       // null for a failed request (its case is gone, its memory returned)
  size_t bytes;
  int killed;
  int failed;        // set-up or a step of the request returned an error
  std::string err;   // that error's message
};

// same-shaped running cases stepped together (fz_case_batch), kept from
// round to round while its members stay the same
struct Group {
  std::vector<int> ids;
  fz_case_batch *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
};

struct Prepared {  // a set-up finished by the set-up thread
  Running r;
  int rc;
  std::string err;
};
}  // namespace

struct fz_server {
  size_t budget = 0, in_use = 0;
  int next_id = 1;
  int device = 0;
  int max_batch = 1;
  std::deque<Request> queue;
  std::vector<Running> active;
  std::map<int, Finished> done;
  std::map<int, std::vector<float>> step_ms;  // per case, every step's duration
  std::vector<float> all_ms;
  // the set-up thread: requests admitted and not yet set up, the results
  std::mutex mu;
  std::condition_variable cv_work, cv_ready;
  std::deque<Request> to_prepare;
  std::deque<Prepared> ready;
  std::set<int> preparing;  // ids admitted and not yet taken by a round
  bool stop = false;
  // held by the set-up thread for a whole set-up, and by a round while it
  // enqueues steps that may capture graphs: a capture on one thread and
  // allocations or copies on another have invalidated the capture
  std::mutex hip_mu;
  std::thread worker;
  // batched steps of same-shaped small cases (FZ_SERVE_GROUP=1: on). Off by
  // default: on the serving mix, whose time goes to its large cases, the
  // batches gave no throughput and a higher per-case step latency (one
  // stream, and a batched step's time counts for each of its cases); on an
  // ensemble of small 2-D cases they raised throughput by 1.2x
  bool group = false;
  // the largest case (cells) that is batched: the steps of small cases are
  // bound by launch latency, which a batch pays once for all its cases
  // (measured: eight 64x2x64 cases 0.25 ms per batched step against
  // 0.55 ms stepped on their own streams); a larger case's launches are
  // long enough for the streams of separate cases to overlap them, which a
  // batch (one stream) gives up (FZ_SERVE_GROUP_MAX_CELLS)
  long group_max_cells = 1L << 17;
  std::vector<Group> groups;
};

static void drop_group(Group &g) {
  if (g.b) fz_case_batch_destroy(g.b);
  if (g.e0) (void)hipEventDestroy(g.e0);
  if (g.e1) (void)hipEventDestroy(g.e1);
  g = Group();
}

static void drop_running(Running &r) {
  for (hipEvent_t &e : r.ev)
    if (e) (void)hipEventDestroy(e);
  fz_case_destroy(r.cs);
}

// create and initialise the case of request q (on its own stream)
static int prepare(const Request &q, Running &r) {
  r = Running();
  r.id = q.id;
  r.left = q.nsteps;
  r.bytes = q.bytes;
  FZ_TRY(fz_case_create_ex(&r.cs, &q.prm, nullptr, q.precision, q.nref));
  int rc = fz_case_ctx(r.cs, &r.ctx);
  if (rc == FZ_OK) rc = fz_case_init(r.cs);
  if (rc == FZ_OK) rc = fz_case_set_graphs(r.cs, 1);  // steps after the first replay captured graphs
  for (hipEvent_t &e : r.ev)
    if (rc == FZ_OK && hipEventCreate(&e) != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "hipEventCreate");
  if (rc != FZ_OK) {
    drop_running(r);
    r.cs = nullptr;
  }
  return rc;
}

static void worker_loop(fz_server *s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    Request q;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [s] { return s->stop || !s->to_prepare.empty(); });
      if (s->stop) return;
      q = s->to_prepare.front();
      s->to_prepare.pop_front();
    }
    Prepared p;
    {
      std::lock_guard<std::mutex> lk(s->hip_mu);
      p.rc = prepare(q, p.r);
    }
    if (p.rc != FZ_OK) p.err = fz_last_error();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->ready.push_back(p);
    }
    s->cv_ready.notify_all();
  }
}

extern "C" int fz_server_create(fz_server **out, size_t hbm_budget) {
  if (!out || hbm_budget == 0) return fz_set_error(FZ_ERR_ARG, "fz_server_create: bad argument");
  int dev = 0;
  FZ_HIP(hipGetDevice(&dev));
  fz_server *s = new fz_server();
  s->budget = hbm_budget;
  s->device = dev;
  if (const char *e = getenv("FZ_SERVE_BATCH")) s->max_batch = std::max(1, std::min(FZ_SERVE_MAX_BATCH, atoi(e)));
  if (const char *e = getenv("FZ_SERVE_GROUP")) s->group = atoi(e) != 0;
  if (const char *e = getenv("FZ_SERVE_GROUP_MAX_CELLS")) s->group_max_cells = atol(e);
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return FZ_OK;
}

extern "C" int fz_server_destroy(fz_server *s) {
  if (!s) return FZ_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Group &g : s->groups) drop_group(g);
  for (Prepared &p : s->ready)
    if (p.rc == FZ_OK) drop_running(p.r);
  for (Running &r : s->active) drop_running(r);
  for (auto &kv : s->done)
    if (kv.second.cs) fz_case_destroy(kv.second.cs);
  delete s;
  return FZ_OK;
}

extern "C" int fz_server_submit(fz_server *s, const fz_params *prm, int nsteps, int *id) {
  return fz_server_submit_ex(s, prm, nsteps, FZ_PREC_FP64, 0, id);
}

extern "C" int fz_server_submit_ex(fz_server *s, const fz_params *prm, int nsteps, int precision, int nref,
                                   int *id) {
  if (!s || !prm || nsteps < 0 || !id) return fz_set_error(FZ_ERR_ARG, "fz_server_submit: bad argument");
  Request q;
  q.id = s->next_id++;
  q.prm = *prm;
  q.nsteps = nsteps;
  q.precision = precision;
  q.nref = nref;
  FZ_TRY(footprint(prm, precision, nref, &q.bytes));
  if (q.bytes > s->budget)
    return fz_set_error(FZ_ERR_ARG, "fz_server_submit: the case needs %zu bytes, more than the whole budget %zu",
                        q.bytes, s->budget);
  s->queue.push_back(q);
  *id = q.id;
  return FZ_OK;
}

// queued requests that fit, in arrival order, to the set-up thread
static void admit(fz_server *s) {
  bool any = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    while (!s->queue.empty() && s->in_use + s->queue.front().bytes <= s->budget) {
      const Request q = s->queue.front();
      s->queue.pop_front();
      s->in_use += q.bytes;
      s->preparing.insert(q.id);
      s->to_prepare.push_back(q);
      any = true;
    }
  }
  if (any) s->cv_work.notify_one();
}

// a request that failed: it finishes with its error, and its memory returns
static void fail_request(fz_server *s, int id, size_t bytes, const std::string &err) {
  s->in_use -= bytes;
  s->done[id] = Finished{nullptr, 0, 0, 1, err};
}

// the finished set-ups join the active set; a failed one finishes as failed
// (fz_server_status, fz_server_error) without stopping the others
static void take_ready(fz_server *s, bool wait) {
  std::unique_lock<std::mutex> lk(s->mu);
  if (wait) s->cv_ready.wait(lk, [s] { return !s->ready.empty() || s->preparing.empty(); });
  while (!s->ready.empty()) {
    Prepared p = s->ready.front();
    s->ready.pop_front();
    s->preparing.erase(p.r.id);
    if (p.rc == FZ_OK) s->active.push_back(p.r);
    else fail_request(s, p.r.id, p.r.bytes, "set-up: " + p.err);
  }
}

extern "C" int fz_server_round(fz_server *s, int *n_active, int *n_queued) {
  if (!s) return fz_set_error(FZ_ERR_ARG, "fz_server_round: null server");
  admit(s);
  take_ready(s, s->active.empty());  // with nothing to step, wait for a set-up
  // steps of every active case, each case on its own stream: all are
  // enqueued before any is waited for. A case gets about as many steps as
  // fit in the slowest case's step (at most max_batch), so that
  // small cases keep the device busy beside large ones; its steps stop at
  // one whose checks must be read before the next step (its dt). Steps are
  // enqueued step by step across the cases, which interleaves the streams'
  // work on the hardware queues they share. A case whose graph is still to
  // be captured runs the step's launches directly while the set-up thread
  // is busy.
  //
  // Running cases of one shape (grid, boundary conditions, precision) that
  // have taken their first step are stepped together instead, one batched
  // step per round (fz_case_batch: one launch per stage for all of them),
  // after the single cases' steps are enqueued.
  const size_t na = s->active.size();
  std::vector<int> gid(na, -1);
  std::vector<Group> groups;
  if (s->group) {
    std::vector<std::vector<size_t>> members;
    for (size_t q = 0; q < na; q++) {
      const Running &r = s->active[q];
      if (r.left <= 0 || r.last_ms <= 0.0f || r.ctx->ni > s->group_max_cells || fz_case_batch_refusal(r.cs, r.cs))
        continue;
      bool placed = false;
      for (size_t g = 0; g < members.size() && !placed; g++)
        if (!fz_case_batch_refusal(s->active[members[g][0]].cs, r.cs)) {
          members[g].push_back(q);
          placed = true;
        }
      if (!placed) members.push_back({q});
    }
    for (const std::vector<size_t> &m : members) {
      if (m.size() < 2) continue;
      Group g;
      for (size_t q : m) g.ids.push_back(s->active[q].id);
      bool found = false;
      for (Group &o : s->groups)
        if (o.b && o.ids == g.ids) {
          g = o;
          o = Group();
          found = true;
          break;
        }
      if (!found) {
        std::vector<fz_case *> cs;
        for (size_t q : m) cs.push_back(s->active[q].cs);
        int rc = fz_case_batch_create(cs.data(), (int)cs.size(), &g.b);
        if (rc == FZ_OK && hipEventCreate(&g.e0) != hipSuccess) rc = FZ_ERR_HIP;
        if (rc == FZ_OK && hipEventCreate(&g.e1) != hipSuccess) rc = FZ_ERR_HIP;
        g.st = s->active[m[0]].ctx->st;
        if (rc != FZ_OK) {  // the cases step alone
          drop_group(g);
          continue;
        }
      }
      for (size_t q : m) gid[q] = (int)groups.size();
      groups.push_back(g);
    }
  }
  for (Group &o : s->groups)
    if (o.b) drop_group(o);  // groups whose members changed
  s->groups.clear();
  float tmax = 0.0f;
  for (const Running &r : s->active) tmax = std::max(tmax, r.last_ms);
  std::vector<int> want(s->active.size()), open(s->active.size());
  for (size_t q = 0; q < s->active.size(); q++) {
    Running &r = s->active[q];
    int d = 1;
    if (r.last_ms > 0.0f && tmax > 0.0f) d = std::max(1, std::min(s->max_batch, (int)(tmax / r.last_ms)));
    want[q] = gid[q] >= 0 ? 0 : std::min(d, std::max(r.left, 0));
    open[q] = want[q] > 0;
    r.nq = 0;
    r.rc = FZ_OK;
  }
  // An error in one case's step stops that case only: the others' steps go
  // on, and every step that was enqueued is waited for and ended below.
  const bool may_capture = s->hip_mu.try_lock();
  for (int k = 0; k < s->max_batch; k++) {
    for (size_t q = 0; q < s->active.size(); q++) {
      Running &r = s->active[q];
      if (!open[q] || r.nq >= want[q]) continue;
      fz_case_block_capture(r.cs, !may_capture);
      int pending = 0, rc = FZ_OK;
      if (hipEventRecord(r.ev[2 * r.nq], r.ctx->st) != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "hipEventRecord");
      if (rc == FZ_OK) rc = fz_case_step_begin(r.cs, &pending);
      if (rc == FZ_OK && hipEventRecord(r.ev[2 * r.nq + 1], r.ctx->st) != hipSuccess)
        rc = fz_set_error(FZ_ERR_HIP, "hipEventRecord");
      if (rc != FZ_OK) {
        r.rc = rc;
        r.err = std::string("step: ") + fz_last_error();
        open[q] = 0;
        continue;
      }
      r.nq++;
      if (pending) open[q] = 0;  // its checks set the next dt: no more steps this round
    }
  }
  if (may_capture) s->hip_mu.unlock();
  // the groups' batched steps: all enqueued, then each waited for and ended
  std::vector<float> gms(groups.size(), 0.0f);
  std::vector<int> grc(groups.size(), FZ_OK);
  std::vector<std::string> gerr(groups.size());
  for (size_t g = 0; g < groups.size(); g++) {
    Group &G = groups[g];
    int rc = hipEventRecord(G.e0, G.st) == hipSuccess ? FZ_OK : fz_set_error(FZ_ERR_HIP, "hipEventRecord");
    if (rc == FZ_OK) rc = fz_case_batch_step_begin(G.b, nullptr);
    if (rc == FZ_OK && hipEventRecord(G.e1, G.st) != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "hipEventRecord");
    grc[g] = rc;
    if (rc != FZ_OK) gerr[g] = std::string("batched step: ") + fz_last_error();
  }
  for (size_t g = 0; g < groups.size(); g++) {
    Group &G = groups[g];
    std::vector<int> killed(G.ids.size(), 0);
    int rc = grc[g];
    if (rc == FZ_OK && hipEventSynchronize(G.e1) != hipSuccess) rc = fz_set_error(FZ_ERR_HIP, "hipEventSynchronize");
    if (rc == FZ_OK) rc = fz_case_batch_step_end(G.b, killed.data());
    if (rc == FZ_OK && hipEventElapsedTime(&gms[g], G.e0, G.e1) != hipSuccess)
      rc = fz_set_error(FZ_ERR_HIP, "hipEventElapsedTime");
    const std::string err =
        rc == FZ_OK ? std::string() : (grc[g] != FZ_OK ? gerr[g] : std::string("batched step: ") + fz_last_error());
    for (size_t q = 0; q < na; q++) {
      if (gid[q] != (int)g) continue;
      Running &r = s->active[q];
      size_t k = 0;
      while (G.ids[k] != r.id) k++;
      if (rc != FZ_OK) {
        r.rc = rc;
        r.err = err;
        continue;
      }
      s->step_ms[r.id].push_back(gms[g]);
      s->all_ms.push_back(gms[g]);
      r.last_ms = gms[g];
      r.killed = r.killed || killed[k];
      r.left = killed[k] ? 0 : r.left - 1;
    }
  }
  s->groups.swap(groups);
  std::vector<Running> still;
  for (Running &r : s->active) {
    if (r.rc == FZ_OK && r.nq > 0) {  // stepped in this round (a request of 0 steps is not)
      int killed = 0;
      int rc = hipEventSynchronize(r.ev[2 * r.nq - 1]) == hipSuccess
                   ? FZ_OK
                   : fz_set_error(FZ_ERR_HIP, "hipEventSynchronize");
      if (rc == FZ_OK) rc = fz_case_step_end(r.cs, &killed);
      for (int q = 0; q < r.nq && rc == FZ_OK; q++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.ev[2 * q], r.ev[2 * q + 1]) != hipSuccess) {
          rc = fz_set_error(FZ_ERR_HIP, "hipEventElapsedTime");
          break;
        }
        s->step_ms[r.id].push_back(ms);
        s->all_ms.push_back(ms);
        r.last_ms = ms;
      }
      if (rc != FZ_OK) {
        r.rc = rc;
        r.err = std::string("step: ") + fz_last_error();
      }
      r.killed = r.killed || killed;
      r.left = killed ? 0 : r.left - r.nq;
    }
    if (r.rc != FZ_OK || r.left <= 0) {  // its group, if any, is made again without it
      for (Group &G : s->groups)
        if (G.b && std::find(G.ids.begin(), G.ids.end(), r.id) != G.ids.end()) drop_group(G);
    }
    if (r.rc != FZ_OK) {
      // the failed case is dropped (its stream drained first), the request
      // finishes as failed and its memory returns
      (void)hipStreamSynchronize(r.ctx->st);
      fz_case_block_capture(r.cs, false);
      drop_running(r);
      fail_request(s, r.id, r.bytes, r.err);
    } else if (r.left <= 0) {
      for (hipEvent_t &e : r.ev) (void)hipEventDestroy(e);
      fz_case_block_capture(r.cs, false);
      s->done[r.id] = Finished{r.cs, r.bytes, r.killed ? 1 : 0, 0, std::string()};
    } else {
      still.push_back(r);
    }
  }
  s->active.swap(still);
  if (n_active) *n_active = (int)s->active.size();
  if (n_queued) {
    std::lock_guard<std::mutex> lk(s->mu);
    *n_queued = (int)(s->queue.size() + s->preparing.size());
  }
  return FZ_OK;
}

// state of a request: 0 queued, 1 running, 2 finished (killed set when a
// check stopped it), 3 failed (fz_server_error has the message), -1 unknown
// or released
extern "C" int fz_server_status(fz_server *s, int id, int *state, int *killed) {
  if (!s || !state) return fz_set_error(FZ_ERR_ARG, "fz_server_status: null argument");
  *state = -1;
  if (killed) *killed = 0;
  for (const Request &q : s->queue)
    if (q.id == id) *state = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->preparing.count(id)) *state = 0;  // admitted, being set up
  }
  for (const Running &r : s->active)
    if (r.id == id) *state = 1;
  auto it = s->done.find(id);
  if (it != s->done.end()) {
    *state = it->second.failed ? 3 : 2;
    if (killed) *killed = it->second.killed;
  }
  return FZ_OK;
}

// the error message of a failed request (empty for any other)
extern "C" int fz_server_error(fz_server *s, int id, char *buf, size_t cap) {
  if (!s || !buf || cap == 0) return fz_set_error(FZ_ERR_ARG, "fz_server_error: bad argument");
  buf[0] = 0;
  auto it = s->done.find(id);
  if (it != s->done.end() && it->second.failed) snprintf(buf, cap, "%s", it->second.err.c_str());
  return FZ_OK;
}

// the case of a finished request (valid until fz_server_release)
extern "C" int fz_server_case(fz_server *s, int id, fz_case **cs) {
  if (!s || !cs) return fz_set_error(FZ_ERR_ARG, "fz_server_case: null argument");
  auto it = s->done.find(id);
  if (it == s->done.end()) return fz_set_error(FZ_ERR_STATE, "fz_server_case: request %d has not finished", id);
  if (it->second.failed)
    return fz_set_error(FZ_ERR_STATE, "fz_server_case: request %d failed: %s", id, it->second.err.c_str());
  *cs = it->second.cs;
  return FZ_OK;
}

extern "C" int fz_server_release(fz_server *s, int id) {
  if (!s) return fz_set_error(FZ_ERR_ARG, "fz_server_release: null server");
  auto it = s->done.find(id);
  if (it == s->done.end()) return fz_set_error(FZ_ERR_STATE, "fz_server_release: request %d has not finished", id);
  if (it->second.cs) fz_case_destroy(it->second.cs);
  s->in_use -= it->second.bytes;
  s->done.erase(it);
  return FZ_OK;
}

// step durations in ms: of request id, or of every step so far (id = 0);
// *count receives the total, at most cap are copied
extern "C" int fz_server_step_times(fz_server *s, int id, float *ms, int cap, int *count) {
  if (!s || !count) return fz_set_error(FZ_ERR_ARG, "fz_server_step_times: null argument");
  const std::vector<float> *v = &s->all_ms;
  std::vector<float> none;
  if (id != 0) {
    auto it = s->step_ms.find(id);
    v = it == s->step_ms.end() ? &none : &it->second;
  }
  *count = (int)v->size();
  if (ms)
    for (int q = 0; q < std::min(cap, *count); q++) ms[q] = (*v)[q];
  return FZ_OK;
}

extern "C" int fz_server_memory(fz_server *s, size_t *in_use, size_t *budget) {
  if (!s) return fz_set_error(FZ_ERR_ARG, "fz_server_memory: null server");
  if (in_use) *in_use = s->in_use;
  if (budget) *budget = s->budget;
  return FZ_OK;
}

// free and total device memory of the current device (for choosing a budget)
extern "C" int fz_device_memory(size_t *free_bytes, size_t *total_bytes) {
  size_t f = 0, t = 0;
  FZ_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return FZ_OK;
}
