// tp_serve.cpp -- request loop over one context: cases of different sizes
// arrive over time, are set up while their device bytes fit an HBM budget,
// are admitted while a slot is free, advance together in the same launches
// while others join and leave (continuous batching), and leave the device
// with their final state copied to the host.  Set-up runs on a staging
// thread, each request in a context (and stream) of its own, beside the
// stepping of the resident cases; a set-up case joins them between steps
// without a copy (tp_adopt_cases).  Host code over the C ABI only; the
// device work is the contexts'.  No counterpart in the reference, which runs
// one case per MPI job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tandem_hip.h"

namespace {

// host buffer of page-locked memory (hipHostMalloc), which the device writes
// at full link speed: a fresh pageable buffer takes a page fault per page
// during the copy, about 4x slower for a state of hundreds of MB.  Plain heap
// memory if the page-locked allocation fails.
struct HostBuf {
  double *p = nullptr;
  size_t n = 0;
  bool pinned = false;
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  HostBuf(HostBuf &&o) noexcept { swap(o); }
  HostBuf &operator=(HostBuf &&o) noexcept {
    swap(o);
    return *this;
  }
  ~HostBuf() { reset(); }
  void swap(HostBuf &o) noexcept {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(pinned, o.pinned);
  }
  void reset() {
    if (p) {
      if (pinned) (void)hipHostFree(p);
      else delete[] p;
    }
    p = nullptr;
    n = 0;
    pinned = false;
  }
  bool alloc(size_t count) {
    reset();
    if (count == 0) return true;
    void *q = nullptr;
    if (hipHostMalloc(&q, sizeof(double) * count, hipHostMallocDefault) == hipSuccess) {
      p = static_cast<double *>(q);
      pinned = true;
    } else {
      (void)hipGetLastError();
      p = new (std::nothrow) double[count];
      if (!p) return false;
    }
    n = count;
    return true;
  }
};

// request states: queued, on the device, done, failed, being set up, set up
enum { RQ_QUEUED = 0, RQ_ACTIVE = 1, RQ_DONE = 2, RQ_FAILED = 3, RQ_STAGING = 4, RQ_STAGED = 5 };

struct Request {
  tp_context *sctx = nullptr;    // its own context until it joins the server's
  int cid = -1;                  // case handle in sctx, then in the server's context
  int nb = 0;
  int64_t points = 0;
  int nsteps = 0, steps = 0;     // requested and made
  int state = RQ_QUEUED;
  double t_arr = 0, t_adm = -1, t_done = -1;
  int64_t bytes = 0;
  HostBuf qa;                    // final qa of every sub-block, concatenated
  std::string error;
};

// the request table: requests never move, so a request's slot stays valid
// while others are appended (a client thread submitting while the loop
// runs); appends happen under the server's mutex, and the count is published
// after the slot is filled
class ReqTable {
 public:
  static constexpr size_t kChunk = 1024, kChunks = 4096;   // up to 4 M requests
  size_t size() const { return n_.load(std::memory_order_acquire); }
  Request &operator[](size_t i) { return chunks_[i / kChunk][i % kChunk]; }
  const Request &operator[](size_t i) const { return chunks_[i / kChunk][i % kChunk]; }
  bool push_back(Request &&q) {
    const size_t i = n_.load(std::memory_order_relaxed);
    if (i >= kChunk * kChunks) return false;
    if (!chunks_[i / kChunk]) chunks_[i / kChunk].reset(new Request[kChunk]);
    chunks_[i / kChunk][i % kChunk] = std::move(q);
    n_.store(i + 1, std::memory_order_release);
    return true;
  }
  template <class T, class R>
  struct Iter {
    T *t;
    size_t i;
    R &operator*() const { return (*t)[i]; }
    Iter &operator++() { ++i; return *this; }
    bool operator!=(const Iter &o) const { return i != o.i; }
  };
  Iter<ReqTable, Request> begin() { return {this, 0}; }
  Iter<ReqTable, Request> end() { return {this, size()}; }
  Iter<const ReqTable, const Request> begin() const { return {this, 0}; }
  Iter<const ReqTable, const Request> end() const { return {this, size()}; }

 private:
  std::unique_ptr<Request[]> chunks_[kChunks];
  std::atomic<size_t> n_{0};
};

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  // nearest rank
  size_t r = (size_t)std::ceil(q * (double)v.size());
  if (r < 1) r = 1;
  return v[std::min(r, v.size()) - 1];
}

}  // namespace

struct tp_server {
  tp_context *ctx = nullptr;
  int device = 0;
  // what the staging threads, the loop and submitting threads share, under
  // mu: the request table (appended to by submissions), the requests'
  // states, the bytes held by staged requests
  std::mutex mu;
  std::condition_variable cv;
  std::vector<std::thread> stagers;
  int nstagers = 2;              // staging threads
  // shared by the staging threads (and submissions) while they set requests
  // up, exclusive for the loop while it captures a step graph: a capture
  // must not overlap the set-up's allocations and synchronous copies, which
  // would invalidate it
  std::shared_mutex hip_mu;
  double gather_s = 0.05;        // an idle device waits for requests arriving within this of the first
  bool stop = false;
  int ahead = 1, staged_now = 0, staged_peak = 0;
  int64_t staged_bytes = 0, staged_count = 0, stage_failed = 0;
  double stage_s = 0;
  int64_t budget = 0, used = 0, peak_used = 0;
  int64_t fill = 0, active_points = 0;   // admission stops at fill points on the device (0: no limit)
  int max_active = 0, max_seen = 0;
  bool graph = true;
  std::chrono::steady_clock::time_point t0;
  ReqTable reqs;
  std::vector<int> active;       // request ids on the device, in admission order
  int64_t admitted_while_busy = 0;   // admissions that joined running cases
  double busy_s = 0;             // wall time spent stepping
  double admit_s = 0, retire_s = 0, capture_s = 0;   // wall time of admissions, retirements, graph captures
  int64_t captures = 0;
  bool changed = true;           // membership changed since the last captured step
  int64_t eager_iters = 0;       // iterations launched eagerly while a set-up held off the capture
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// the message of the last failure of a server call on this thread (the
// loop, a staging thread and a submitting thread may fail at once)
thread_local std::string g_srv_err;

int srv_fail(tp_server *, int code, const std::string &m) {
  g_srv_err = m;
  return code;
}

int64_t qa_count(tp_context *ctx, int cid, int nb) {
  int64_t n = 0;
  for (int b = 0; b < nb; b++) n += tp_array_count(ctx, cid, b, TP_QA);
  return n;
}

// copy the finished cases' states to the host and take them off the device
// (one rebuild of the launch tables for all of them)
int retire(tp_server *s, const std::vector<int> &fin) {
  std::vector<int32_t> ids;
  for (int r : fin) {
    Request &q = s->reqs[r];
    if (!q.qa.alloc((size_t)qa_count(s->ctx, q.cid, q.nb)))
      return srv_fail(s, TP_E_SIZE, "retire: no host memory for a final state");
    int64_t off = 0;
    for (int b = 0; b < q.nb; b++) {
      const int64_t n = tp_array_count(s->ctx, q.cid, b, TP_QA);
      if (int rc = tp_download(s->ctx, q.cid, b, TP_QA, q.qa.p + off, n))
        return srv_fail(s, rc, std::string("retire: ") + tp_last_error());
      off += n;
    }
    ids.push_back(q.cid);
  }
  if (int rc = tp_remove_cases(s->ctx, (int32_t)ids.size(), ids.data()))
    return srv_fail(s, rc, std::string("retire: ") + tp_last_error());
  s->changed = true;
  std::lock_guard<std::mutex> lk(s->mu);
  for (int r : fin) {
    Request &q = s->reqs[r];
    q.t_done = s->now();
    q.state = RQ_DONE;
    s->used -= q.bytes;
    s->active_points -= q.points;
    s->active.erase(std::find(s->active.begin(), s->active.end(), r));
  }
  s->cv.notify_all();   // room in the budget for staging
  return 0;
}

// a staging thread (there are nstagers): takes arrived requests in arrival
// order and sets each up in its own context (build, geometry, initial
// state), while the device bytes of the cases on the device and of those set
// up fit the budget and fewer than `ahead` requests wait set up
void stager_main(tp_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    const double t = s->now();
    int pick = -1;
    double next = 1e300;
    if (s->staged_now < s->ahead) {
      for (size_t r = 0; r < s->reqs.size(); r++) {
        const Request &q = s->reqs[r];
        if (q.state != RQ_QUEUED) continue;
        if (q.t_arr > t) { next = std::min(next, q.t_arr); continue; }
        if (s->used + s->staged_bytes + q.bytes > s->budget) continue;   // passed over for smaller ones
        if (pick < 0 || q.t_arr < s->reqs[pick].t_arr) pick = (int)r;
      }
    }
    if (pick < 0) {
      if (next < 1e300) s->cv.wait_for(lk, std::chrono::duration<double>(std::min(next - t, 0.05)));
      else s->cv.wait_for(lk, std::chrono::milliseconds(50));
      continue;
    }
    Request &q = s->reqs[pick];
    q.state = RQ_STAGING;
    s->staged_bytes += q.bytes;
    s->staged_now++;
    s->staged_peak = std::max(s->staged_peak, s->staged_now);
    tp_context *sc = q.sctx;
    const int cid = q.cid;
    lk.unlock();
    const double t0 = s->now();
    // the line-solver options of the server's context, so that the set-up
    // (metrics) and the steps use the same solves
    int rc = 0;
    {
      std::shared_lock<std::shared_mutex> hl(s->hip_mu);
      for (int opt : {TP_OPT_EXACT_LINES, TP_OPT_CHUNKS, TP_OPT_STRESSES}) {
        int64_t v = 0;
        if (!rc) rc = tp_get_option(s->ctx, opt, &v);
        if (!rc) rc = tp_set_option(sc, opt, v);
      }
      if (!rc) rc = tp_admit(sc, cid);
      if (!rc) rc = tp_initialo(sc, cid);
      if (!rc) rc = tp_sync(sc);
    }
    const std::string e = rc ? tp_last_error() : "";
    const double t1 = s->now();
    lk.lock();
    s->stage_s += t1 - t0;
    Request &q2 = s->reqs[pick];
    if (rc) {
      q2.state = RQ_FAILED;
      q2.error = "set-up: " + e;
      s->staged_bytes -= q2.bytes;
      s->staged_now--;
      s->stage_failed++;
    } else {
      q2.state = RQ_STAGED;
      s->staged_count++;
    }
    s->cv.notify_all();
  }
}

void start_stager(tp_server *s) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stagers.empty() && !s->stop)
    for (int i = 0; i < s->nstagers; i++) s->stagers.emplace_back(stager_main, s);
}

// an idle device takes the requests that arrive together: while requests
// that arrived within gather_s of the earliest set-up one are still queued
// (and fit the budget) or being set up, and those set up fill neither the
// slots, nor the fill target, nor the staging window, admission waits for
// them (a burst then runs as one batch; a stream is not held back).  A device
// that is stepping never waits: it takes what is set up between steps.
bool gathering(const tp_server *s) {
  const double t = s->now();
  int64_t pts = 0;
  int nst = 0;
  double first = 1e300;
  for (const Request &q : s->reqs)
    if (q.state == RQ_STAGED) { pts += q.points; nst++; first = std::min(first, q.t_arr); }
  if (nst == 0) return false;
  bool pending = false;
  for (const Request &q : s->reqs) {
    if (q.t_arr > first + s->gather_s) continue;   // a later arrival: not waited for
    if (q.state == RQ_STAGING) pending = true;
    else if (q.state == RQ_QUEUED && q.t_arr <= t && s->used + s->staged_bytes + q.bytes <= s->budget)
      pending = true;
  }
  if (!pending) return false;
  return nst < s->max_active && nst < s->ahead && !(s->fill > 0 && pts >= s->fill);
}

// admit set-up requests in arrival order while a slot is free and the cases
// on the device hold fewer than the fill target's points (one case is always
// admitted onto an idle device); the chosen ones join in one
// tp_adopt_cases call (one rebuild of the launch tables)
int admit(tp_server *s, int *nadm) {
  *nadm = 0;
  std::vector<int> pick;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->active.empty() && gathering(s)) return 0;
    std::vector<int> order;
    for (size_t r = 0; r < s->reqs.size(); r++)
      if (s->reqs[r].state == RQ_STAGED) order.push_back((int)r);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return s->reqs[a].t_arr < s->reqs[b].t_arr; });
    int64_t points = s->active_points;
    size_t nact = s->active.size();
    for (int r : order) {
      if ((int)nact >= s->max_active) break;
      if (s->fill > 0 && nact > 0 && points >= s->fill) break;
      pick.push_back(r);
      points += s->reqs[r].points;
      nact++;
    }
  }
  if (pick.empty()) return 0;
  const bool busy = !s->active.empty();
  const double ta = s->now();
  std::vector<tp_context *> srcs;
  std::vector<int32_t> ids, nids(pick.size(), -1);
  for (int r : pick) { srcs.push_back(s->reqs[r].sctx); ids.push_back(s->reqs[r].cid); }
  if (int rc = tp_adopt_cases(s->ctx, (int32_t)pick.size(), srcs.data(), ids.data(), nids.data()))
    return srv_fail(s, rc, std::string("admit: ") + tp_last_error());
  for (tp_context *c : srcs) tp_destroy(c);
  s->admit_s += s->now() - ta;
  s->changed = true;
  std::lock_guard<std::mutex> lk(s->mu);
  for (size_t i = 0; i < pick.size(); i++) {
    Request &q = s->reqs[pick[i]];
    q.sctx = nullptr;
    q.cid = nids[i];
    q.state = RQ_ACTIVE;
    q.t_adm = s->now();
    s->staged_bytes -= q.bytes;
    s->staged_now--;
    s->used += q.bytes;
    s->active_points += q.points;
    s->peak_used = std::max(s->peak_used, s->used + s->staged_bytes);
    s->active.push_back(pick[i]);
    s->max_seen = std::max(s->max_seen, (int)s->active.size());
    if (busy) s->admitted_while_busy++;
    (*nadm)++;
  }
  s->cv.notify_all();
  return 0;
}

int iterate(tp_server *s, int max_steps, int32_t *status) {
  start_stager(s);
  int nadm = 0;
  if (int rc = admit(s, &nadm)) return rc;
  int stepped = 0;
  if (!s->active.empty()) {
    int k = max_steps;
    for (int r : s->active) k = std::min(k, s->reqs[r].nsteps - s->reqs[r].steps);
    k = std::max(k, 0);
    if (k > 0) {
      bool replay = s->graph && !s->changed;
      if (s->graph && s->changed) {
        // capture the step of the new set of cases (no replay yet) unless a
        // request is being set up: its allocations and copies must not
        // overlap a capture, so these steps launch eagerly and the capture
        // waits for an iteration without set-up
        std::unique_lock<std::shared_mutex> hl(s->hip_mu, std::try_to_lock);
        if (hl.owns_lock()) {
          const double tc = s->now();
          if (int rc = tp_step_graph(s->ctx, 0))
            return srv_fail(s, rc, std::string("capture: ") + tp_last_error());
          s->capture_s += s->now() - tc;
          s->captures++;
          s->changed = false;
          replay = true;
        } else {
          s->eager_iters++;
        }
      }
      const double t = s->now();
      int rc = replay ? tp_step_graph(s->ctx, k) : tp_step(s->ctx, k);
      if (!rc) rc = tp_sync(s->ctx);
      if (rc) return srv_fail(s, rc, std::string("step: ") + tp_last_error());
      s->busy_s += s->now() - t;
      for (int r : s->active) s->reqs[r].steps += k;
      stepped = k;
    }
    std::vector<int> fin;
    for (int r : s->active)
      if (s->reqs[r].steps >= s->reqs[r].nsteps) fin.push_back(r);
    const double tr = s->now();
    if (!fin.empty())
      if (int rc = retire(s, fin)) return rc;
    s->retire_s += s->now() - tr;
  }
  if (status) {
    int queued = 0, waiting = 0, done = 0;
    const double t = s->now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (const Request &q : s->reqs) {
      if (q.state == RQ_QUEUED) (q.t_arr <= t ? queued : waiting)++;
      if (q.state == RQ_STAGING || q.state == RQ_STAGED) queued++;
      if (q.state == RQ_DONE || q.state == RQ_FAILED) done++;
    }
    status[0] = (int32_t)s->active.size();
    status[1] = queued;
    status[2] = waiting;
    status[3] = done;
    status[4] = nadm;
    status[5] = stepped;
  }
  return 0;
}

}  // namespace

extern "C" {

int32_t tp_server_create(int64_t budget_bytes, int32_t max_active, int32_t use_graph, tp_server **out) {
  if (!out || max_active < 1) return TP_E_ARG;
  tp_context *ctx = nullptr;
  if (int rc = tp_create(&ctx)) return rc;
  if (int rc = tp_finalize(ctx)) {
    tp_destroy(ctx);
    return rc;
  }
  if (budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      tp_destroy(ctx);
      return TP_E_HIP;
    }
    budget_bytes = (int64_t)(0.9 * (double)fr);
  }
  tp_server *s = new tp_server();
  s->ctx = ctx;
  (void)hipGetDevice(&s->device);
  s->budget = budget_bytes;
  s->max_active = max_active;
  s->ahead = max_active;   // set up as many ahead as there are slots (within the budget)
  s->graph = use_graph != 0;
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return 0;
}

int32_t tp_server_destroy(tp_server *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  for (std::thread &t : s->stagers) t.join();
  s->stagers.clear();
  for (Request &q : s->reqs)
    if (q.sctx) tp_destroy(q.sctx);
  tp_destroy(s->ctx);
  delete s;
  return 0;
}

int32_t tp_server_set_staging(tp_server *s, int32_t ahead) {
  if (!s || ahead < 1) return TP_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->ahead = ahead;
  s->cv.notify_all();
  return 0;
}

tp_context *tp_server_context(tp_server *s) { return s ? s->ctx : nullptr; }

int32_t tp_server_set_fill(tp_server *s, int64_t points) {
  if (!s || points < 0) return TP_E_ARG;
  s->fill = points;
  return 0;
}

const char *tp_server_error(tp_server *s) { return s ? g_srv_err.c_str() : "null server"; }

double tp_server_clock(tp_server *s) { return s ? s->now() : 0.0; }

int32_t tp_server_submit(tp_server *s, int32_t nblocks, const int32_t *dims, const int32_t *nbc,
                         const int32_t *ncd, const int32_t *nsz, int32_t njct, const int32_t *jct,
                         const int32_t *jvel, const int32_t *iparams, const double *dparams,
                         const double *const *xyz, int32_t nsteps, double t_arrival, int32_t *req) {
  if (!s || !xyz || !req || nsteps < 1) return s ? srv_fail(s, TP_E_ARG, "tp_server_submit: bad arguments") : TP_E_ARG;
  const double t_call = s->now();   // the arrival of a request submitted "now"
  for (int b = 0; b < nblocks; b++)
    if (!xyz[b]) return srv_fail(s, TP_E_ARG, "tp_server_submit: missing grid");
  // the request's own context, where the staging thread sets it up (its
  // HIP calls kept apart from a graph capture of the loop, as the stager's)
  std::shared_lock<std::shared_mutex> hl(s->hip_mu);
  Request q;
  if (int rc = tp_create(&q.sctx)) return srv_fail(s, rc, tp_last_error());
  auto drop = [&](int rc, const std::string &m) {
    tp_destroy(q.sctx);
    return srv_fail(s, rc, m);
  };
  if (int rc = tp_finalize(q.sctx)) return drop(rc, tp_last_error());
  if (int rc = tp_add_case(q.sctx, nblocks, dims, nbc, ncd, nsz, njct, jct, jvel, iparams, dparams, &q.cid))
    return drop(rc, tp_last_error());
  for (int b = 0; b < nblocks; b++)
    if (int rc = tp_set_grid(q.sctx, q.cid, b, xyz[b])) return drop(rc, tp_last_error());
  if (int rc = tp_case_bytes(q.sctx, q.cid, &q.bytes)) return drop(rc, tp_last_error());
  if (q.bytes > s->budget)
    return drop(TP_E_SIZE, "case needs " + std::to_string(q.bytes) + " device bytes, over the budget of " +
                               std::to_string(s->budget));
  q.nb = nblocks;
  q.nsteps = nsteps;
  for (int b = 0; b < nblocks; b++)
    q.points += (int64_t)(dims[3 * b] + 1) * (dims[3 * b + 1] + 1) * (dims[3 * b + 2] + 1);
  q.t_arr = t_arrival >= 0 ? t_arrival : t_call;
  hl.unlock();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const int32_t id = (int32_t)s->reqs.size();
    if (!s->reqs.push_back(std::move(q))) {
      tp_destroy(q.sctx);
      return srv_fail(s, TP_E_SIZE, "tp_server_submit: the request table is full");
    }
    *req = id;
    s->cv.notify_all();
  }
  start_stager(s);   // set-up starts at arrival, not at the next iteration
  return 0;
}

int32_t tp_server_iterate(tp_server *s, int32_t max_steps, int32_t *status6) {
  if (!s || max_steps < 1) return TP_E_ARG;
  return iterate(s, max_steps, status6);
}

int32_t tp_server_run(tp_server *s, int32_t max_steps) {
  if (!s || max_steps < 1) return TP_E_ARG;
  for (;;) {
    int32_t st[6];
    if (int rc = iterate(s, max_steps, st)) return rc;
    if (st[0] > 0) continue;   // cases on the device: keep stepping
    // nothing on the device: wait for a set-up request or the end
    std::unique_lock<std::mutex> lk(s->mu);
    bool open = false, staged = false;
    for (const Request &q : s->reqs) {
      if (q.state == RQ_QUEUED || q.state == RQ_STAGING) open = true;
      if (q.state == RQ_STAGED) staged = true;
    }
    if (!open && !staged) return 0;
    if (!staged && s->staged_now == 0 && s->used == 0) {
      // nothing runs and nothing is being set up: is any arrived request
      // able to fit the budget at all?  (submission refuses larger ones)
      bool waiting = false;
      const double t = s->now();
      for (const Request &q : s->reqs)
        if (q.state == RQ_QUEUED && (q.t_arr > t || q.bytes <= s->budget)) waiting = true;
      if (!waiting) return srv_fail(s, TP_E_SIZE, "queued requests do not fit the budget");
    }
    // woken by the staging thread when a request is set up
    s->cv.wait_for(lk, std::chrono::milliseconds(staged ? 5 : 20));
  }
}

int32_t tp_server_result(tp_server *s, int32_t req, double *qa, int64_t count, double *info6) {
  if (!s || req < 0 || req >= (int)s->reqs.size()) return s ? srv_fail(s, TP_E_ARG, "bad request") : TP_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  const Request &q = s->reqs[req];
  if (info6) {
    info6[0] = q.state == RQ_STAGING || q.state == RQ_STAGED ? RQ_QUEUED : q.state;
    info6[1] = q.t_arr;
    info6[2] = q.t_adm;
    info6[3] = q.t_done;
    info6[4] = q.steps;
    info6[5] = (double)q.bytes;
  }
  if (qa) {
    if (q.state != RQ_DONE) return srv_fail(s, TP_E_STATE, "request not finished");
    if (count != (int64_t)q.qa.n)
      return srv_fail(s, TP_E_SIZE, "expected " + std::to_string(q.qa.n) + " values");
    if (q.qa.n) std::memcpy(qa, q.qa.p, sizeof(double) * q.qa.n);
  }
  return 0;
}

int64_t tp_server_result_count(tp_server *s, int32_t req) {
  if (!s || req < 0) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  if (req >= (int)s->reqs.size()) return -1;
  return (int64_t)s->reqs[req].qa.n;
}

int32_t tp_server_stats(tp_server *s, double *out16) {
  if (!s || !out16) return TP_E_ARG;
  double *out12 = out16;
  std::vector<double> tts;
  double first = 1e300, last = 0;
  int failed = 0;
  std::lock_guard<std::mutex> lk(s->mu);
  for (const Request &q : s->reqs) {
    if (q.state == RQ_FAILED) failed++;
    if (q.state != RQ_DONE) continue;
    tts.push_back(q.t_done - q.t_arr);
    first = std::min(first, q.t_arr);
    last = std::max(last, q.t_done);
  }
  const double wall = tts.empty() ? 0.0 : last - first;
  double mean = 0;
  for (double v : tts) mean += v;
  if (!tts.empty()) mean /= (double)tts.size();
  out12[0] = (double)tts.size();
  out12[1] = failed;
  out12[2] = percentile(tts, 0.50);
  out12[3] = percentile(tts, 0.99);
  out12[4] = mean;
  out12[5] = tts.empty() ? 0.0 : *std::max_element(tts.begin(), tts.end());
  out12[6] = wall;
  out12[7] = wall > 0 ? 3600.0 * (double)tts.size() / wall : 0.0;
  out12[8] = (double)s->peak_used;
  out12[9] = (double)s->max_seen;
  out12[10] = (double)s->admitted_while_busy;
  out12[11] = s->busy_s;
  out16[12] = s->admit_s;
  out16[13] = s->retire_s;
  out16[14] = s->capture_s;
  out16[15] = (double)s->captures;
  return 0;
}

int32_t tp_server_stats2(tp_server *s, double *out6) {
  if (!s || !out6) return TP_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  out6[0] = s->stage_s;
  out6[1] = (double)s->staged_count;
  out6[2] = (double)s->stage_failed;
  out6[3] = (double)s->staged_peak;
  out6[4] = (double)s->eager_iters;
  out6[5] = (double)s->ahead;
  return 0;
}

}  // extern "C"
