// tp_serve.cpp -- request loop over one context: cases of different sizes
// arrive over time, are admitted while their device bytes fit an HBM budget
// and a slot is free, advance together in the same launches while others
// join and leave (continuous batching), and leave the device with their
// final state copied to the host.  Host code over the C ABI only; the device
// work is the context's.  No counterpart in the reference, which runs one
// case per MPI job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
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

struct Request {
  int cid = -1;                  // case handle in the context
  int nb = 0;
  int64_t points = 0;
  int nsteps = 0, steps = 0;     // requested and made
  int state = 0;                 // 0 queued, 1 on the device, 2 done, 3 failed
  double t_arr = 0, t_adm = -1, t_done = -1;
  int64_t bytes = 0;
  HostBuf qa;                    // final qa of every sub-block, concatenated
  std::string error;
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
  int64_t budget = 0, used = 0, peak_used = 0;
  int64_t fill = 0, active_points = 0;   // admission stops at fill points on the device (0: no limit)
  int max_active = 0, max_seen = 0;
  bool graph = true;
  std::chrono::steady_clock::time_point t0;
  std::vector<Request> reqs;
  std::vector<int> active;       // request ids on the device, in admission order
  int64_t admitted_while_busy = 0;   // admissions that joined running cases
  double busy_s = 0;             // wall time spent stepping
  double admit_s = 0, retire_s = 0, capture_s = 0;   // wall time of admissions, retirements, graph captures
  int64_t captures = 0;
  bool changed = true;           // membership changed since the last captured step
  std::string err;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

int srv_fail(tp_server *s, int code, const std::string &m) {
  s->err = m;
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
  for (int r : fin) {
    Request &q = s->reqs[r];
    q.t_done = s->now();
    q.state = 2;
    s->used -= q.bytes;
    s->active_points -= q.points;
    s->active.erase(std::find(s->active.begin(), s->active.end(), r));
  }
  return 0;
}

// admit arrived requests in arrival order while they fit and the cases on
// the device hold fewer than the fill target's points; a request that does
// not fit the budget is passed over for later, smaller ones.  The chosen
// requests join in one tp_admit_cases call (one rebuild of the launch tables).
int admit(tp_server *s, int *nadm) {
  *nadm = 0;
  const double t = s->now();
  std::vector<int> order;
  for (size_t r = 0; r < s->reqs.size(); r++)
    if (s->reqs[r].state == 0 && s->reqs[r].t_arr <= t) order.push_back((int)r);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return s->reqs[a].t_arr < s->reqs[b].t_arr; });
  const bool busy = !s->active.empty();
  std::vector<int> pick;
  std::vector<int32_t> ids;
  int64_t used = s->used, points = s->active_points;
  size_t nact = s->active.size();
  for (int r : order) {
    const Request &q = s->reqs[r];
    if ((int)nact >= s->max_active) break;
    if (s->fill > 0 && nact > 0 && points >= s->fill) break;
    if (used + q.bytes > s->budget) continue;
    pick.push_back(r);
    ids.push_back(q.cid);
    used += q.bytes;
    points += q.points;
    nact++;
  }
  if (pick.empty()) return 0;
  const double ta = s->now();
  std::vector<int32_t> st(pick.size(), 0);
  tp_admit_cases(s->ctx, (int32_t)ids.size(), ids.data(), st.data());
  std::vector<std::string> err(pick.size());
  for (size_t i = 0; i < pick.size(); i++) {
    if (st[i]) err[i] = tp_last_error();
    else if ((st[i] = tp_initialo(s->ctx, ids[i]))) err[i] = tp_last_error();
  }
  s->admit_s += s->now() - ta;
  s->changed = true;
  for (size_t i = 0; i < pick.size(); i++) {
    Request &q = s->reqs[pick[i]];
    if (st[i]) {
      q.state = 3;
      q.error = err[i];
      tp_remove_case(s->ctx, q.cid);
      continue;
    }
    q.state = 1;
    q.t_adm = s->now();
    s->used += q.bytes;
    s->active_points += q.points;
    s->peak_used = std::max(s->peak_used, s->used);
    s->active.push_back(pick[i]);
    s->max_seen = std::max(s->max_seen, (int)s->active.size());
    if (busy) s->admitted_while_busy++;
    (*nadm)++;
  }
  return 0;
}

int iterate(tp_server *s, int max_steps, int32_t *status) {
  int nadm = 0;
  if (int rc = admit(s, &nadm)) return rc;
  int stepped = 0;
  if (!s->active.empty()) {
    int k = max_steps;
    for (int r : s->active) k = std::min(k, s->reqs[r].nsteps - s->reqs[r].steps);
    k = std::max(k, 0);
    if (k > 0) {
      if (s->graph && s->changed) {
        // capture the step of the new set of cases (no replay yet)
        const double tc = s->now();
        if (int rc = tp_step_graph(s->ctx, 0)) return srv_fail(s, rc, std::string("capture: ") + tp_last_error());
        s->capture_s += s->now() - tc;
        s->captures++;
      }
      s->changed = false;
      const double t = s->now();
      int rc = s->graph ? tp_step_graph(s->ctx, k) : tp_step(s->ctx, k);
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
    for (const Request &q : s->reqs) {
      if (q.state == 0) (q.t_arr <= t ? queued : waiting)++;
      if (q.state >= 2) done++;
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
  s->budget = budget_bytes;
  s->max_active = max_active;
  s->graph = use_graph != 0;
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return 0;
}

int32_t tp_server_destroy(tp_server *s) {
  if (!s) return 0;
  tp_destroy(s->ctx);
  delete s;
  return 0;
}

tp_context *tp_server_context(tp_server *s) { return s ? s->ctx : nullptr; }

int32_t tp_server_set_fill(tp_server *s, int64_t points) {
  if (!s || points < 0) return TP_E_ARG;
  s->fill = points;
  return 0;
}

const char *tp_server_error(tp_server *s) { return s ? s->err.c_str() : "null server"; }

double tp_server_clock(tp_server *s) { return s ? s->now() : 0.0; }

int32_t tp_server_submit(tp_server *s, int32_t nblocks, const int32_t *dims, const int32_t *nbc,
                         const int32_t *ncd, const int32_t *nsz, int32_t njct, const int32_t *jct,
                         const int32_t *jvel, const int32_t *iparams, const double *dparams,
                         const double *const *xyz, int32_t nsteps, double t_arrival, int32_t *req) {
  if (!s || !xyz || !req || nsteps < 1) return s ? srv_fail(s, TP_E_ARG, "tp_server_submit: bad arguments") : TP_E_ARG;
  Request q;
  if (int rc = tp_add_case(s->ctx, nblocks, dims, nbc, ncd, nsz, njct, jct, jvel, iparams, dparams, &q.cid))
    return srv_fail(s, rc, tp_last_error());
  for (int b = 0; b < nblocks; b++) {
    if (!xyz[b]) {
      tp_remove_case(s->ctx, q.cid);
      return srv_fail(s, TP_E_ARG, "tp_server_submit: missing grid");
    }
    if (int rc = tp_set_grid(s->ctx, q.cid, b, xyz[b])) {
      tp_remove_case(s->ctx, q.cid);
      return srv_fail(s, rc, tp_last_error());
    }
  }
  if (int rc = tp_case_bytes(s->ctx, q.cid, &q.bytes)) {
    tp_remove_case(s->ctx, q.cid);
    return srv_fail(s, rc, tp_last_error());
  }
  if (q.bytes > s->budget) {
    tp_remove_case(s->ctx, q.cid);
    return srv_fail(s, TP_E_SIZE, "case needs " + std::to_string(q.bytes) + " device bytes, over the budget of " +
                                      std::to_string(s->budget));
  }
  q.nb = nblocks;
  q.nsteps = nsteps;
  for (int b = 0; b < nblocks; b++)
    q.points += (int64_t)(dims[3 * b] + 1) * (dims[3 * b + 1] + 1) * (dims[3 * b + 2] + 1);
  q.t_arr = t_arrival >= 0 ? t_arrival : s->now();
  *req = (int32_t)s->reqs.size();
  s->reqs.push_back(std::move(q));
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
    if (st[0] == 0 && st[1] == 0 && st[2] == 0) return 0;
    if (st[0] == 0 && st[1] > 0 && st[4] == 0 && st[5] == 0) {
      // nothing ran, nothing is on the device and nothing was admissible: a
      // request cannot fit
      return srv_fail(s, TP_E_SIZE, "queued requests do not fit the free slots and budget");
    }
    if (st[0] == 0 && st[1] == 0 && st[2] > 0) {
      // idle until the next arrival
      double next = 1e300;
      for (const Request &q : s->reqs)
        if (q.state == 0) next = std::min(next, q.t_arr);
      const double wait = next - s->now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1.0)));
    }
  }
}

int32_t tp_server_result(tp_server *s, int32_t req, double *qa, int64_t count, double *info6) {
  if (!s || req < 0 || req >= (int)s->reqs.size()) return s ? srv_fail(s, TP_E_ARG, "bad request") : TP_E_ARG;
  const Request &q = s->reqs[req];
  if (info6) {
    info6[0] = q.state;
    info6[1] = q.t_arr;
    info6[2] = q.t_adm;
    info6[3] = q.t_done;
    info6[4] = q.steps;
    info6[5] = (double)q.bytes;
  }
  if (qa) {
    if (q.state != 2) return srv_fail(s, TP_E_STATE, "request not finished");
    if (count != (int64_t)q.qa.n)
      return srv_fail(s, TP_E_SIZE, "expected " + std::to_string(q.qa.n) + " values");
    if (q.qa.n) std::memcpy(qa, q.qa.p, sizeof(double) * q.qa.n);
  }
  return 0;
}

int64_t tp_server_result_count(tp_server *s, int32_t req) {
  if (!s || req < 0 || req >= (int)s->reqs.size()) return -1;
  return (int64_t)s->reqs[req].qa.n;
}

int32_t tp_server_stats(tp_server *s, double *out16) {
  if (!s || !out16) return TP_E_ARG;
  double *out12 = out16;
  std::vector<double> tts;
  double first = 1e300, last = 0;
  int failed = 0;
  for (const Request &q : s->reqs) {
    if (q.state == 3) failed++;
    if (q.state != 2) continue;
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

}  // extern "C"
