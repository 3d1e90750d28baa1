// gfr_server.cpp — reconstruction serving loop: requests of any geometry,
// mode and count level arrive at any time; each is admitted as a
// device-resident session (gfr_recon) when a slot and HBM budget are free,
// and every admitted session advances pass by pass on its own stream, so the
// passes of different requests overlap on the device (continuous batching: a
// request joins the running set between two passes, it does not wait for the
// set to drain). The reference reconstructs one study per process run
// (GFIRST.cu:577-621); a request here is one such run.
//
// Threads: the stepper enqueues passes (never blocks on the device: at most
// `inflight` passes of a session are queued ahead of the device), the
// admitter creates sessions (uploads, mask), the retirer reads results back
// and frees sessions. Each request's result depends only on its own inputs:
// the same request alone or in any mix gives the same bits.
#include <math.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "gfr_internal.h"

namespace {

struct Req {
  long long id = 0;
  gfr_geometry g{};
  gfr_recon_params p{};
  const float *d2 = nullptr, *t2 = nullptr, *norm = nullptr;
  const int32_t* rec = nullptr;
  long long nrec = 0;
  std::atomic<int> state{GFR_REQ_PENDING};
  std::atomic<bool> cancel{false};  // set by gfr_server_cancel while admitted
  int error = GFR_OK;
  std::string msg;
  double t_submit = 0, t_admit = 0, t_done = 0;
  long long passes = 0;
  long long bytes = 0;
  gfr_recon* r = nullptr;
  std::deque<hipEvent_t> evs;  // passes queued ahead of the device
  bool last = false;           // final pass enqueued
  std::vector<float> xD, xT;
};

}  // namespace

struct gfr_server {
  gfr_ctx* ctx = nullptr;
  int max_active = 0, inflight = 2;
  int creating = 0;       // sessions being created by admitters
  long long reserved = 0; // budget bound of the sessions being created
  int retiring = 0;       // sessions being read back and freed
  bool mem_wait = false;  // a request did not fit: admit nothing until a retirement
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_admit, cv_step, cv_retire, cv_done;
  std::map<long long, std::shared_ptr<Req>> reqs;
  std::deque<std::shared_ptr<Req>> pending, retire;
  std::vector<std::shared_ptr<Req>> active;
  long long next_id = 1;
  bool stop = false;          // admitter and stepper
  bool stop_retire = false;   // retirer (after the last sessions are queued)
  long long n_done = 0, n_failed = 0, n_passes = 0;
  std::vector<hipEvent_t> ev_pool;  // stepper-owned
  std::thread th_step, th_retire;
  std::vector<std::thread> th_admit;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

hipEvent_t take_ev(gfr_server* s) {
  if (!s->ev_pool.empty()) {
    hipEvent_t e = s->ev_pool.back();
    s->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

void fail_req(gfr_server* s, const std::shared_ptr<Req>& q, int code, const std::string& msg) {
  q->error = code;
  q->msg = msg;
  q->t_done = s->now();
  q->state = GFR_REQ_FAILED;
  s->n_failed++;
}

// Upper bound of the device bytes a request's session holds: the session
// plan with both backgrounds, plus the VLOR lists in mode 3 (per record:
// two indices, two event values, up to two bins with list offsets and two
// list entries; BLANKT per bin).
long long request_bound(const std::shared_ptr<Req>& q) {
  const long long f = gfr_recon_footprint(&q->g, &q->p);
  if (f < 0) return 0;  // invalid parameters: creation reports them
  long long v = 0;
  if (q->p.mode == 3) {
    gfr_geometry_info inf{};
    if (gfr_geometry_check(&q->g, &inf) == GFR_OK) v = 48 * q->nrec + 4 * inf.ndata + 64;
  }
  return f + v;
}

// Creates sessions for pending requests in arrival order as slots and
// budget allow (several admitters run at once, so one large upload does not
// hold back the requests behind it); a request that cannot fit even on an
// idle server fails with GFR_ENOMEM instead of blocking the queue.
void admitter(gfr_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    if (s->pending.empty() || s->mem_wait ||
        (int)s->active.size() + s->creating >= s->max_active) {
      s->cv_admit.wait_for(lk, std::chrono::milliseconds(50));
      if (s->mem_wait && s->active.empty() && s->retire.empty() && s->creating == 0 &&
          s->retiring == 0)
        s->mem_wait = false;  // nothing left to retire: retry (and fail if still too large)
      continue;
    }
    std::shared_ptr<Req> q = s->pending.front();
    // Budget decisions are taken here, in arrival order: the head is started
    // only if an upper bound of its device bytes fits next to what is in use
    // and what the sessions being created may take.
    const long long bound = request_bound(q);
    size_t budget = 0, in_use = 0;
    {
      std::lock_guard<std::mutex> lc(s->ctx->mu);
      budget = s->ctx->budget;
      in_use = s->ctx->in_use;
    }
    const bool idle = s->active.empty() && s->retire.empty() && s->creating == 0 &&
                      s->retiring == 0;
    if (budget && (long long)in_use + s->reserved + bound > (long long)budget && !idle) {
      s->mem_wait = true;
      continue;
    }
    s->pending.pop_front();
    s->creating++;
    s->reserved += bound;
    lk.unlock();
    gfr_recon* r = nullptr;
    int rc = gfr_recon_create(s->ctx, &q->g, &q->p, q->d2, q->t2, q->norm, q->rec, q->nrec, &r);
    std::string msg = rc == GFR_OK ? std::string() : std::string(gfr_last_error(s->ctx));
    lk.lock();
    s->creating--;
    s->reserved -= bound;
    if (rc == GFR_ENOMEM &&
        !(s->active.empty() && s->retire.empty() && s->creating == 0 && s->retiring == 0)) {
      // memory frees up when running sessions retire: back into the queue in
      // arrival order, and no later request is admitted before it
      auto it = s->pending.begin();
      while (it != s->pending.end() && (*it)->id < q->id) ++it;
      s->pending.insert(it, q);
      s->mem_wait = true;
      continue;
    }
    if (rc != GFR_OK) {
      fail_req(s, q, rc, msg);
      s->cv_done.notify_all();
      s->cv_admit.notify_all();
      continue;
    }
    q->r = r;
    q->bytes = gfr_recon_footprint(&q->g, &q->p);
    q->t_admit = s->now();
    q->state = GFR_REQ_ACTIVE;
    q->d2 = q->t2 = q->norm = nullptr;  // inputs are no longer referenced
    q->rec = nullptr;
    s->active.push_back(q);
    s->cv_step.notify_one();
  }
}

// Enqueues passes round-robin over the admitted sessions.
void stepper(gfr_server* s) {
  std::vector<std::shared_ptr<Req>> act;
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    act = s->active;
    lk.unlock();
    bool progressed = false;
    std::vector<std::shared_ptr<Req>> finished;
    for (auto& q : act) {
      // release events of passes the device has finished
      while (!q->evs.empty() && hipEventQuery(q->evs.front()) == hipSuccess) {
        s->ev_pool.push_back(q->evs.front());
        q->evs.pop_front();
      }
      if (q->last) {
        if (q->evs.empty()) finished.push_back(q);
        continue;
      }
      if (q->cancel) {
        // no further pass; the request retires once its queued passes end
        q->error = GFR_ECANCELED;
        q->msg = "request cancelled";
        q->last = true;
        if (q->evs.empty()) finished.push_back(q);
        continue;
      }
      if ((int)q->evs.size() >= s->inflight) continue;
      const int rc = gfr_recon_step(q->r);
      if (rc < 0) {
        q->error = rc;
        q->msg = gfr_last_error(s->ctx);
        q->last = true;
        finished.push_back(q);
        continue;
      }
      q->passes++;
      hipEvent_t e = take_ev(s);
      if (e && hipEventRecord(e, (hipStream_t)gfr_recon_stream(q->r)) == hipSuccess)
        q->evs.push_back(e);
      else if (e)
        s->ev_pool.push_back(e);
      if (rc == 1) q->last = true;
      progressed = true;
    }
    lk.lock();
    for (auto& q : finished) {
      for (size_t i = 0; i < s->active.size(); i++)
        if (s->active[i] == q) {
          s->active.erase(s->active.begin() + i);
          break;
        }
      s->retire.push_back(q);
      s->cv_retire.notify_one();
    }
    if (!progressed && finished.empty())
      s->cv_step.wait_for(lk, std::chrono::microseconds(100));
  }
}

// Reads results back and frees sessions.
void retirer(gfr_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop_retire || !s->retire.empty()) {
    if (s->retire.empty()) {
      s->cv_retire.wait(lk);
      continue;
    }
    std::shared_ptr<Req> q = s->retire.front();
    s->retire.pop_front();
    s->retiring++;
    lk.unlock();
    int rc = q->error;
    std::string msg = q->msg;
    if (rc == GFR_OK) {
      gfr_geometry_info inf{};
      gfr_geometry_check(&q->g, &inf);
      q->xD.resize((size_t)inf.nvox);
      rc = gfr_recon_get(q->r, GFR_OUT_XD, q->xD.data());
      if (rc == GFR_OK && q->p.mode >= 2) {
        q->xT.resize((size_t)inf.nvox);
        rc = gfr_recon_get(q->r, GFR_OUT_XT, q->xT.data());
      }
      if (rc != GFR_OK) msg = gfr_last_error(s->ctx);
    }
    // the result is on the host: the request is complete before its session
    // is freed (so later admissions follow t_done)
    lk.lock();
    s->n_passes += q->passes;
    if (rc != GFR_OK) {
      fail_req(s, q, rc, msg);
    } else {
      q->t_done = s->now();
      q->state = GFR_REQ_DONE;
      s->n_done++;
    }
    s->cv_done.notify_all();
    lk.unlock();
    gfr_recon_destroy(q->r);  // waits for the session's stream
    q->r = nullptr;
    for (hipEvent_t e : q->evs) (void)hipEventDestroy(e);
    q->evs.clear();
    lk.lock();
    s->retiring--;
    s->mem_wait = false;
    s->cv_admit.notify_all();
  }
}

}  // namespace

extern "C" {

int gfr_server_create(gfr_ctx* ctx, int max_active, int inflight, int admitters,
                      gfr_server** out) {
  if (!ctx || !out) return GFR_EINVAL;
  *out = nullptr;
  if (max_active < 1 || max_active > 4096 || inflight < 1 || inflight > 64 || admitters < 1 ||
      admitters > 16)
    return gfr::set_err(ctx, GFR_EINVAL,
                        "max_active in [1, 4096], inflight in [1, 64], admitters in [1, 16]");
  gfr_server* s = new gfr_server();
  s->ctx = ctx;
  s->max_active = max_active;
  s->inflight = inflight;
  s->t0 = std::chrono::steady_clock::now();
  s->th_step = std::thread(stepper, s);
  for (int k = 0; k < admitters; k++) s->th_admit.emplace_back(admitter, s);
  s->th_retire = std::thread(retirer, s);
  *out = s;
  return GFR_OK;
}

int gfr_server_submit(gfr_server* s, const gfr_geometry* g, const gfr_recon_params* p,
                      const float* d2, const float* t2, const float* norm, const int32_t* rec,
                      long long nrec, long long* id) {
  if (!s || !id) return GFR_EINVAL;
  *id = 0;
  if (!g || !p || !d2 || !norm)
    return gfr::set_err(s->ctx, GFR_EINVAL, "gfr_server_submit: geometry, params, d2, norm required");
  if (gfr_geometry_check(g, nullptr) != GFR_OK)
    return gfr::set_err(s->ctx, GFR_EINVAL, "gfr_server_submit: invalid geometry");
  auto q = std::make_shared<Req>();
  q->g = *g;
  q->p = *p;
  q->d2 = d2;
  q->t2 = t2;
  q->norm = norm;
  q->rec = rec;
  q->nrec = nrec;
  std::lock_guard<std::mutex> lk(s->mu);
  q->id = s->next_id++;
  q->t_submit = s->now();
  s->reqs[q->id] = q;
  s->pending.push_back(q);
  s->cv_admit.notify_one();
  *id = q->id;
  return GFR_OK;
}

int gfr_server_info(gfr_server* s, long long id, gfr_request_info* info) {
  if (!s || !info) return GFR_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return gfr::set_err(s->ctx, GFR_EINVAL, "unknown request");
  const Req& q = *it->second;
  info->state = q.state;
  info->error = q.error;
  info->t_submit = q.t_submit;
  info->t_admit = q.t_admit;
  info->t_done = q.t_done;
  info->passes = q.passes;
  info->bytes = q.bytes;
  return GFR_OK;
}

int gfr_server_wait(gfr_server* s, long long id, double timeout_s) {
  if (!s) return GFR_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return gfr::set_err(s->ctx, GFR_EINVAL, "unknown request");
  std::shared_ptr<Req> q = it->second;
  auto done = [&] { return q->state == GFR_REQ_DONE || q->state == GFR_REQ_FAILED; };
  if (timeout_s < 0)
    s->cv_done.wait(lk, done);
  else
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  return q->state;
}

int gfr_server_result(gfr_server* s, long long id, float* xD, float* xT) {
  if (!s) return GFR_EINVAL;
  std::shared_ptr<Req> q;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return gfr::set_err(s->ctx, GFR_EINVAL, "unknown request");
    q = it->second;
  }
  if (q->state == GFR_REQ_FAILED) return gfr::set_err(s->ctx, q->error, q->msg);
  if (q->state != GFR_REQ_DONE) return gfr::set_err(s->ctx, GFR_ESTATE, "request not complete");
  if (xD) memcpy(xD, q->xD.data(), sizeof(float) * q->xD.size());
  if (xT) {
    if (q->xT.empty()) return gfr::set_err(s->ctx, GFR_ESTATE, "no triples image in mode 1");
    memcpy(xT, q->xT.data(), sizeof(float) * q->xT.size());
  }
  return GFR_OK;
}

int gfr_server_release(gfr_server* s, long long id) {
  if (!s) return GFR_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return gfr::set_err(s->ctx, GFR_EINVAL, "unknown request");
  const int st = it->second->state;
  if (st != GFR_REQ_DONE && st != GFR_REQ_FAILED)
    return gfr::set_err(s->ctx, GFR_ESTATE, "request still pending or active");
  s->reqs.erase(it);
  return GFR_OK;
}

int gfr_server_cancel(gfr_server* s, long long id) {
  if (!s) return GFR_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return gfr::set_err(s->ctx, GFR_EINVAL, "unknown request");
  std::shared_ptr<Req> q = it->second;
  for (auto pi = s->pending.begin(); pi != s->pending.end(); ++pi) {
    if (*pi == q) {
      // not admitted yet: it fails now and never takes device memory
      s->pending.erase(pi);
      fail_req(s, q, GFR_ECANCELED, "request cancelled");
      s->cv_done.notify_all();
      return GFR_OK;
    }
  }
  const int st = q->state;
  if (st == GFR_REQ_DONE || st == GFR_REQ_FAILED)
    return gfr::set_err(s->ctx, GFR_ESTATE, "request already complete");
  // being created by an admitter, or admitted: the stepper stops it before
  // its next pass (a request still being created is stopped once admitted)
  q->cancel = true;
  s->cv_step.notify_one();
  return GFR_OK;
}

int gfr_server_stats(gfr_server* s, gfr_server_totals* st) {
  if (!s || !st) return GFR_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  st->pending = (long long)s->pending.size();
  st->active = (long long)s->active.size();
  st->done = s->n_done;
  st->failed = s->n_failed;
  st->passes = s->n_passes;
  st->submitted = s->next_id - 1;
  st->now = s->now();
  return GFR_OK;
}

int gfr_server_destroy(gfr_server* s) {
  if (!s) return GFR_EINVAL;
  {
    // requests never admitted fail now, so that a waiter wakes up
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& q : s->pending) fail_req(s, q, GFR_ESTATE, "server destroyed");
    s->pending.clear();
  }
  s->cv_done.notify_all();
  s->cv_admit.notify_all();
  s->cv_step.notify_all();
  for (std::thread& t : s->th_admit) t.join();
  s->th_step.join();
  {
    // sessions still admitted: retire them (their results are dropped)
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& q : s->active) {
      q->error = GFR_ESTATE;
      q->msg = "server destroyed";
      s->retire.push_back(q);
    }
    s->active.clear();
    s->stop_retire = true;
  }
  s->cv_retire.notify_all();
  s->th_retire.join();
  for (auto& kv : s->reqs)
    for (hipEvent_t e : kv.second->evs) (void)hipEventDestroy(e);
  for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
  delete s;
  return GFR_OK;
}

}  // extern "C"
