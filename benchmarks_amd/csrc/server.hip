// Request server: continuous batching of diffusion solves on one device.
//
//   submit()  -> queue (FIFO)                         [any thread]
//   scheduler thread, each round:
//     admit   : pop queued requests while a slot is free and the request's
//               footprint (two pitched buffers + conversion staging) fits
//               in the server's device pool (first-fit, coalescing free list)
//     upload  : H2D (+ dtype conversion) on the copy stream, marked by an
//               event of the request's own; a request joins the sweeps once
//               its event has completed, so the requests already on the
//               device do not wait for a newcomer's transfer (unless more
//               cells are uploading than are on the device: then the sweep
//               waits for the uploads)
//     sweep   : every active request advances min(remaining, sweep_steps)
//               steps; one launch per kernel class covers all of them (the
//               batched sweep planner), so requests that arrived while others
//               were running share the next launches
//     finish  : requests that reach their step count are copied back on the
//               copy stream after their last sweep; a host callback stamps
//               the completion time and wakes waiters; the scheduler then
//               returns their memory to the pool
//   Two sweeps may be in flight (double-buffered launch tables), so host-side
//   planning of sweep k+1 overlaps the device running sweep k.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "plan.hpp"

using namespace d3h;

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) with coalescing (scheduler thread only).
class Pool {
 public:
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
  }
  bool allocate(size_t bytes, size_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t roff = it->first + bytes;
        free_.erase(it);
        if (rest) free_[roff] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<size_t, size_t> free_;
};

enum State { kQueued, kActive, kDraining, kDone, kFailed };

struct Request {
  int64_t id = 0;
  d3h_problem p{};
  const void *in = nullptr;
  int in_dt = 0;
  void *out = nullptr;
  int out_dt = 0;
  double t_submit = 0, t_done = 0;
  std::atomic<int> state{kQueued};
  int status = D3H_OK;
  size_t off = 0, footprint = 0, buf_bytes = 0, stage_off = 0;
  int pitch = 0;
  int cur = 0;
  int remaining = 0;
  hipEvent_t up_ev = nullptr;  // marks the end of the upload (scheduler)
  bool uploaded = false;       // up_ev has completed: the request can sweep
  bool freed = false;     // pool memory returned (scheduler)
  bool released = false;  // record released by d3h_server_wait before `freed`
};

}  // namespace

struct d3h_server {
  d3h_context *ctx = nullptr;
  d3h_server_config cfg{};
  void *pool = nullptr;
  size_t pool_bytes = 0;
  Pool alloc;
  hipStream_t compute = nullptr, copy = nullptr;
  std::vector<hipStream_t> extra;
  std::vector<hipEvent_t> fork_events;
  std::vector<hipEvent_t> ev_free;  // upload events not in use (scheduler)
  hipEvent_t sweep_ev[2] = {nullptr, nullptr};
  bool sweep_pending[2] = {false, false};
  void *dev_tables[2] = {nullptr, nullptr};
  char *host_tables[2] = {nullptr, nullptr};
  size_t table_cap = 0;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int fatal = D3H_OK;
  std::deque<Request *> queue;
  std::vector<Request *> active, draining;
  std::unordered_map<int64_t, Request *> reqs;
  int64_t next_id = 1;
  int64_t completed = 0, rejected = 0, sweeps = 0;
  double busy_ms = 0;
};

namespace {

void on_request_done(void *arg) {
  auto *pair = static_cast<std::pair<d3h_server *, Request *> *>(arg);
  d3h_server *s = pair->first;
  Request *r = pair->second;
  delete pair;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->t_done = now_ms();
    r->state = kDone;
    ++s->completed;
  }
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

int ensure_tables(d3h_server *s, size_t bytes) {
  if (bytes <= s->table_cap) return D3H_OK;
  // no sweep may still read the old tables
  D3H_HIP(hipStreamSynchronize(s->compute));
  for (int i = 0; i < 2; ++i) {
    if (s->dev_tables[i]) (void)hipFree(s->dev_tables[i]);
    if (s->host_tables[i]) (void)hipHostFree(s->host_tables[i]);
    s->dev_tables[i] = nullptr;
    s->host_tables[i] = nullptr;
    s->sweep_pending[i] = false;
  }
  const size_t cap = std::max(bytes * 2, (size_t)1 << 20);
  for (int i = 0; i < 2; ++i) {
    if (hipMalloc(&s->dev_tables[i], cap) != hipSuccess) return fail(D3H_ERR_NOMEM, "server tables");
    if (hipHostMalloc((void **)&s->host_tables[i], cap, hipHostMallocDefault) != hipSuccess)
      return fail(D3H_ERR_NOMEM, "server pinned tables");
  }
  s->table_cap = cap;
  return D3H_OK;
}

char *field_ptr(d3h_server *s, Request *r, int which) {
  return static_cast<char *>(s->pool) + r->off + which * r->buf_bytes;
}

int upload(d3h_server *s, Request *r) {
  const d3h_problem &p = r->p;
  char *stage = static_cast<char *>(s->pool) + r->stage_off;
  D3H_TRY(copy_in(r->in, r->in_dt, field_ptr(s, r, 0), p.dtype, r->pitch, p.nx, p.ny, p.nz, stage, s->copy));
  r->cur = 0;
  if (s->ev_free.empty()) {
    hipEvent_t ev = nullptr;
    D3H_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    s->ev_free.push_back(ev);
  }
  r->up_ev = s->ev_free.back();
  s->ev_free.pop_back();
  r->uploaded = false;
  D3H_HIP(hipEventRecord(r->up_ev, s->copy));
  return D3H_OK;
}

// Has r's upload completed?  Its event then goes back to the free list, and
// the compute stream is ordered after it (already satisfied: no wait).
int poll_upload(d3h_server *s, Request *r) {
  if (r->uploaded) return D3H_OK;
  const hipError_t e = hipEventQuery(r->up_ev);
  if (e == hipErrorNotReady) return D3H_OK;
  if (e != hipSuccess) return fail(D3H_ERR_HIP, "upload of request %lld: %s", (long long)r->id,
                                   hipGetErrorString(e));
  D3H_HIP(hipStreamWaitEvent(s->compute, r->up_ev, 0));
  s->ev_free.push_back(r->up_ev);
  r->up_ev = nullptr;
  r->uploaded = true;
  return D3H_OK;
}

int download(d3h_server *s, Request *r) {
  const d3h_problem &p = r->p;
  char *stage = static_cast<char *>(s->pool) + r->stage_off;
  D3H_TRY(copy_out(field_ptr(s, r, r->cur), p.dtype, r->pitch, r->out, r->out_dt, p.nx, p.ny, p.nz, stage,
                   s->copy));
  auto *arg = new std::pair<d3h_server *, Request *>(s, r);
  hipError_t e = hipLaunchHostFunc(s->copy, on_request_done, arg);
  if (e != hipSuccess) {
    delete arg;
    return fail(D3H_ERR_HIP, "hipLaunchHostFunc: %s", hipGetErrorString(e));
  }
  return D3H_OK;
}

// One scheduling round; returns true when it issued device work.
int round_once(d3h_server *s, bool *did_work) {
  *did_work = false;
  std::vector<Request *> admitted;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    // return memory of finished requests
    for (auto it = s->draining.begin(); it != s->draining.end();) {
      Request *r = *it;
      if (r->state.load() == kDone) {
        s->alloc.release(r->off, r->footprint);
        r->freed = true;
        it = s->draining.erase(it);
        if (r->released) delete r;  // its waiter has already returned
      } else {
        ++it;
      }
    }
    while (!s->queue.empty() && (int)(s->active.size() + admitted.size()) < s->cfg.max_active) {
      Request *r = s->queue.front();
      size_t off = 0;
      if (!s->alloc.allocate(r->footprint, &off)) break;  // FIFO: wait for memory
      s->queue.pop_front();
      r->off = off;
      r->stage_off = off + 2 * r->buf_bytes;
      r->state = kActive;
      admitted.push_back(r);
    }
  }
  if (!admitted.empty()) {
    for (Request *r : admitted) {
      D3H_TRY(upload(s, r));
      r->remaining = r->p.steps;
    }
    s->active.insert(s->active.end(), admitted.begin(), admitted.end());
    *did_work = true;
  }
  std::vector<Request *> members, pending, finishing;
  double ready_cells = 0, pending_cells = 0;
  for (Request *r : s->active) {
    if (r->remaining == 0) {
      finishing.push_back(r);  // zero-step requests (copied back after their upload)
      continue;
    }
    D3H_TRY(poll_upload(s, r));
    const double cells = (double)r->p.nx * r->p.ny * r->p.nz;
    if (r->uploaded) {
      members.push_back(r);
      ready_cells += cells;
    } else {
      pending.push_back(r);
      pending_cells += cells;
    }
  }
  // A burst of arrivals (more cells still uploading than on the device): the
  // sweep waits for those uploads instead of running the few ready requests
  // alone, whose small sweeps would leave the device mostly idle and finish
  // the burst later.  A single newcomer among running requests joins once
  // its upload has completed, without holding them up.
  if (!pending.empty() && ready_cells < pending_cells) {
    for (Request *r : pending) {
      D3H_HIP(hipStreamWaitEvent(s->compute, r->up_ev, 0));
      s->ev_free.push_back(r->up_ev);  // a later record does not affect this wait
      r->up_ev = nullptr;
      r->uploaded = true;
      members.push_back(r);
    }
  }
  if (!members.empty()) {
    const int slot = (int)(s->sweeps & 1);
    if (s->sweep_pending[slot]) D3H_HIP(hipEventSynchronize(s->sweep_ev[slot]));
    std::vector<PlanMember> ms;
    for (Request *r : members) {
      PlanMember m{};
      m.nx = r->p.nx; m.ny = r->p.ny; m.nz = r->p.nz; m.dtype = r->p.dtype; m.pitch = r->pitch;
      m.buf[0] = field_ptr(s, r, 0);
      m.buf[1] = field_ptr(s, r, 1);
      m.parity0 = r->cur;
      m.steps = std::min(r->remaining, s->cfg.sweep_steps);
      for (int k = 0; k < 7; ++k) m.coef[k] = r->p.coef[k];
      m.sr = r->p.rounding == D3H_ROUND_STOCHASTIC;
      m.seed = r->p.seed;
      m.step_base = r->p.steps - r->remaining;  // the request's own step numbers
      ms.push_back(m);
    }
    SweepPlan plan;
    D3H_TRY(plan_sweep(s->ctx, ms, s->cfg.temporal_depth, s->cfg.arith == D3H_ARITH_FMA, &plan));
    D3H_TRY(ensure_tables(s, plan.table_bytes));
    plan_fill(plan, s->host_tables[slot]);
    D3H_HIP(hipMemcpyAsync(s->dev_tables[slot], s->host_tables[slot], plan.table_bytes,
                           hipMemcpyHostToDevice, s->compute));
    int launches = 0;
    D3H_TRY(plan_enqueue(plan.A, s->dev_tables[slot], s->compute, s->extra, s->fork_events, &launches));
    D3H_TRY(plan_enqueue(plan.B, s->dev_tables[slot], s->compute, s->extra, s->fork_events, &launches));
    D3H_HIP(hipEventRecord(s->sweep_ev[slot], s->compute));
    s->sweep_pending[slot] = true;
    ++s->sweeps;
    for (size_t i = 0; i < members.size(); ++i) {
      Request *r = members[i];
      r->cur = plan.parity_after[i];
      r->remaining -= ms[i].steps;
      if (r->remaining == 0) finishing.push_back(r);
    }
    // results are copied after this sweep
    if (!finishing.empty()) D3H_HIP(hipStreamWaitEvent(s->copy, s->sweep_ev[slot], 0));
    *did_work = true;
  }
  if (!finishing.empty()) {
    for (Request *r : finishing) {
      if (r->up_ev) {  // a zero-step request: its download follows the upload on the copy stream
        s->ev_free.push_back(r->up_ev);
        r->up_ev = nullptr;
      }
      D3H_TRY(download(s, r));
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request *r : finishing) {
      r->state = r->state.load() == kDone ? (int)kDone : (int)kDraining;
      s->draining.push_back(r);
      s->active.erase(std::find(s->active.begin(), s->active.end(), r));
    }
  }
  return D3H_OK;
}

void scheduler(d3h_server *s) {
  (void)hipSetDevice(s->ctx->device);
  double busy_start = -1;
  while (true) {
    bool work = false;
    int rc = round_once(s, &work);
    std::unique_lock<std::mutex> lk(s->mu);
    if (rc != D3H_OK) {
      s->fatal = rc;
      // fail everything that has not completed
      for (auto &kv : s->reqs) {
        Request *r = kv.second;
        const int st = r->state.load();
        if (st != kDone && st != kFailed) {
          r->state = kFailed;
          r->status = rc;
        }
      }
      s->queue.clear();
      s->cv_done.notify_all();
      return;
    }
    const bool idle = s->queue.empty() && s->active.empty();
    if (!idle && busy_start < 0) busy_start = now_ms();
    if (idle && s->draining.empty() && busy_start >= 0) {
      s->busy_ms += now_ms() - busy_start;
      busy_start = -1;
    }
    if (s->stop && idle && s->draining.empty()) return;
    if (!work) {
      if (!s->active.empty())  // uploads in flight
        s->cv_work.wait_for(lk, std::chrono::microseconds(20));
      else if (!s->draining.empty() || !s->queue.empty())
        s->cv_work.wait_for(lk, std::chrono::microseconds(100));
      else
        s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
    }
  }
}

}  // namespace

extern "C" {

int d3h_server_create(d3h_context *ctx, const d3h_server_config *cfg_in, d3h_server **out) {
  if (!ctx || !out) return fail(D3H_ERR_INVALID, "ctx/out is NULL");
  *out = nullptr;
  d3h_server_config cfg{};
  if (cfg_in) cfg = *cfg_in;
  if (cfg.arith != D3H_ARITH_EXACT && cfg.arith != D3H_ARITH_FMA) return fail(D3H_ERR_INVALID, "arith");
  if (cfg.max_active <= 0) cfg.max_active = 4096;
  if (cfg.temporal_depth < 0 || cfg.temporal_depth == 3 || cfg.temporal_depth > 4)
    return fail(D3H_ERR_INVALID, "server depth %d (0 = tuned per size class, 1, 2 or 4)",
                cfg.temporal_depth);
  if (cfg.sweep_steps <= 0) cfg.sweep_steps = 8;
  D3H_HIP(hipSetDevice(ctx->device));
  size_t fr = 0, tot = 0;
  D3H_HIP(hipMemGetInfo(&fr, &tot));
  if (cfg.hbm_budget_bytes == 0) cfg.hbm_budget_bytes = fr / 10 * 9;
  if (cfg.hbm_budget_bytes > fr) return fail(D3H_ERR_NOMEM, "budget %zu > free %zu", cfg.hbm_budget_bytes, fr);
  auto *s = new d3h_server();
  s->ctx = ctx;
  s->cfg = cfg;
  s->pool_bytes = align_up(cfg.hbm_budget_bytes);
  bool ok = hipMalloc(&s->pool, s->pool_bytes) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) == hipSuccess;
  for (int i = 0; ok && i < 2; ++i) {
    hipStream_t st;
    ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    if (ok) s->extra.push_back(st);
  }
  for (int i = 0; ok && i < 3; ++i) {
    hipEvent_t ev;
    ok = hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    if (ok) s->fork_events.push_back(ev);
  }
  ok = ok && hipEventCreateWithFlags(&s->sweep_ev[0], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&s->sweep_ev[1], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    const size_t pool_bytes = s->pool_bytes;
    d3h_server_destroy(s);
    return fail(D3H_ERR_NOMEM, "server resources (pool of %zu bytes)", pool_bytes);
  }
  s->alloc.reset(s->pool_bytes);
  s->th = std::thread(scheduler, s);
  *out = s;
  return D3H_OK;
}

int d3h_server_destroy(d3h_server *s) {
  if (!s) return D3H_OK;
  if (s->th.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    s->th.join();
  }
  (void)hipSetDevice(s->ctx->device);
  if (s->compute) (void)hipStreamSynchronize(s->compute);
  if (s->copy) (void)hipStreamSynchronize(s->copy);
  for (auto st : s->extra) (void)hipStreamDestroy(st);
  for (auto ev : s->fork_events) (void)hipEventDestroy(ev);
  for (auto ev : s->ev_free) (void)hipEventDestroy(ev);
  for (Request *r : s->active)
    if (r->up_ev) (void)hipEventDestroy(r->up_ev);
  for (int i = 0; i < 2; ++i) {
    if (s->sweep_ev[i]) (void)hipEventDestroy(s->sweep_ev[i]);
    if (s->dev_tables[i]) (void)hipFree(s->dev_tables[i]);
    if (s->host_tables[i]) (void)hipHostFree(s->host_tables[i]);
  }
  if (s->compute) (void)hipStreamDestroy(s->compute);
  if (s->copy) (void)hipStreamDestroy(s->copy);
  if (s->pool) (void)hipFree(s->pool);
  // released records still draining are no longer in `reqs`
  for (Request *r : s->draining)
    if (r->released) delete r;
  for (auto &kv : s->reqs) delete kv.second;
  delete s;
  return D3H_OK;
}

int d3h_problem_check(const d3h_problem *p) {
  if (!p) return fail(D3H_ERR_INVALID, "problem is NULL");
  return check_problem(*p);
}

int d3h_problem_footprint(const d3h_problem *p, int host_dtype, int out_dtype, size_t *bytes) {
  if (!p || !bytes) return fail(D3H_ERR_INVALID, "NULL argument");
  if (!dtype_valid(host_dtype) || !dtype_valid(out_dtype)) return fail(D3H_ERR_INVALID, "dtype");
  D3H_TRY(check_problem(*p));
  const int pitch = pitch_elems(p->nx, p->dtype);
  const size_t buf = align_up((size_t)pitch * p->ny * p->nz * dtype_bytes(p->dtype));
  const size_t n = (size_t)p->nx * p->ny * p->nz;
  size_t stage = 0;
  if (needs_staging(host_dtype, p->dtype, pitch, p->nx)) stage = std::max(stage, n * dtype_bytes(host_dtype));
  if (needs_staging(out_dtype, p->dtype, pitch, p->nx)) stage = std::max(stage, n * dtype_bytes(out_dtype));
  *bytes = 2 * buf + align_up(stage);
  return D3H_OK;
}

int d3h_server_submit(d3h_server *s, const d3h_problem *p, const void *host_in, int host_dtype,
                      void *host_out, int out_dtype, int64_t *id) {
  if (!s || !p || !host_in || !host_out || !id) return fail(D3H_ERR_INVALID, "NULL argument");
  size_t footprint = 0;
  D3H_TRY(d3h_problem_footprint(p, host_dtype, out_dtype, &footprint));
  auto *r = new Request();
  r->p = *p;
  r->in = host_in;
  r->in_dt = host_dtype;
  r->out = host_out;
  r->out_dt = out_dtype;
  r->pitch = pitch_elems(p->nx, p->dtype);
  r->buf_bytes = align_up((size_t)r->pitch * p->ny * p->nz * dtype_bytes(p->dtype));
  r->footprint = footprint;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fatal != D3H_OK) {
    delete r;
    return fail(s->fatal, "server stopped after an error");
  }
  if (r->footprint > s->pool_bytes) {
    ++s->rejected;
    delete r;
    return fail(D3H_ERR_NOMEM, "request needs %zu bytes, server pool is %zu", r->footprint, s->pool_bytes);
  }
  r->id = s->next_id++;
  r->t_submit = now_ms();
  s->reqs[r->id] = r;
  s->queue.push_back(r);
  *id = r->id;
  s->cv_work.notify_all();
  return D3H_OK;
}

int d3h_server_poll(d3h_server *s, int64_t id, int *done, double *latency_ms) {
  if (!s || !done) return fail(D3H_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(D3H_ERR_NOT_FOUND, "request %lld", (long long)id);
  Request *r = it->second;
  const int st = r->state.load();
  *done = (st == kDone || st == kFailed) ? 1 : 0;
  if (latency_ms) *latency_ms = st == kDone ? r->t_done - r->t_submit : 0.0;
  if (st == kFailed) return fail(r->status, "request %lld failed", (long long)id);
  return D3H_OK;
}

int d3h_server_wait(d3h_server *s, int64_t id, double *latency_ms) {
  if (!s) return fail(D3H_ERR_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(D3H_ERR_NOT_FOUND, "request %lld", (long long)id);
  Request *r = it->second;
  s->cv_done.wait(lk, [&] {
    const int st = r->state.load();
    return st == kDone || st == kFailed;
  });
  if (r->state.load() == kFailed) {
    if (latency_ms) *latency_ms = 0.0;
    return fail(r->status, "request %lld failed", (long long)id);
  }
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  // release the record: at once when its memory is back in the pool, else
  // when the scheduler returns it (the record must outlive its draining)
  s->reqs.erase(it);
  if (r->freed) delete r;
  else r->released = true;
  return D3H_OK;
}

int d3h_server_stats(d3h_server *s, int64_t *completed, int64_t *rejected, int64_t *sweeps,
                     double *busy_ms) {
  if (!s) return fail(D3H_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (completed) *completed = s->completed;
  if (rejected) *rejected = s->rejected;
  if (sweeps) *sweeps = s->sweeps;
  if (busy_ms) *busy_ms = s->busy_ms;
  return D3H_OK;
}

}  // extern "C"
