// Serving scheduler: continuous batching of many independent simulation runs
// on one device.  The reference runs one solver per process; the steps each
// request takes are those of atm_lsrk_run / atm_ark23_run (src/odesolvers.jl:28-67,
// 231-388), so a served result is bitwise the result of the run alone.
//
// Threads: callers submit, poll and wait under one mutex; a worker thread owns
// the device side (admission, ticks, retirement) and launches only on the
// server's own streams, through the per-thread stream override, so the
// context's stream stays with the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "handles.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  int64_t id = 0;
  atm_request r{};
  std::vector<double> q0;  // host copy of the initial state until the upload has landed
  int64_t n = 0;           // doubles in the state
  uint64_t state_bytes = 0;  // the request's own device buffers while resident
  // the operators it runs (dg, lin, band; null where unused) and their device bytes:
  // counted against the budget once, while any resident request runs them
  const void* ops[3] = {nullptr, nullptr, nullptr};
  uint64_t op_bytes[3] = {0, 0, 0};
  int64_t left = 0;        // steps still to take
  int state = ATM_REQ_QUEUED;  // published under the server's mutex
  std::string error;           // published with the state
  int frc = 0;                 // worker: failure status of this request
  std::string fmsg;            // worker: failure message
  hipEvent_t ready = nullptr;  // recorded on the submitting thread's stream
  double t_submit = -1, t_admit = -1, t_done = -1;
  double* q = nullptr;
  double* dq = nullptr;
  double* scratch = nullptr;
  double* work = nullptr;
};

int stages_of(const atm_request& r) {
  if (r.kind == ATM_REQ_ARK23) return 3;
  return r.scheme == ATM_LSRK144 ? 14 : 5;
}

constexpr int MAX_SIDE_STREAMS = 8;

}  // namespace

struct atm_server {
  atm_ctx* ctx = nullptr;
  atm_server_desc desc{};
  Clock::time_point t0;
  hipStream_t main = nullptr;
  std::vector<hipStream_t> side;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr, ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  std::thread worker;
  // shared state (under mu)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int64_t next_id = 1;
  std::deque<Req*> pending;
  std::deque<int64_t> finished;  // not yet handed out by atm_server_next_done
  std::unordered_map<int64_t, Req*> reqs;
  uint64_t state_in_use = 0;  // device bytes of resident requests' buffers
  uint64_t op_in_use = 0;     // device bytes of the operators resident requests run
  std::unordered_map<const void*, int64_t> op_refs;  // resident requests per operator
  uint64_t pool_idle = 0;     // published copy of pool_bytes
  int64_t unfinished = 0;
  atm_server_stats stats{};
  // worker-owned
  std::vector<Req*> active;
  atm_batch* batch[2] = {nullptr, nullptr};  // one per LSRK scheme, regrouped as members change
  std::vector<int64_t> members[2];
  // idle device buffers of retired requests, by rounded size, reused by later
  // admissions instead of hipMalloc/hipFree per request
  std::multimap<size_t, void*> pool;
  uint64_t pool_bytes = 0;
  int64_t pool_hits = 0, pool_misses = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

constexpr size_t POOL_GRAIN = (size_t)1 << 16;
size_t pool_round(size_t bytes) { return (bytes + POOL_GRAIN - 1) / POOL_GRAIN * POOL_GRAIN; }

void pool_trim(atm_server* s, uint64_t limit) {
  while (s->pool_bytes > limit && !s->pool.empty()) {
    auto it = std::prev(s->pool.end());  // largest first
    hipFree(it->second);
    s->pool_bytes -= it->first;
    s->pool.erase(it);
  }
}

// A device buffer of at least `bytes` from the pool (same rounded size) or hipMalloc;
// on an allocation failure the idle pool is released and the allocation retried.
hipError_t pool_get(atm_server* s, size_t bytes, double** out) {
  const size_t rb = pool_round(bytes);
  auto it = s->pool.find(rb);
  if (it != s->pool.end()) {
    *out = static_cast<double*>(it->second);
    s->pool.erase(it);
    s->pool_bytes -= rb;
    s->pool_hits += 1;
    return hipSuccess;
  }
  s->pool_misses += 1;
  hipError_t e = hipMalloc((void**)out, rb);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pool_trim(s, 0);
    e = hipMalloc((void**)out, rb);
  }
  return e;
}

void pool_put(atm_server* s, double* p, size_t bytes) {
  if (!p) return;
  const size_t rb = pool_round(bytes);
  s->pool.emplace(rb, p);
  s->pool_bytes += rb;
}

// Buffers back to the pool, once the work on them has finished (after the tick's
// synchronisation).
void release_buffers(atm_server* s, Req* r) {
  const size_t bytes = (size_t)r->n * sizeof(double);
  pool_put(s, r->q, bytes);
  pool_put(s, r->dq, bytes);
  pool_put(s, r->scratch, bytes);
  pool_put(s, r->work, 9 * bytes);
  r->q = r->dq = r->scratch = r->work = nullptr;
}

void free_buffers(Req* r) {
  for (double** p : {&r->q, &r->dq, &r->scratch, &r->work}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
}

// Device bytes admitting r adds: its buffers and the operators no resident
// request runs yet (under the server's mutex).
uint64_t admission_bytes(const atm_server* s, const Req* r) {
  uint64_t need = r->state_bytes;
  for (int k = 0; k < 3; ++k) {
    if (!r->ops[k]) continue;
    auto it = s->op_refs.find(r->ops[k]);
    if (it == s->op_refs.end() || it->second == 0) need += r->op_bytes[k];
  }
  return need;
}

void charge(atm_server* s, const Req* r) {
  s->state_in_use += r->state_bytes;
  for (int k = 0; k < 3; ++k) {
    if (!r->ops[k]) continue;
    if (s->op_refs[r->ops[k]]++ == 0) s->op_in_use += r->op_bytes[k];
  }
}

void uncharge(atm_server* s, const Req* r) {
  s->state_in_use -= r->state_bytes;
  for (int k = 0; k < 3; ++k) {
    if (!r->ops[k]) continue;
    auto it = s->op_refs.find(r->ops[k]);
    if (it != s->op_refs.end() && --it->second == 0) {
      s->op_in_use -= r->op_bytes[k];
      s->op_refs.erase(it);
    }
  }
}

void mark_failed(atm_server* s, Req* r, int rc, const char* what) {
  r->frc = rc < 0 ? rc : ATM_ERR_HIP;
  r->fmsg = std::string(what) + ": " + last_error(s->ctx);
}

// Device side of an admitted request, on the server's main stream.
int admit_one(atm_server* s, Req* r) {
  atm_ctx* ctx = s->ctx;
  const size_t bytes = (size_t)r->n * sizeof(double);
  if (r->ready) {
    HIPCHK(ctx, hipStreamWaitEvent(s->main, r->ready, 0));
    hipEventDestroy(r->ready);
    r->ready = nullptr;
  }
  HIPCHK(ctx, pool_get(s, bytes, &r->q));
  if (r->r.kind == ATM_REQ_LSRK) {
    HIPCHK(ctx, pool_get(s, bytes, &r->dq));
    HIPCHK(ctx, pool_get(s, bytes, &r->scratch));
    HIPCHK(ctx, hipMemsetAsync(r->dq, 0, bytes, s->main));
  } else {
    HIPCHK(ctx, pool_get(s, 9 * bytes, &r->work));
  }
  HIPCHK(ctx, hipMemcpyAsync(r->q, r->q0.data(), bytes, hipMemcpyHostToDevice, s->main));
  return ATM_OK;
}

// LSRK requests of one scheme: one batched launch per kernel group and stage.
void run_explicit(atm_server* s, int scheme, const std::vector<Req*>& rs, int64_t step) {
  if (rs.empty()) return;
  std::vector<int64_t> ids;
  std::vector<atm_dg*> dgs;
  std::vector<double*> q, dq, scratch;
  std::vector<double> dt;
  for (Req* r : rs) {
    ids.push_back(r->id);
    dgs.push_back(r->r.dg);
    q.push_back(r->q);
    dq.push_back(r->dq);
    scratch.push_back(r->scratch);
    dt.push_back(r->r.dt);
  }
  int rc = ATM_OK;
  if (!s->batch[scheme]) {
    rc = atm_batch_create(s->ctx, (int32_t)dgs.size(), dgs.data(), &s->batch[scheme]);
    if (rc == ATM_OK) s->members[scheme] = ids;
  } else if (ids != s->members[scheme]) {
    // same batch, new members: its tables keep their memory
    s->members[scheme].clear();
    rc = batch_set_members(s->batch[scheme], (int32_t)dgs.size(), dgs.data());
    if (rc == ATM_OK) s->members[scheme] = ids;
  }
  // eager launches, no graph capture: while a stream captures, HIP refuses the
  // legacy-stream copies that other threads make (operator creation, uploads)
  if (rc == ATM_OK)
    rc = atm_batch_lsrk_run(s->batch[scheme], scheme, q.data(), dq.data(), scratch.data(), dt.data(), step, 0);
  if (rc != ATM_OK) {
    for (Req* r : rs) mark_failed(s, r, rc, "batched LSRK");
    if (s->batch[scheme]) atm_batch_destroy(s->batch[scheme]);
    s->batch[scheme] = nullptr;
    s->members[scheme].clear();
  }
}

// ARK23 requests, round-robin over the side streams forked from the main stream.
void run_implicit(atm_server* s, const std::vector<Req*>& rs, int64_t step) {
  if (rs.empty()) return;
  const int used = std::min<int>((int)s->side.size(), (int)rs.size());
  if (used > 0) {
    if (hipEventRecord(s->ev_fork, s->main) != hipSuccess) {
      for (Req* r : rs) mark_failed(s, r, ATM_ERR_HIP, "fork of the side streams");
      return;
    }
    for (int j = 0; j < used; ++j) hipStreamWaitEvent(s->side[j], s->ev_fork, 0);
  }
  for (size_t i = 0; i < rs.size(); ++i) {
    Req* r = rs[i];
    tl_stream = used > 0 ? s->side[i % used] : s->main;
    const int rc = atm_ark23_run(r->r.dg, r->r.lin, r->r.band, r->q, r->work, r->r.dt, 0.0, step, r->r.split_rhs,
                                 r->r.paperversion);
    if (rc != ATM_OK) mark_failed(s, r, rc, "ARK23");
  }
  tl_stream = s->main;
  for (int j = 0; j < used; ++j) {
    hipEventRecord(s->ev_join[j], s->side[j]);
    hipStreamWaitEvent(s->main, s->ev_join[j], 0);
  }
}

void worker_loop(atm_server* s) {
  hipSetDevice(s->ctx->device);
  tl_stream = s->main;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty() || !s->active.empty(); });
    if (s->pending.empty() && s->active.empty()) break;  // stopping, nothing left
    const double tb = s->now();
    std::vector<Req*> admitted;
    // in submission order while the slots and the budget allow (the first request
    // always fits an empty device: submit rejects requests above the budget)
    while (!s->pending.empty() && (int64_t)(s->active.size() + admitted.size()) < s->desc.max_active &&
           ((double)(s->state_in_use + s->op_in_use + admission_bytes(s, s->pending.front())) <= s->desc.hbm_budget ||
            (s->active.empty() && admitted.empty()))) {
      Req* r = s->pending.front();
      s->pending.pop_front();
      charge(s, r);
      r->state = ATM_REQ_ACTIVE;
      r->t_admit = tb;
      admitted.push_back(r);
    }
    s->stats.peak_bytes = std::max(s->stats.peak_bytes, (double)(s->state_in_use + s->op_in_use));
    s->stats.peak_operator_bytes = std::max(s->stats.peak_operator_bytes, (double)s->op_in_use);
    // idle pooled buffers and resident requests together stay within the budget
    const uint64_t pool_limit = (uint64_t)std::max(0.0, s->desc.hbm_budget - (double)(s->state_in_use + s->op_in_use));
    lk.unlock();
    pool_trim(s, pool_limit);

    atm_server_stats d{};
    std::vector<Req*> out;  // leaving this tick (finished or failed)
    for (Req* r : admitted) {
      const int rc = admit_one(s, r);
      if (rc != ATM_OK) {
        mark_failed(s, r, rc, "admission");
        out.push_back(r);
      } else {
        s->active.push_back(r);
      }
    }
    const double te = s->now();
    d.admit_s = te - tb;

    int64_t step = 0;
    for (Req* r : s->active)
      if (r->left > 0) step = step == 0 ? r->left : std::min(step, r->left);
    step = std::min<int64_t>(step, s->desc.chunk);
    float span_ms = 0.f;
    double tw = te;
    if (step > 0) {
      std::vector<Req*> lsrk[2], ark;
      for (Req* r : s->active) {
        if (r->left == 0) continue;
        if (r->r.kind == ATM_REQ_ARK23)
          ark.push_back(r);
        else
          lsrk[r->r.scheme == ATM_LSRK144 ? 1 : 0].push_back(r);
      }
      hipEventRecord(s->ev_begin, s->main);
      run_explicit(s, ATM_LSRK54, lsrk[0], step);
      run_explicit(s, ATM_LSRK144, lsrk[1], step);
      run_implicit(s, ark, step);
      hipEventRecord(s->ev_end, s->main);
      tw = s->now();
      d.enqueue_s = tw - te;
      hipEventSynchronize(s->ev_end);
      hipEventElapsedTime(&span_ms, s->ev_begin, s->ev_end);
      d.gpu_span_s = span_ms / 1e3;
      d.ticks = 1;
    }
    const hipError_t sticky = hipStreamSynchronize(s->main);
    const double tr = s->now();
    d.wait_s = tr - tw;

    std::vector<Req*> keep;
    for (Req* r : s->active) {
      if (r->frc < 0) {
        out.push_back(r);
        continue;
      }
      if (sticky != hipSuccess) {
        r->frc = ATM_ERR_HIP;
        r->fmsg = std::string("device: ") + hipGetErrorString(sticky);
        out.push_back(r);
        continue;
      }
      if (r->left > 0) {
        r->left -= step;
        d.dof_updates += (double)r->n * (double)step * stages_of(r->r);
      }
      if (!r->q0.empty()) std::vector<double>().swap(r->q0);  // the upload has landed
      if (r->left == 0) {
        out.push_back(r);
      } else {
        keep.push_back(r);
      }
    }
    s->active.swap(keep);
    for (Req* r : out) {
      if (r->frc == 0 && r->r.result) {
        const hipError_t e = hipMemcpyAsync(r->r.result, r->q, (size_t)r->n * sizeof(double), hipMemcpyDeviceToHost,
                                            s->main);
        if (e != hipSuccess) {
          r->frc = ATM_ERR_HIP;
          r->fmsg = std::string("result download: ") + hipGetErrorString(e);
        }
      }
    }
    if (!out.empty()) {
      const hipError_t e = hipStreamSynchronize(s->main);
      if (e != hipSuccess)
        for (Req* r : out)
          if (r->frc == 0) {
            r->frc = ATM_ERR_HIP;
            r->fmsg = std::string("result download: ") + hipGetErrorString(e);
          }
    }
    for (Req* r : out) {
      release_buffers(s, r);
      std::vector<double>().swap(r->q0);
      if (r->ready) {
        hipEventDestroy(r->ready);
        r->ready = nullptr;
      }
    }
    const double tf = s->now();
    d.retire_s = tf - tr;

    lk.lock();
    for (Req* r : out) {
      if (r->frc == 0) {
        r->state = ATM_REQ_DONE;
        s->stats.completed += 1;
      } else {
        r->state = r->frc;
        r->error = r->fmsg;
        s->stats.failed += 1;
      }
      r->t_done = tf;
      uncharge(s, r);
      s->unfinished -= 1;
      s->finished.push_back(r->id);
    }
    s->stats.ticks += d.ticks;
    s->stats.max_active = std::max<int64_t>(s->stats.max_active, (int64_t)(s->active.size() + out.size()));
    s->stats.dof_updates += d.dof_updates;
    s->stats.busy_s += tf - tb;
    s->stats.gpu_span_s += d.gpu_span_s;
    s->stats.admit_s += d.admit_s;
    s->stats.enqueue_s += d.enqueue_s;
    s->stats.wait_s += d.wait_s;
    s->stats.retire_s += d.retire_s;
    s->stats.pool_hits = s->pool_hits;
    s->stats.pool_misses = s->pool_misses;
    s->stats.pool_bytes = (double)s->pool_bytes;
    if (!out.empty()) s->cv_done.notify_all();
  }
  tl_stream = nullptr;
}

Req* find(atm_server* s, int64_t id) {
  auto it = s->reqs.find(id);
  return it == s->reqs.end() ? nullptr : it->second;
}

bool finished_state(int st) { return st == ATM_REQ_DONE || st < 0; }

}  // namespace

extern "C" {

int atm_server_create(atm_ctx* ctx, const atm_server_desc* d, atm_server** out) {
  if (!out) return ATM_ERR_ARG;
  *out = nullptr;
  if (!ctx || !d) return fail(ctx, ATM_ERR_ARG, "atm_server_create: null argument");
  if (d->max_active < 1 || d->chunk < 1 || d->imex_streams < 0 || d->imex_streams > MAX_SIDE_STREAMS ||
      !(d->hbm_budget > 0) || d->reserved != 0)
    return fail(ctx, ATM_ERR_ARG, "atm_server_create: max_active, chunk >= 1, imex_streams in [0, 8], hbm_budget > 0");
  hipSetDevice(ctx->device);
  atm_server* s = new atm_server();
  s->ctx = ctx;
  s->desc = *d;
  s->t0 = Clock::now();
  auto cleanup = [s]() {
    for (hipStream_t st : s->side) hipStreamDestroy(st);
    for (hipEvent_t e : s->ev_join) hipEventDestroy(e);
    if (s->main) hipStreamDestroy(s->main);
    if (s->ev_begin) hipEventDestroy(s->ev_begin);
    if (s->ev_end) hipEventDestroy(s->ev_end);
    if (s->ev_fork) hipEventDestroy(s->ev_fork);
    delete s;
  };
  hipError_t e = hipStreamCreateWithFlags(&s->main, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&s->ev_begin);
  if (e == hipSuccess) e = hipEventCreate(&s->ev_end);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming);
  for (int j = 0; j < d->imex_streams && e == hipSuccess; ++j) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) s->side.push_back(st);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) s->ev_join.push_back(ev);
  }
  if (e != hipSuccess) {
    cleanup();
    return hipfail(ctx, e, "atm_server_create");
  }
  try {
    s->worker = std::thread(worker_loop, s);
  } catch (...) {
    cleanup();
    return fail(ctx, ATM_ERR_NOMEM, "atm_server_create: cannot start the worker thread");
  }
  *out = s;
  return ATM_OK;
}

int atm_server_destroy(atm_server* s) {
  if (!s) return ATM_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  hipSetDevice(s->ctx->device);
  for (int k = 0; k < 2; ++k)
    if (s->batch[k]) atm_batch_destroy(s->batch[k]);
  hipStreamSynchronize(s->main);
  pool_trim(s, 0);
  for (auto& kv : s->reqs) {
    Req* r = kv.second;
    free_buffers(r);
    if (r->ready) hipEventDestroy(r->ready);
    delete r;
  }
  for (hipStream_t st : s->side) {
    hipStreamSynchronize(st);
    hipStreamDestroy(st);
  }
  for (hipEvent_t e : s->ev_join) hipEventDestroy(e);
  hipStreamSynchronize(s->main);
  hipStreamDestroy(s->main);
  hipEventDestroy(s->ev_begin);
  hipEventDestroy(s->ev_end);
  hipEventDestroy(s->ev_fork);
  delete s;
  return ATM_OK;
}

int atm_server_submit(atm_server* s, const atm_request* req, int64_t* id) {
  if (!s || !req || !id) return s ? fail(s->ctx, ATM_ERR_ARG, "atm_server_submit: null argument") : ATM_ERR_ARG;
  atm_ctx* ctx = s->ctx;
  const atm_request& q = *req;
  const atm_dg* shape = nullptr;
  if (q.kind == ATM_REQ_LSRK) {
    if (!q.dg || (q.scheme != ATM_LSRK54 && q.scheme != ATM_LSRK144))
      return fail(ctx, ATM_ERR_ARG, "atm_server_submit: an LSRK request needs dg and scheme LSRK54 or LSRK144");
    if (q.dg->ctx != ctx) return fail(ctx, ATM_ERR_ARG, "atm_server_submit: operator of another context");
    shape = q.dg;
  } else if (q.kind == ATM_REQ_ARK23) {
    if (!q.lin || !q.band) return fail(ctx, ATM_ERR_ARG, "atm_server_submit: an ARK23 request needs lin and band");
    if (banded_lin(q.band) != q.lin) return fail(ctx, ATM_ERR_ARG, "atm_server_submit: band was not built from lin");
    if (q.lin->ctx != ctx || (q.dg && q.dg->ctx != ctx))
      return fail(ctx, ATM_ERR_ARG, "atm_server_submit: operator of another context");
    if (q.dg && (q.dg->Ntot != q.lin->Ntot || q.dg->Ns != q.lin->Ns))
      return fail(ctx, ATM_ERR_ARG, "atm_server_submit: operators differ in shape");
    shape = q.lin;
  } else {
    return fail(ctx, ATM_ERR_ARG, "atm_server_submit: unknown request kind");
  }
  if (!q.q0 || q.steps < 0 || !std::isfinite(q.dt))
    return fail(ctx, ATM_ERR_ARG, "atm_server_submit: q0, finite dt and steps >= 0 required");
  Req* r = new Req();
  r->r = q;
  r->n = (int64_t)shape->Ns * shape->Ntot;
  const size_t sb = (size_t)r->n * sizeof(double);
  r->state_bytes = q.kind == ATM_REQ_LSRK ? 3 * pool_round(sb) : pool_round(sb) + pool_round(9 * sb);
  uint64_t total = r->state_bytes;
  if (q.kind == ATM_REQ_LSRK) {
    r->ops[0] = q.dg;
    r->op_bytes[0] = q.dg->dev_bytes;
  } else {
    r->ops[0] = q.dg;
    r->op_bytes[0] = q.dg ? q.dg->dev_bytes : 0;
    r->ops[1] = q.lin;
    r->op_bytes[1] = q.lin->dev_bytes;
    r->ops[2] = q.band;
    r->op_bytes[2] = (uint64_t)atm_banded_device_bytes(q.band);
  }
  for (int k = 0; k < 3; ++k) total += r->op_bytes[k];
  if ((double)total > s->desc.hbm_budget) {
    delete r;
    return fail(ctx, ATM_ERR_ARG,
                "atm_server_submit: the request (its buffers and operators) needs more device memory than the budget");
  }
  r->q0.assign(q.q0, q.q0 + r->n);
  r->left = q.steps;
  hipSetDevice(ctx->device);
  hipError_t e = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(r->ready, stream_of(ctx));
  if (e != hipSuccess) {
    if (r->ready) hipEventDestroy(r->ready);
    delete r;
    return hipfail(ctx, e, "atm_server_submit");
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->t_submit = s->now();
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    s->unfinished += 1;
    s->stats.submitted += 1;
    *id = r->id;
  }
  s->cv_work.notify_one();
  return ATM_OK;
}

int atm_server_poll(atm_server* s, int64_t id) {
  if (!s) return ATM_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  Req* r = find(s, id);
  return r ? r->state : fail(s->ctx, ATM_ERR_ARG, "atm_server_poll: unknown request");
}

int atm_server_wait(atm_server* s, int64_t id) {
  if (!s) return ATM_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (!find(s, id)) return fail(s->ctx, ATM_ERR_ARG, "atm_server_wait: unknown request");
  // look the request up again at every wake-up: another thread may release it
  // as soon as it has finished
  s->cv_done.wait(lk, [s, id] {
    const Req* x = find(s, id);
    return !x || finished_state(x->state);
  });
  const Req* r = find(s, id);
  return r ? r->state : ATM_REQ_DONE;
}

int atm_server_next_done(atm_server* s, double timeout_s, int64_t* id) {
  if (!s || !id) return ATM_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [s] { return !s->finished.empty(); };
  if (timeout_s < 0)
    s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return 0;
  *id = s->finished.front();
  s->finished.pop_front();
  return 1;
}

int atm_server_drain(atm_server* s) {
  if (!s) return ATM_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [s] { return s->unfinished == 0; });
  return ATM_OK;
}

double atm_server_clock(atm_server* s) { return s ? s->now() : -1.0; }

int atm_server_times(atm_server* s, int64_t id, double* t) {
  if (!s || !t) return ATM_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  Req* r = find(s, id);
  if (!r) return fail(s->ctx, ATM_ERR_ARG, "atm_server_times: unknown request");
  t[0] = r->t_submit;
  t[1] = r->t_admit;
  t[2] = r->t_done;
  return ATM_OK;
}

const char* atm_server_error(atm_server* s, int64_t id) {
  // a copy per calling thread: the request may be released while the caller reads it
  static thread_local std::string msg;
  if (!s) return "null server";
  std::lock_guard<std::mutex> lk(s->mu);
  Req* r = find(s, id);
  msg = r ? r->error : std::string("unknown request");
  return msg.c_str();
}

int atm_server_release(atm_server* s, int64_t id) {
  if (!s) return ATM_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  Req* r = find(s, id);
  if (!r || !finished_state(r->state))
    return fail(s->ctx, ATM_ERR_ARG, "atm_server_release: unknown or unfinished request");
  s->reqs.erase(id);
  s->finished.erase(std::remove(s->finished.begin(), s->finished.end(), id), s->finished.end());
  delete r;
  return ATM_OK;
}

int atm_server_get_stats(atm_server* s, atm_server_stats* out) {
  if (!s || !out) return ATM_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return ATM_OK;
}

}  // extern "C"
