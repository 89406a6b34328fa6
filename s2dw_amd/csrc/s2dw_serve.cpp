// s2dw_serve.cpp -- request loop for many analyses and syntheses of uneven
// (B, N, J, alpha) on one GPU (continuous batching).
//
// A worker thread owns the device work.  Requests wait in a FIFO queue; the
// worker admits the head whenever an execution slot is free and the device
// footprint it needs (a plan of its shape, unless an idle one exists, plus
// its input and output buffers) fits the HBM budget, evicting idle plans
// when that makes room.  Each plan owns a stream, so admitted requests of
// any shapes run concurrently; a request holds its plan until its work has
// finished (plans carry work buffers).  A synthesis may consume the wavelet
// coefficients an earlier analysis left resident: its stream waits on the
// analysis's completion event, so it can be admitted while the analysis
// still runs.  Results do not depend on what else is in flight: every
// request runs the same kernels, in the same order, on buffers of its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "s2dw_hip.h"
#include "s2dw_internal.h"

namespace s2dw {
namespace {

using Clock = std::chrono::steady_clock;

struct Key {
  int B, N, J;
  double alpha;
  bool operator<(const Key &o) const { return std::tie(B, N, J, alpha) < std::tie(o.B, o.N, o.J, o.alpha); }
};

struct PlanEntry {
  Plan p;
  Key key{};
  Clock::time_point last_used;
};

struct Request {
  long long ticket = 0;
  int kind = 0;  // 0 analysis, 1 synthesis
  Key key{};
  const double *in_host = nullptr;   // analysis: flm; synthesis: wav (if no resident source)
  const double *in_host2 = nullptr;  // synthesis: scoeff
  double *out_host = nullptr;        // analysis: wav (optional); synthesis: flm
  double *out_host2 = nullptr;       // analysis: scoeff
  long long wav_from = -1;           // synthesis: ticket of the analysis whose wav to use
  bool keep_resident = false;        // analysis: keep wav + scoeff on the device
  // execution state
  PlanEntry *plan = nullptr;
  void *d_in = nullptr, *d_out = nullptr, *d_sc = nullptr;
  long long bytes = 0;  // device bytes held by this request's buffers
  hipEvent_t done = nullptr;
  int status = kErrNone;
  std::string error;  // detail text of a failure (raised on the worker thread)
  bool finished = false;
  Clock::time_point t_submit, t_start, t_end;
};

struct Resident {
  void *d_wav = nullptr, *d_sc = nullptr;
  long long bytes = 0;
  hipEvent_t ready = nullptr;  // recorded by the producing analysis
  Key key{};
};

}  // namespace
}  // namespace s2dw

struct s2dw_hip_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  int device = 0;
  long long budget = 0, in_use = 0;
  int max_inflight = 1;
  long long next_ticket = 1;
  std::deque<std::shared_ptr<s2dw::Request>> queue;
  std::vector<std::shared_ptr<s2dw::Request>> inflight;
  std::map<long long, std::shared_ptr<s2dw::Request>> all;  // by ticket, until waited on
  std::map<long long, s2dw::Resident> resident;             // analysis outputs kept on the device
  std::multimap<s2dw::Key, s2dw::PlanEntry *> idle;         // idle plans by shape
  long long completed = 0;
};

namespace s2dw {
namespace {

// 0 on success, else the library's error code with the HIP error recorded
int hchk(hipError_t e, const char *what) { return e == hipSuccess ? kErrNone : hip_error(e, what); }

long long wav_doubles_of(const Key &k) {
  long long n = 0;
  for (int j = 0; j <= k.J; ++j) {
    int hi = 0, lo = 0;
    band_limits(j, k.B, k.alpha, &hi, &lo);
    n += (long long)(2 * hi - 1) * 2 * hi * k.N;
  }
  return n;
}

int bl_of(const Key &k) {
  int hi = 0, lo = 0;
  band_limits(k.J + 1, k.B, k.alpha, &hi, &lo);
  return hi;
}

// Builds a plan for the key (host setup + device tables) on its own stream.
int build_plan(const Key &k, PlanEntry **out) {
  auto *e = new (std::nothrow) PlanEntry;
  if (!e) return kErrMemAlloc;
  std::vector<double> K((size_t)(k.J + 1) * k.B), Phi2(k.B), S((size_t)2 * k.B * k.N);
  int bl = 0;
  int err = init_kernels(K.data(), Phi2.data(), &bl, k.J, k.B, k.alpha);
  if (!err) err = init_directionality(S.data(), k.B, k.N);
  if (!err) err = plan_build(&e->p, k.B, k.N, k.J, k.alpha, K.data(), Phi2.data(), S.data(), S2DW_PREC_FP64, nullptr, true);
  if (!err) err = hchk(hipStreamSynchronize(e->p.stream), "plan set-up");
  if (err) {
    plan_free(&e->p);
    delete e;
    return err;
  }
  e->key = k;
  *out = e;
  return kErrNone;
}

void free_plan(s2dw_hip_server *s, PlanEntry *e) {
  s->in_use -= e->p.device_bytes;
  plan_free(&e->p);
  delete e;
}

// Device bytes a request needs besides its plan.
long long request_bytes(const Request &r) {
  const long long wav = wav_doubles_of(r.key) * 8, sc = 16LL * bl_of(r.key) * bl_of(r.key);
  const long long flm = 16LL * r.key.B * r.key.B;
  if (r.kind == 0) return flm + wav + sc;
  return flm + (r.wav_from >= 0 ? 0 : wav + sc);
}

// Tries to admit the queue head; returns true if it was launched (or failed).
bool admit_one(s2dw_hip_server *s) {
  if (s->queue.empty() || (int)s->inflight.size() >= s->max_inflight) return false;
  auto r = s->queue.front();
  // a synthesis reading an analysis result: the producer was submitted
  // earlier, so it is resident (finished) or still running; if neither, it
  // failed and so does this request
  const void *src_wav = nullptr, *src_sc = nullptr;
  hipEvent_t src_ready = nullptr;
  if (r->kind == 1 && r->wav_from >= 0) {
    auto rt = s->resident.find(r->wav_from);
    if (rt != s->resident.end()) {
      src_wav = rt->second.d_wav;
      src_sc = rt->second.d_sc;
      src_ready = rt->second.ready;
    } else {
      for (auto &q : s->inflight)
        if (q->ticket == r->wav_from && q->status == kErrNone) {
          src_wav = q->d_out;
          src_sc = q->d_sc;
          src_ready = q->done;
        }
    }
    if (!src_wav) {
      s->queue.pop_front();
      r->error = "the analysis this synthesis reads failed or was released";
      r->status = kErrArgInvalid;
      r->finished = true;
      s->cv_done.notify_all();
      return true;
    }
  }
  const long long need_buf = request_bytes(*r);
  PlanEntry *plan = nullptr;
  auto it = s->idle.find(r->key);
  if (it != s->idle.end()) {
    plan = it->second;
    s->idle.erase(it);
  }
  auto evict_for = [&](long long need) {
    while (s->in_use + need > s->budget && !s->idle.empty()) {
      auto lru = s->idle.begin();
      for (auto jt = s->idle.begin(); jt != s->idle.end(); ++jt)
        if (jt->second->last_used < lru->second->last_used) lru = jt;
      PlanEntry *victim = lru->second;
      s->idle.erase(lru);
      free_plan(s, victim);
    }
    return s->in_use + need <= s->budget;
  };
  if (!plan) {
    // a plan's size is known only once built: build, then check the budget
    if (!evict_for(need_buf)) {
      if (s->inflight.empty()) {  // nothing will free memory: the request cannot fit
        s->queue.pop_front();
        r->error = "request buffers exceed the server's HBM budget";
        r->status = kErrMemAlloc;
        r->finished = true;
        s->cv_done.notify_all();
        return true;
      }
      return false;
    }
    const int err = build_plan(r->key, &plan);
    if (err) {
      s->queue.pop_front();
      r->error = s2dw_hip_last_error();
      r->status = err;
      r->finished = true;
      s->cv_done.notify_all();
      return true;
    }
    s->in_use += plan->p.device_bytes;
  }
  if (!evict_for(need_buf)) {
    if (!s->inflight.empty()) {
      // over budget for now: park the plan idle and retry when work retires
      plan->last_used = Clock::now();
      s->idle.emplace(r->key, plan);
      return false;
    }
    // nothing in flight and still over budget: the request can never fit
    free_plan(s, plan);
    s->queue.pop_front();
    set_error("request B=%d N=%d needs more than the server's HBM budget of %lld bytes", r->key.B, r->key.N,
              s->budget);
    r->error = s2dw_hip_last_error();
    r->status = kErrMemAlloc;
    r->finished = true;
    s->cv_done.notify_all();
    return true;
  }
  s->queue.pop_front();
  r->plan = plan;
  r->t_start = Clock::now();
  Plan &p = plan->p;
  const long long wav_b = p.wav_doubles * 8, sc_b = 16LL * p.bl_scoeff * p.bl_scoeff, flm_b = 16LL * p.B * p.B;
  int err = kErrNone;
  auto alloc = [&](void **ptr, long long n) {
    if (!err && hipMalloc(ptr, (size_t)std::max(1LL, n)) != hipSuccess) err = kErrMemAlloc;
  };
  if (r->kind == 0) {
    alloc(&r->d_in, flm_b);
    alloc(&r->d_out, wav_b);
    alloc(&r->d_sc, sc_b);
  } else {
    alloc(&r->d_out, flm_b);
    if (r->wav_from < 0) {
      alloc(&r->d_in, wav_b);
      alloc(&r->d_sc, sc_b);
    }
  }
  r->bytes = need_buf;
  s->in_use += need_buf;
  if (!err) err = hchk(hipEventCreateWithFlags(&r->done, hipEventDisableTiming), "event");
  hipStream_t st = p.stream;
  if (!err && r->kind == 0) {
    err = hchk(hipMemcpyAsync(r->d_in, r->in_host, flm_b, hipMemcpyHostToDevice, st), "copy in");
    if (!err) err = run_analysis(&p, (double *)r->d_out, (double2 *)r->d_sc, (const double2 *)r->d_in);
    if (!err && r->out_host)
      err = hchk(hipMemcpyAsync(r->out_host, r->d_out, wav_b, hipMemcpyDeviceToHost, st), "copy out");
    if (!err && r->out_host2)
      err = hchk(hipMemcpyAsync(r->out_host2, r->d_sc, sc_b, hipMemcpyDeviceToHost, st), "copy out");
  } else if (!err) {
    const void *wav = r->d_in, *sc = r->d_sc;
    if (r->wav_from >= 0) {
      wav = src_wav;
      sc = src_sc;
      err = hchk(hipStreamWaitEvent(st, src_ready, 0), "wait producer");
    } else {
      err = hchk(hipMemcpyAsync(r->d_in, r->in_host, wav_b, hipMemcpyHostToDevice, st), "copy in");
      if (!err) err = hchk(hipMemcpyAsync(r->d_sc, r->in_host2, sc_b, hipMemcpyHostToDevice, st), "copy in");
    }
    if (!err) err = run_synthesis(&p, (double2 *)r->d_out, (const double *)wav, (const double2 *)sc);
    if (!err) err = hchk(hipMemcpyAsync(r->out_host, r->d_out, flm_b, hipMemcpyDeviceToHost, st), "copy out");
  }
  if (!err) err = hchk(hipEventRecord(r->done, st), "event");
  r->status = err;
  if (err) r->error = s2dw_hip_last_error();
  s->inflight.push_back(r);
  return true;
}

void finish(s2dw_hip_server *s, const std::shared_ptr<Request> &r) {
  r->t_end = Clock::now();
  if (r->kind == 0 && r->keep_resident && r->status == kErrNone) {
    Resident res;
    res.d_wav = r->d_out;
    res.d_sc = r->d_sc;
    res.bytes = r->bytes;
    res.ready = r->done;  // already complete; later consumers wait on it trivially
    res.key = r->key;
    s->resident[r->ticket] = res;
    r->d_out = r->d_sc = nullptr;
    r->done = nullptr;
    (void)hipFree(r->d_in);
    r->d_in = nullptr;
    const long long flm_b = 16LL * r->key.B * r->key.B;
    s->resident[r->ticket].bytes = r->bytes - flm_b;
    s->in_use -= flm_b;
  } else {
    (void)hipFree(r->d_in);
    (void)hipFree(r->d_out);
    (void)hipFree(r->d_sc);
    s->in_use -= r->bytes;
    if (r->done) (void)hipEventDestroy(r->done);
    r->done = nullptr;
  }
  if (r->plan) {
    r->plan->last_used = Clock::now();
    s->idle.emplace(r->key, r->plan);
    r->plan = nullptr;
  }
  r->finished = true;
  ++s->completed;
}

void worker_loop(s2dw_hip_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    // retire finished work
    for (size_t i = 0; i < s->inflight.size();) {
      auto &r = s->inflight[i];
      const bool done = r->status != kErrNone || hipEventQuery(r->done) == hipSuccess;
      if (done) {
        finish(s, r);
        s->inflight.erase(s->inflight.begin() + (long)i);
        s->cv_done.notify_all();
      } else {
        ++i;
      }
    }
    if (s->stop && s->queue.empty() && s->inflight.empty()) break;
    bool progressed = false;
    while (admit_one(s)) progressed = true;
    if (!progressed) {
      if (s->inflight.empty()) s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      else s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
  }
}

int submit(s2dw_hip_server *s, std::shared_ptr<Request> r, long long *ticket) {
  if (r->key.B < 1 || r->key.N < 1 || r->key.N > r->key.B || !params_valid(r->key.J, r->key.B, r->key.alpha, r->key.N)) {
    set_error("invalid sizes B=%d N=%d J=%d alpha=%g", r->key.B, r->key.N, r->key.J, r->key.alpha);
    return kErrSizeInvalid;
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (r->kind == 1 && r->wav_from >= 0) {
    auto it = s->resident.find(r->wav_from);
    auto pend = s->all.find(r->wav_from);
    const bool ok = (it != s->resident.end() && it->second.key.B == r->key.B && it->second.key.N == r->key.N &&
                     it->second.key.J == r->key.J && it->second.key.alpha == r->key.alpha) ||
                    (pend != s->all.end() && pend->second->kind == 0 && pend->second->keep_resident &&
                     !(pend->second->key < r->key) && !(r->key < pend->second->key));
    if (!ok) {
      set_error("ticket %lld holds no resident wavelet coefficients of this shape", r->wav_from);
      return kErrArgInvalid;
    }
  }
  r->ticket = s->next_ticket++;
  r->t_submit = Clock::now();
  s->all[r->ticket] = r;
  s->queue.push_back(r);
  *ticket = r->ticket;
  s->cv_work.notify_all();
  return kErrNone;
}

}  // namespace
}  // namespace s2dw

using namespace s2dw;

extern "C" {

int s2dw_hip_server_create(s2dw_hip_server_t *srv, long long hbm_budget_bytes, int max_inflight) {
  if (!srv || hbm_budget_bytes <= 0 || max_inflight < 1) return kErrArgInvalid;
  auto *s = new (std::nothrow) s2dw_hip_server;
  if (!s) return kErrMemAlloc;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return kErrInitFail;
  }
  s->budget = hbm_budget_bytes;
  s->max_inflight = max_inflight;
  s->worker = std::thread(worker_loop, s);
  *srv = s;
  return kErrNone;
}

int s2dw_hip_server_destroy(s2dw_hip_server_t s) {
  if (!s) return kErrArgInvalid;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  s->worker.join();
  for (auto &kv : s->resident) {
    (void)hipFree(kv.second.d_wav);
    (void)hipFree(kv.second.d_sc);
    if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
  }
  for (auto &kv : s->idle) free_plan(s, kv.second);
  delete s;
  return kErrNone;
}

int s2dw_hip_server_submit_analysis(s2dw_hip_server_t s, int B, int N, int J, double alpha, const double *flm,
                                    double *wav, double *scoeff, int keep_resident, long long *ticket) {
  if (!s || !flm || !ticket || (!scoeff && !keep_resident)) return kErrArgInvalid;
  auto r = std::make_shared<Request>();
  r->kind = 0;
  r->key = Key{B, N, J, alpha};
  r->in_host = flm;
  r->out_host = wav;
  r->out_host2 = scoeff;
  r->keep_resident = keep_resident != 0;
  return submit(s, r, ticket);
}

int s2dw_hip_server_submit_synthesis(s2dw_hip_server_t s, int B, int N, int J, double alpha, long long wav_from,
                                     const double *wav, const double *scoeff, double *flm, long long *ticket) {
  if (!s || !flm || !ticket || (wav_from < 0 && (!wav || !scoeff))) return kErrArgInvalid;
  auto r = std::make_shared<Request>();
  r->kind = 1;
  r->key = Key{B, N, J, alpha};
  r->in_host = wav;
  r->in_host2 = scoeff;
  r->out_host = flm;
  r->wav_from = wav_from;
  return submit(s, r, ticket);
}

int s2dw_hip_server_wait(s2dw_hip_server_t s, long long ticket, double *latency_ms, double *run_ms) {
  if (!s) return kErrArgInvalid;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return kErrArgInvalid;
  auto r = it->second;
  s->cv_done.wait(lk, [&] { return r->finished; });
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_submit).count();
  if (run_ms) *run_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_start).count();
  s->all.erase(ticket);
  if (r->status != kErrNone) set_error("%s", r->error.c_str());
  return r->status;
}

int s2dw_hip_server_release(s2dw_hip_server_t s, long long ticket) {
  if (!s) return kErrArgInvalid;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->resident.find(ticket);
  if (it == s->resident.end()) return kErrArgInvalid;
  // a queued or running synthesis may still read it
  for (auto &q : s->queue)
    if (q->wav_from == ticket) return kErrArgInvalid;
  for (auto &q : s->inflight)
    if (q->wav_from == ticket) return kErrArgInvalid;
  (void)hipFree(it->second.d_wav);
  (void)hipFree(it->second.d_sc);
  if (it->second.ready) (void)hipEventDestroy(it->second.ready);
  s->in_use -= it->second.bytes;
  s->resident.erase(it);
  return kErrNone;
}

int s2dw_hip_server_stats(s2dw_hip_server_t s, long long *in_use_bytes, int *inflight, int *queued,
                          long long *completed) {
  if (!s) return kErrArgInvalid;
  std::lock_guard<std::mutex> g(s->mu);
  if (in_use_bytes) *in_use_bytes = s->in_use;
  if (inflight) *inflight = (int)s->inflight.size();
  if (queued) *queued = (int)s->queue.size();
  if (completed) *completed = s->completed;
  return kErrNone;
}

}  // extern "C"
