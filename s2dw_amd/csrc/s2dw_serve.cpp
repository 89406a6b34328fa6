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
//
// Requests run on a fixed set of lanes, one stream each, at most
// kMaxLanes (the hardware queues a process gets by default): each request
// keeps all its launches on its lane's stream, so concurrent requests never
// share a hardware queue and a synthesis waiting for its analysis stalls
// only its own lane.
//
// Each lane has a launcher thread: the worker admits a request (plan,
// buffers, lane) and hands it to the lane's thread, which issues its copies
// and kernels, so the launches of concurrent requests proceed in parallel
// (a small transform's launch sequence costs more host time than its
// kernels take on the device).
//
// The worker never blocks on a transfer: host data goes through pinned
// staging buffers that the client threads fill (submit) and drain (wait),
// so every copy the worker issues is asynchronous; the client threads split
// large copies over threads.  Device and pinned
// buffers are pooled by size; pooled device bytes count against the budget
// and are the first thing freed when room is needed.
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
  hipStream_t own = nullptr;  // the plan's stream (its launches go to a lane's)
};

constexpr int kMaxLanes = 4;

// Buffers cached by exact size.
struct Pool {
  std::multimap<size_t, void *> free;
  long long bytes = 0;
  void *take(size_t n) {
    auto it = free.find(n);
    if (it == free.end()) return nullptr;
    void *q = it->second;
    free.erase(it);
    bytes -= (long long)n;
    return q;
  }
  void put(void *q, size_t n) {
    if (!q) return;
    free.emplace(n, q);
    bytes += (long long)n;
  }
};

struct Buf {
  void *ptr = nullptr;
  size_t n = 0;
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
  // pinned staging: inputs filled at submit, outputs drained at wait
  Buf h_in, h_in2, h_out, h_out2;
  // execution state
  PlanEntry *plan = nullptr;
  int lane = -1;
  void *d_in = nullptr, *d_out = nullptr, *d_sc = nullptr;
  size_t n_in = 0, n_out = 0, n_sc = 0;
  hipEvent_t done = nullptr;
  const void *src_wav = nullptr, *src_sc = nullptr;  // synthesis of a resident analysis: its outputs
  hipEvent_t src_ready = nullptr;                    // and their completion event
  bool issued = false;  // the lane thread has enqueued the work and recorded done
  int status = kErrNone;
  std::string error;  // detail text of a failure (raised on the thread that met it)
  bool finished = false;
  Clock::time_point t_submit, t_start, t_end;
};

struct Resident {
  void *d_wav = nullptr, *d_sc = nullptr;
  size_t n_wav = 0, n_sc = 0;
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
  int precision = S2DW_PREC_FP64;  // precision (and rounding form) of every plan the server builds
  long long next_ticket = 1;
  std::deque<std::shared_ptr<s2dw::Request>> queue;
  std::vector<std::shared_ptr<s2dw::Request>> inflight;
  std::map<long long, std::shared_ptr<s2dw::Request>> all;  // by ticket, until waited on
  std::map<long long, s2dw::Resident> resident;             // analysis outputs kept on the device
  std::multimap<s2dw::Key, s2dw::PlanEntry *> idle;         // idle plans by shape
  std::map<s2dw::Key, int> nplans;                          // plans built, per shape (idle or busy)
  long long completed = 0;
  std::vector<hipStream_t> lanes;  // one stream per concurrent request
  std::vector<char> lane_busy;
  std::vector<std::thread> lane_thr;                      // launcher thread per lane
  std::vector<std::shared_ptr<s2dw::Request>> lane_job;   // request handed to the lane, until issued
  std::condition_variable cv_lane;
  bool stop_lanes = false;
  s2dw::Pool dpool;               // device buffers (counted in in_use)
  std::mutex pin_mu;              // guards hpool (client threads and the worker)
  s2dw::Pool hpool;               // pinned host buffers
  long long pin_cap = 8LL << 30;  // pinned bytes kept cached at most
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
int build_plan(const Key &k, int precision, PlanEntry **out) {
  auto *e = new (std::nothrow) PlanEntry;
  if (!e) return kErrMemAlloc;
  std::vector<double> K((size_t)(k.J + 1) * k.B), Phi2(k.B), S((size_t)2 * k.B * k.N);
  int bl = 0;
  int err = init_kernels(K.data(), Phi2.data(), &bl, k.J, k.B, k.alpha);
  if (!err) err = init_directionality(S.data(), k.B, k.N);
  if (!err) err = plan_build(&e->p, k.B, k.N, k.J, k.alpha, K.data(), Phi2.data(), S.data(), precision, nullptr, true);
  if (!err) err = hchk(hipStreamSynchronize(e->p.stream), "plan set-up");
  if (err) {
    plan_free(&e->p);
    delete e;
    return err;
  }
  e->key = k;
  e->own = e->p.stream;
  e->p.single_stream = true;
  *out = e;
  return kErrNone;
}

void free_plan(s2dw_hip_server *s, PlanEntry *e) {
  s->in_use -= e->p.device_bytes;
  --s->nplans[e->key];
  e->p.stream = e->own;
  plan_free(&e->p);
  delete e;
}

// Byte sizes of a request's host-side arrays (and device buffers).
struct Sizes {
  size_t flm, wav, sc;
};
Sizes sizes_of(const Key &k) {
  const int bl = bl_of(k);
  return Sizes{(size_t)16 * k.B * k.B, (size_t)wav_doubles_of(k) * 8, (size_t)16 * bl * bl};
}

// Pinned host buffers, taken and returned by client threads and the worker.
bool pin_get(s2dw_hip_server *s, Buf &b, size_t n) {
  b = Buf{};
  if (n == 0) return true;
  {
    std::lock_guard<std::mutex> g(s->pin_mu);
    b.ptr = s->hpool.take(n);
  }
  if (!b.ptr && hipHostMalloc(&b.ptr, n, hipHostMallocPortable) != hipSuccess) {
    b.ptr = nullptr;
    return false;
  }
  b.n = n;
  return true;
}

void pin_put(s2dw_hip_server *s, Buf &b) {
  if (!b.ptr) return;
  bool keep;
  {
    std::lock_guard<std::mutex> g(s->pin_mu);
    keep = s->hpool.bytes + (long long)b.n <= s->pin_cap;
    if (keep) s->hpool.put(b.ptr, b.n);
  }
  if (!keep) (void)hipHostFree(b.ptr);
  b = Buf{};
}

void pin_release_all(s2dw_hip_server *s, Request &r) {
  pin_put(s, r.h_in);
  pin_put(s, r.h_in2);
  pin_put(s, r.h_out);
  pin_put(s, r.h_out2);
}

// Device buffers (worker, under the server lock): pooled bytes stay counted
// in in_use until freed.
void dev_put(s2dw_hip_server *s, void *q, size_t n) {
  if (q) s->dpool.put(q, n);
}

// Frees pooled device buffers, then idle plans (least recently used first),
// until need more bytes fit the budget.
bool evict_for(s2dw_hip_server *s, long long need) {
  while (s->in_use + need > s->budget && !s->dpool.free.empty()) {
    auto big = std::prev(s->dpool.free.end());
    (void)hipFree(big->second);
    s->in_use -= (long long)big->first;
    s->dpool.bytes -= (long long)big->first;
    s->dpool.free.erase(big);
  }
  while (s->in_use + need > s->budget && !s->idle.empty()) {
    auto lru = s->idle.begin();
    for (auto jt = s->idle.begin(); jt != s->idle.end(); ++jt)
      if (jt->second->last_used < lru->second->last_used) lru = jt;
    PlanEntry *victim = lru->second;
    s->idle.erase(lru);
    free_plan(s, victim);
  }
  return s->in_use + need <= s->budget;
}

// Device buffer sizes a request needs (0 = none): input, output, scoeff.
void dev_sizes(const Request &r, size_t &n_in, size_t &n_out, size_t &n_sc) {
  const Sizes z = sizes_of(r.key);
  if (r.kind == 0) {
    n_in = z.flm;
    n_out = z.wav;
    n_sc = z.sc;
  } else {
    n_out = z.flm;
    n_in = r.wav_from >= 0 ? 0 : z.wav;
    n_sc = r.wav_from >= 0 ? 0 : z.sc;
  }
}

// Bytes of those sizes the pool cannot supply.
long long new_bytes(s2dw_hip_server *s, size_t a, size_t b, size_t c) {
  std::vector<std::pair<size_t, void *>> took;
  long long miss = 0;
  for (size_t n : {a, b, c}) {
    if (!n) continue;
    void *q = s->dpool.take(n);
    if (q) took.emplace_back(n, q);
    else miss += (long long)n;
  }
  for (auto &t : took) s->dpool.put(t.second, t.first);
  return miss;
}

// Takes a pooled buffer of n bytes or allocates one (n = 0: none).
int dev_get(s2dw_hip_server *s, void **q, size_t n) {
  *q = nullptr;
  if (!n) return kErrNone;
  *q = s->dpool.take(n);
  if (*q) return kErrNone;
  if (hipMalloc(q, n) != hipSuccess) {
    *q = nullptr;
    set_error("hipMalloc of %zu bytes failed", n);
    return kErrMemAlloc;
  }
  s->in_use += (long long)n;
  return kErrNone;
}

// Tries to admit the queue head; returns true if it was launched (or failed).
// Where the wavelet coefficients a synthesis reads come from.
enum SrcState { kSrcReady, kSrcQueued, kSrcGone };

SrcState source_of(s2dw_hip_server *s, const Request &r, const void **wav, const void **sc, hipEvent_t *ready) {
  *wav = *sc = nullptr;
  *ready = nullptr;
  if (r.kind != 1 || r.wav_from < 0) return kSrcReady;
  auto rt = s->resident.find(r.wav_from);
  if (rt != s->resident.end()) {
    *wav = rt->second.d_wav;
    *sc = rt->second.d_sc;
    *ready = rt->second.ready;
    return kSrcReady;
  }
  for (auto &q : s->inflight)
    if (q->ticket == r.wav_from) {
      if (!q->issued) return kSrcQueued;  // its completion event is not recorded yet
      if (q->status != kErrNone) return kSrcGone;
      *wav = q->d_out;
      *sc = q->d_sc;
      *ready = q->done;
      return kSrcReady;
    }
  for (auto &q : s->queue)
    if (q->ticket == r.wav_from) return kSrcQueued;
  return kSrcGone;  // failed or released
}

bool admit_at(s2dw_hip_server *s, size_t qi, const void *src_wav, const void *src_sc, hipEvent_t src_ready);

// Admits the first request, within a bounded look-ahead, that can start now:
// its source (if it reads a resident analysis) is running or finished, and
// an idle plan of its shape exists or none of its shape does (a request
// waits for a busy plan of its shape rather than building a second one).
// Returns true if a request was launched or failed.
bool admit_one(s2dw_hip_server *s) {
  if (s->queue.empty() || (int)s->inflight.size() >= (int)s->lanes.size()) return false;
  constexpr size_t kLookahead = 64;
  const size_t look = std::min(s->queue.size(), kLookahead);
  for (size_t qi = 0; qi < look; ++qi) {
    const auto &r = s->queue[qi];
    const void *wav, *sc;
    hipEvent_t ready;
    const SrcState st = source_of(s, *r, &wav, &sc, &ready);
    if (st == kSrcQueued) continue;
    if (st == kSrcGone) {
      auto rr = r;
      s->queue.erase(s->queue.begin() + (long)qi);
      rr->error = "the analysis this synthesis reads failed or was released";
      rr->status = kErrArgInvalid;
      rr->finished = true;
      s->cv_done.notify_all();
      return true;
    }
    if (s->idle.find(r->key) == s->idle.end()) {
      auto np = s->nplans.find(r->key);
      if (np != s->nplans.end() && np->second > 0) continue;  // its plan is busy
    }
    return admit_at(s, qi, wav, sc, ready);
  }
  return false;
}

bool admit_at(s2dw_hip_server *s, size_t qi, const void *src_wav, const void *src_sc, hipEvent_t src_ready) {
  auto r = s->queue[qi];
  size_t n_in = 0, n_out = 0, n_sc = 0;
  dev_sizes(*r, n_in, n_out, n_sc);
  auto fail = [&](int code, const std::string &msg) {
    s->queue.erase(s->queue.begin() + (long)qi);
    r->error = msg;
    r->status = code;
    r->finished = true;
    s->cv_done.notify_all();
    return true;
  };
  PlanEntry *plan = nullptr;
  auto it = s->idle.find(r->key);
  if (it != s->idle.end()) {
    plan = it->second;
    s->idle.erase(it);
  }
  if (!plan) {
    // a plan's size is known only once built: make room for the buffers,
    // build, then check the budget again
    if (!evict_for(s, new_bytes(s, n_in, n_out, n_sc))) {
      if (s->inflight.empty()) return fail(kErrMemAlloc, "request buffers exceed the server's HBM budget");
      return false;
    }
    const int err = build_plan(r->key, s->precision, &plan);
    if (err) return fail(err, s2dw_hip_last_error());
    s->in_use += plan->p.device_bytes;
    ++s->nplans[r->key];
  }
  if (!evict_for(s, new_bytes(s, n_in, n_out, n_sc))) {
    if (!s->inflight.empty()) {
      // over budget for now: park the plan idle and retry when work retires
      plan->last_used = Clock::now();
      s->idle.emplace(r->key, plan);
      return false;
    }
    // nothing in flight and still over budget: the request can never fit
    free_plan(s, plan);
    set_error("request B=%d N=%d needs more than the server's HBM budget of %lld bytes", r->key.B, r->key.N,
              s->budget);
    return fail(kErrMemAlloc, s2dw_hip_last_error());
  }
  s->queue.erase(s->queue.begin() + (long)qi);
  r->plan = plan;
  r->t_start = Clock::now();
  Plan &p = plan->p;
  for (int i = 0; i < (int)s->lanes.size(); ++i)
    if (!s->lane_busy[i]) {
      r->lane = i;
      break;
    }
  s->lane_busy[r->lane] = 1;  // a lane is free: inflight < lanes
  p.stream = s->lanes[r->lane];
  int err = dev_get(s, &r->d_in, n_in);
  if (!err) err = dev_get(s, &r->d_out, n_out);
  if (!err) err = dev_get(s, &r->d_sc, n_sc);
  r->n_in = n_in;
  r->n_out = n_out;
  r->n_sc = n_sc;
  if (!err) err = hchk(hipEventCreateWithFlags(&r->done, hipEventDisableTiming), "event");
  r->src_wav = src_wav;
  r->src_sc = src_sc;
  r->src_ready = src_ready;
  if (err) {
    r->status = err;
    r->error = s2dw_hip_last_error();
    r->issued = true;  // nothing to issue
  } else {
    s->lane_job[r->lane] = r;  // the lane's thread issues the work
    s->cv_lane.notify_all();
  }
  s->inflight.push_back(r);
  return true;
}

// Enqueues a request's copies and kernels on its lane's stream and records
// its completion event (on the lane's thread, without the server lock: the
// request owns its plan, buffers and lane until it finishes).
int issue(Request *r) {
  Plan &p = r->plan->p;
  const hipStream_t st = p.stream;
  int err = kErrNone;
  // every copy is between device and pinned memory: asynchronous
  if (r->kind == 0) {
    err = hchk(hipMemcpyAsync(r->d_in, r->h_in.ptr, r->n_in, hipMemcpyHostToDevice, st), "copy in");
    if (!err) err = run_analysis(&p, (double *)r->d_out, (double2 *)r->d_sc, (const double2 *)r->d_in);
    if (!err && r->h_out.ptr)
      err = hchk(hipMemcpyAsync(r->h_out.ptr, r->d_out, r->n_out, hipMemcpyDeviceToHost, st), "copy out");
    if (!err && r->h_out2.ptr)
      err = hchk(hipMemcpyAsync(r->h_out2.ptr, r->d_sc, r->n_sc, hipMemcpyDeviceToHost, st), "copy out");
  } else {
    const void *wav = r->d_in, *sc = r->d_sc;
    if (r->wav_from >= 0) {
      wav = r->src_wav;
      sc = r->src_sc;
      err = hchk(hipStreamWaitEvent(st, r->src_ready, 0), "wait producer");
    } else {
      err = hchk(hipMemcpyAsync(r->d_in, r->h_in.ptr, r->n_in, hipMemcpyHostToDevice, st), "copy in");
      if (!err) err = hchk(hipMemcpyAsync(r->d_sc, r->h_in2.ptr, r->n_sc, hipMemcpyHostToDevice, st), "copy in");
    }
    if (!err) err = run_synthesis(&p, (double2 *)r->d_out, (const double *)wav, (const double2 *)sc);
    if (!err) err = hchk(hipMemcpyAsync(r->h_out.ptr, r->d_out, r->n_out, hipMemcpyDeviceToHost, st), "copy out");
  }
  if (!err) err = hchk(hipEventRecord(r->done, st), "event");
  return err;
}

void lane_loop(s2dw_hip_server *s, int lane) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    s->cv_lane.wait(lk, [&] { return s->stop_lanes || s->lane_job[lane] != nullptr; });
    if (!s->lane_job[lane]) break;  // stop_lanes and nothing handed over
    std::shared_ptr<Request> r = s->lane_job[lane];
    lk.unlock();
    const int err = issue(r.get());
    std::string msg = err ? s2dw_hip_last_error() : std::string();
    lk.lock();
    r->status = err;
    if (err) r->error = msg;
    r->issued = true;
    s->lane_job[lane] = nullptr;
    s->cv_work.notify_all();
  }
}

void finish(s2dw_hip_server *s, const std::shared_ptr<Request> &r) {
  r->t_end = Clock::now();
  if (r->status != kErrNone && r->lane >= 0) {
    // the lane may still run what was enqueued before the failure
    (void)hipStreamSynchronize(s->lanes[r->lane]);
  }
  if (r->lane >= 0) s->lane_busy[r->lane] = 0;
  r->lane = -1;
  if (r->kind == 0 && r->keep_resident && r->status == kErrNone) {
    Resident res;
    res.d_wav = r->d_out;
    res.d_sc = r->d_sc;
    res.n_wav = r->n_out;
    res.n_sc = r->n_sc;
    res.ready = r->done;  // already complete; later consumers wait on it trivially
    res.key = r->key;
    s->resident[r->ticket] = res;
    dev_put(s, r->d_in, r->n_in);
    r->d_in = r->d_out = r->d_sc = nullptr;
    r->done = nullptr;
  } else {
    dev_put(s, r->d_in, r->n_in);
    dev_put(s, r->d_out, r->n_out);
    dev_put(s, r->d_sc, r->n_sc);
    r->d_in = r->d_out = r->d_sc = nullptr;
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
      const bool done = r->issued && (r->status != kErrNone || hipEventQuery(r->done) == hipSuccess);
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
  const Clock::time_point t0 = Clock::now();
  if (r->kind == 1 && r->wav_from >= 0) {
    // checked again at admission (the source may be released meanwhile)
    std::lock_guard<std::mutex> g(s->mu);
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
  // stage the host data in pinned buffers here, on the client's thread
  const Sizes z = sizes_of(r->key);
  bool ok = true;
  if (r->kind == 0) {
    ok = pin_get(s, r->h_in, z.flm) && pin_get(s, r->h_out, r->out_host ? z.wav : 0) &&
         pin_get(s, r->h_out2, r->out_host2 ? z.sc : 0);
    if (ok) host_copy_mt(r->h_in.ptr, r->in_host, z.flm);
  } else {
    ok = pin_get(s, r->h_out, z.flm);
    if (ok && r->wav_from < 0) {
      ok = pin_get(s, r->h_in, z.wav) && pin_get(s, r->h_in2, z.sc);
      if (ok) {
        host_copy_mt(r->h_in.ptr, r->in_host, z.wav);
        host_copy_mt(r->h_in2.ptr, r->in_host2, z.sc);
      }
    }
  }
  if (!ok) {
    pin_release_all(s, *r);
    set_error("pinned host staging of %zu bytes failed", z.wav + z.flm + z.sc);
    return kErrMemAlloc;
  }
  std::lock_guard<std::mutex> g(s->mu);
  r->ticket = s->next_ticket++;
  r->t_submit = t0;  // latency includes the staging copy
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
  return s2dw_hip_server_create_ex(srv, hbm_budget_bytes, max_inflight, S2DW_PREC_FP64);
}

int s2dw_hip_server_create_ex(s2dw_hip_server_t *srv, long long hbm_budget_bytes, int max_inflight, int precision) {
  if (!srv || hbm_budget_bytes <= 0 || max_inflight < 1) return kErrArgInvalid;
  const int prec = precision & ~S2DW_ROUNDING_REFERENCE;
  if (prec != S2DW_PREC_FP64 && prec != S2DW_PREC_SPLIT) return kErrArgInvalid;
  auto *s = new (std::nothrow) s2dw_hip_server;
  if (!s) return kErrMemAlloc;
  s->precision = precision;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return kErrInitFail;
  }
  s->budget = hbm_budget_bytes;
  s->max_inflight = max_inflight;
  const int nl = std::min(max_inflight, kMaxLanes);
  for (int i = 0; i < nl; ++i) {
    hipStream_t q = nullptr;
    if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t o : s->lanes) (void)hipStreamDestroy(o);
      delete s;
      return kErrInitFail;
    }
    s->lanes.push_back(q);
    s->lane_busy.push_back(0);
  }
  s->lane_job.assign(nl, nullptr);
  for (int i = 0; i < nl; ++i) s->lane_thr.emplace_back(lane_loop, s, i);
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
  s->worker.join();  // the queue and the work in flight have drained
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop_lanes = true;
    s->cv_lane.notify_all();
  }
  for (auto &t : s->lane_thr) t.join();
  for (auto &kv : s->resident) {
    (void)hipFree(kv.second.d_wav);
    (void)hipFree(kv.second.d_sc);
    if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
  }
  for (auto &kv : s->idle) free_plan(s, kv.second);
  for (auto &kv : s->all) pin_release_all(s, *kv.second);  // never waited for
  for (auto &kv : s->dpool.free) (void)hipFree(kv.second);
  for (hipStream_t q : s->lanes) (void)hipStreamDestroy(q);
  for (auto &kv : s->hpool.free) (void)hipHostFree(kv.second);
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
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return kErrArgInvalid;
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->finished; });
    s->all.erase(ticket);
  }
  // outputs from the pinned staging buffers, on the caller's thread
  if (r->status == kErrNone) {
    if (r->h_out.ptr && r->out_host) host_copy_mt(r->out_host, r->h_out.ptr, r->h_out.n);
    if (r->h_out2.ptr && r->out_host2) host_copy_mt(r->out_host2, r->h_out2.ptr, r->h_out2.n);
  }
  pin_release_all(s, *r);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_submit).count();
  if (run_ms) *run_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_start).count();
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
  dev_put(s, it->second.d_wav, it->second.n_wav);
  dev_put(s, it->second.d_sc, it->second.n_sc);
  if (it->second.ready) (void)hipEventDestroy(it->second.ready);
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
