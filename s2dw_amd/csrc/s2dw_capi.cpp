// s2dw_capi.cpp -- the C ABI of libs2dw_hip.so (declared in include/s2dw_hip.h).
//
// Host-pointer entry points mirror the reference's s2dw_core_mod procedures
// (synchronous, same arguments and layouts).  They run on a cached plan per
// (device, B, N, J, alpha, K_gamma, Phi2, Slm): the plan's device tables are
// built once, later calls with the same configuration only move data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <list>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/s2dw_hip.h"
#include "s2dw_internal.h"

struct s2dw_hip_plan {
  s2dw::Plan p;
};

namespace s2dw {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_error(hipError_t e, const char *what) {
  set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return e == hipErrorOutOfMemory ? kErrMemAlloc : kErrInitFail;
}

// Host memcpy split over OpenMP threads in 1 MB pieces (one thread below
// 2 MB): a single thread copies ~10 GB/s, well below what the host and PCIe
// carry.
void host_copy_mt(void *dst, const void *src, size_t n) {
  constexpr size_t kPiece = 1u << 20;
  const long long np = (long long)((n + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static) if (np > 2)
  for (long long i = 0; i < np; ++i) {
    const size_t o = (size_t)i * kPiece;
    std::memcpy(static_cast<char *>(dst) + o, static_cast<const char *>(src) + o, std::min(kPiece, n - o));
  }
}

namespace {

// Cached plans and staging buffers for the host-pointer entry points.
struct HostEntry {
  Plan plan;
  double2 *d_flm = nullptr, *d_sc = nullptr;
  double *d_wav = nullptr, *d_static = nullptr;
  size_t static_doubles = 0;
  ~HostEntry() {
    if (d_flm) (void)hipFree(d_flm);
    if (d_sc) (void)hipFree(d_sc);
    if (d_wav) (void)hipFree(d_wav);
    if (d_static) (void)hipFree(d_static);
    plan_free(&plan);
  }
};

std::mutex g_mu;
std::list<std::unique_ptr<HostEntry>> g_cache;  // most recent first
constexpr size_t kCacheMax = 4;

// Copies between the caller's (pageable) arrays and the device.  Large ones
// go through two pinned staging buffers: the DMA of one chunk overlaps the
// host copy of the other, and host copies are split over threads.  The
// buffers (2 x kStageBytes of pinned memory) are allocated on first use and
// kept until the process ends; g_mu guards them.
constexpr size_t kStageBytes = 32u << 20, kStageMin = 8u << 20;
void *g_stage[2] = {nullptr, nullptr};
hipEvent_t g_stage_ev[2] = {nullptr, nullptr};
bool g_stage_rec[2] = {false, false};  // the event has been recorded

int stage_ready() {
  for (int i = 0; i < 2; ++i) {
    if (!g_stage[i]) S2DW_HIP_CHECK(hipHostMalloc(&g_stage[i], kStageBytes, hipHostMallocDefault));
    if (!g_stage_ev[i]) S2DW_HIP_CHECK(hipEventCreateWithFlags(&g_stage_ev[i], hipEventDisableTiming));
    // a previous copy's transfers out of the buffer have finished
    if (g_stage_rec[i]) S2DW_HIP_CHECK(hipEventSynchronize(g_stage_ev[i]));
  }
  return kErrNone;
}

int stage_record(int i, hipStream_t st) {
  S2DW_HIP_CHECK(hipEventRecord(g_stage_ev[i], st));
  g_stage_rec[i] = true;
  return kErrNone;
}

void host_copy(void *dst, const void *src, size_t n) { host_copy_mt(dst, src, n); }

// Device to host, in stream order after the work already on st; large
// copies have completed on return, small ones at the next synchronisation.
int copy_d2h(void *dst, const void *src, size_t n, hipStream_t st) {
  if (n < kStageMin) {
    S2DW_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, st));
    return kErrNone;
  }
  if (int e = stage_ready()) return e;
  const size_t nch = (n + kStageBytes - 1) / kStageBytes;
  for (size_t i = 0; i <= nch; ++i) {
    if (i < nch) {
      const size_t o = i * kStageBytes, len = std::min(kStageBytes, n - o);
      S2DW_HIP_CHECK(hipMemcpyAsync(g_stage[i & 1], static_cast<const char *>(src) + o, len, hipMemcpyDeviceToHost, st));
      if (int e = stage_record(i & 1, st)) return e;
    }
    if (i > 0) {
      const size_t k = i - 1, o = k * kStageBytes, len = std::min(kStageBytes, n - o);
      S2DW_HIP_CHECK(hipEventSynchronize(g_stage_ev[k & 1]));
      host_copy(static_cast<char *>(dst) + o, g_stage[k & 1], len);
    }
  }
  return kErrNone;
}

// Host to device, in stream order on st; the source may be reused on return
// (large copies have left it for the staging buffers).
int copy_h2d(void *dst, const void *src, size_t n, hipStream_t st) {
  if (n < kStageMin) {
    S2DW_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st));
    return kErrNone;
  }
  if (int e = stage_ready()) return e;
  const size_t nch = (n + kStageBytes - 1) / kStageBytes;
  for (size_t i = 0; i < nch; ++i) {
    const size_t o = i * kStageBytes, len = std::min(kStageBytes, n - o);
    if (i >= 2) S2DW_HIP_CHECK(hipEventSynchronize(g_stage_ev[i & 1]));  // chunk i - 2 has left the buffer
    host_copy(g_stage[i & 1], static_cast<const char *>(src) + o, len);
    S2DW_HIP_CHECK(hipMemcpyAsync(static_cast<char *>(dst) + o, g_stage[i & 1], len, hipMemcpyHostToDevice, st));
    if (int e = stage_record(i & 1, st)) return e;
  }
  return kErrNone;
}

// Rounding form of the host-buffer entry points: the reference's own
// (default), or S2DW_HIP_ROUNDING=streamed / table for the other forms of
// the plain transform (read at each call).
int host_form() {
  const char *ev = std::getenv("S2DW_HIP_ROUNDING");
  if (ev && !std::strcmp(ev, "streamed")) return S2DW_WIGNER_STREAMED;
  if (ev && !std::strcmp(ev, "table")) return S2DW_WIGNER_TABLE;
  return S2DW_ROUNDING_REFERENCE;
}

int plan_form(const Plan &p) {
  return p.refround ? S2DW_ROUNDING_REFERENCE : p.d_Sa ? S2DW_WIGNER_TABLE : S2DW_WIGNER_STREAMED;
}

bool same_config(const Plan &p, int dev, int B, int N, int J, double alpha, const double *K, const double *Phi2,
                 const double *S) {
  if (p.device != dev || p.B != B || p.N != N || p.J != J || p.alpha != alpha) return false;
  if (p.form_req != host_form()) return false;
  if (std::memcmp(p.K.data(), K, sizeof(double) * (size_t)(J + 1) * B)) return false;
  if (Phi2 && std::memcmp(p.Phi2.data(), Phi2, sizeof(double) * (size_t)B)) return false;
  if (std::memcmp(p.S.data(), S, sizeof(double) * 2 * (size_t)B * N)) return false;
  return true;
}

// Returns a plan for the configuration; Phi2 may be NULL for analysis (it is
// not used there), in which case a matching cached plan of any Phi2 is taken.
int get_entry(int B, int N, int J, double alpha, const double *K, const double *Phi2, const double *S,
              HostEntry **out) {
  int dev = 0;
  S2DW_HIP_CHECK(hipGetDevice(&dev));
  for (auto it = g_cache.begin(); it != g_cache.end(); ++it) {
    if (same_config((*it)->plan, dev, B, N, J, alpha, K, Phi2, S)) {
      g_cache.splice(g_cache.begin(), g_cache, it);
      *out = g_cache.front().get();
      return kErrNone;
    }
  }
  auto e = std::make_unique<HostEntry>();
  std::vector<double> phi(B, 0.0);
  if (Phi2) std::memcpy(phi.data(), Phi2, sizeof(double) * B);
  // the drop-in reproduces the reference's own rounding unless asked otherwise
  const int form = host_form();
  int err = plan_build(&e->plan, B, N, J, alpha, K, phi.data(), S, S2DW_PREC_FP64 | form, nullptr, true);
  e->plan.form_req = form;
  if (err) return err;
  const Plan &p = e->plan;
  S2DW_HIP_CHECK(hipMalloc(&e->d_flm, sizeof(double2) * (size_t)B * B));
  S2DW_HIP_CHECK(hipMalloc(&e->d_sc, sizeof(double2) * (size_t)p.bl_scoeff * p.bl_scoeff));
  S2DW_HIP_CHECK(hipMalloc(&e->d_wav, sizeof(double) * (size_t)p.wav_doubles));
  g_cache.push_front(std::move(e));
  while (g_cache.size() > kCacheMax) g_cache.pop_back();
  *out = g_cache.front().get();
  return kErrNone;
}

int ensure_static(HostEntry *e) {
  const Plan &p = e->plan;
  const size_t n = (size_t)(p.J + 1) * (2 * p.B - 1) * (2 * p.B) * p.N;
  if (e->static_doubles < n) {
    if (e->d_static) (void)hipFree(e->d_static);
    e->d_static = nullptr;
    S2DW_HIP_CHECK(hipMalloc(&e->d_static, sizeof(double) * n));
    e->static_doubles = n;
  }
  return kErrNone;
}

int check_bl(const Plan &p, int bl_scoeff) {
  if (bl_scoeff != p.bl_scoeff) {
    set_error("bl_scoeff=%d does not match the kernels' band limit %d", bl_scoeff, p.bl_scoeff);
    return kErrSizeInvalid;
  }
  return kErrNone;
}

}  // namespace
}  // namespace s2dw

using namespace s2dw;

extern "C" {

const char *s2dw_hip_last_error(void) { return g_err; }

const char *s2dw_hip_version(void) { return "s2dw_hip 0.1 (gfx950)"; }

int s2dw_hip_init_kernels(double *K_gamma, double *Phi2, int *bl_scoeff, int J, int B, double alpha) {
  if (!K_gamma || !Phi2 || !bl_scoeff) return kErrArgInvalid;
  const int e = init_kernels(K_gamma, Phi2, bl_scoeff, J, B, alpha);
  if (e == kErrSizeInvalid) set_error("s2dw_core_init_kernels: invalid sizes J=%d B=%d alpha=%g", J, B, alpha);
  if (e == kErrQuadStepExceed) set_error("qsimp: exceeded number of steps");
  return e;
}

int s2dw_hip_init_directionality(double *Slm, int B, int N) {
  if (!Slm) return kErrArgInvalid;
  const int e = init_directionality(Slm, B, N);
  if (e) set_error("s2dw_core_init_directionality: N=%d > B=%d", N, B);
  return e;
}

int s2dw_hip_comp_Jmax(int B, double alpha) { return comp_Jmax(B, alpha); }

int s2dw_hip_admiss(double *a, const double *K_gamma, const double *Phi2, int B, int J, int *pass) {
  if (!a || !K_gamma || !Phi2) return kErrArgInvalid;
  const bool ok = admiss(a, K_gamma, Phi2, B, J);
  if (pass) *pass = ok ? 1 : 0;
  return kErrNone;
}

int s2dw_hip_assimilate_int(double x, int *y) {
  int v = 0;
  const bool ok = assimilate_int(x, &v);
  if (y) *y = v;
  return ok ? 1 : 0;
}

int s2dw_hip_band_limits(int J, int B, double alpha, int *bl_hi, int *bl_lo) {
  for (int j = 0; j <= J; ++j) {
    int h, l;
    band_limits(j, B, alpha, &h, &l);
    if (bl_hi) bl_hi[j] = h;
    if (bl_lo) bl_lo[j] = l;
  }
  return kErrNone;
}

static int analysis_common(double *const *wavdyn, double *wav_static, double *scoeff, const double *flm,
                           const double *K_gamma, const double *Slm, int J, int B, int N, int bl_scoeff,
                           double alpha) {
  if (!scoeff || !flm || !K_gamma || !Slm) return kErrArgInvalid;
  std::lock_guard<std::mutex> lk(g_mu);
  HostEntry *e = nullptr;
  int err = get_entry(B, N, J, alpha, K_gamma, nullptr, Slm, &e);
  if (err) return err;
  Plan &p = e->plan;
  if ((err = check_bl(p, bl_scoeff))) return err;
  S2DW_HIP_CHECK(hipMemcpyAsync(e->d_flm, flm, sizeof(double2) * (size_t)B * B, hipMemcpyHostToDevice, p.stream));
  if ((err = run_analysis(&p, e->d_wav, e->d_sc, e->d_flm))) return err;
  if (wav_static) {
    if ((err = ensure_static(e))) return err;
    if ((err = pack_static(&p, e->d_wav, e->d_static, true))) return err;
    const size_t n = (size_t)(J + 1) * (2 * B - 1) * (2 * B) * N;
    if ((err = copy_d2h(wav_static, e->d_static, sizeof(double) * n, p.stream))) return err;
  } else {
    for (int j = 0; j <= J; ++j) {
      const Scale &s = p.sc[j];
      const size_t n = (size_t)(2 * s.hi - 1) * 2 * s.hi * N;
      if ((err = copy_d2h(wavdyn[j], e->d_wav + s.wav_off, sizeof(double) * n, p.stream))) return err;
    }
  }
  S2DW_HIP_CHECK(hipMemcpyAsync(scoeff, e->d_sc, sizeof(double2) * (size_t)bl_scoeff * bl_scoeff,
                                hipMemcpyDeviceToHost, p.stream));
  S2DW_HIP_CHECK(hipStreamSynchronize(p.stream));
  return kErrNone;
}

int s2dw_hip_analysis_flm2wav(double *wav, double *scoeff, const double *flm, const double *K_gamma,
                              const double *Slm, int J, int B, int N, int bl_scoeff, double alpha) {
  if (!wav) return kErrArgInvalid;
  return analysis_common(nullptr, wav, scoeff, flm, K_gamma, Slm, J, B, N, bl_scoeff, alpha);
}

int s2dw_hip_analysis_flm2wav_dynamic(double *const *wavdyn, double *scoeff, const double *flm,
                                      const double *K_gamma, const double *Slm, int J, int B, int N,
                                      int bl_scoeff, double alpha) {
  if (!wavdyn) return kErrArgInvalid;
  for (int j = 0; j <= J; ++j)
    if (!wavdyn[j]) return kErrArgInvalid;
  return analysis_common(wavdyn, nullptr, scoeff, flm, K_gamma, Slm, J, B, N, bl_scoeff, alpha);
}

static int synthesis_common(double *flm, const double *const *wavdyn, const double *wav_static, const double *scoeff,
                            const double *K_gamma, const double *Phi2, const double *Slm, int J, int B, int N,
                            int bl_scoeff, double alpha) {
  if (!flm || !scoeff || !K_gamma || !Phi2 || !Slm) return kErrArgInvalid;
  std::lock_guard<std::mutex> lk(g_mu);
  HostEntry *e = nullptr;
  int err = get_entry(B, N, J, alpha, K_gamma, Phi2, Slm, &e);
  if (err) return err;
  Plan &p = e->plan;
  if ((err = check_bl(p, bl_scoeff))) return err;
  if (wav_static) {
    if ((err = ensure_static(e))) return err;
    const size_t n = (size_t)(J + 1) * (2 * B - 1) * (2 * B) * N;
    if ((err = copy_h2d(e->d_static, wav_static, sizeof(double) * n, p.stream))) return err;
    if ((err = pack_static(&p, e->d_wav, e->d_static, false))) return err;
  } else {
    for (int j = 0; j <= J; ++j) {
      const Scale &s = p.sc[j];
      const size_t n = (size_t)(2 * s.hi - 1) * 2 * s.hi * N;
      if ((err = copy_h2d(e->d_wav + s.wav_off, wavdyn[j], sizeof(double) * n, p.stream))) return err;
    }
  }
  S2DW_HIP_CHECK(hipMemcpyAsync(e->d_sc, scoeff, sizeof(double2) * (size_t)bl_scoeff * bl_scoeff,
                                hipMemcpyHostToDevice, p.stream));
  if ((err = run_synthesis(&p, e->d_flm, e->d_wav, e->d_sc))) return err;
  S2DW_HIP_CHECK(hipMemcpyAsync(flm, e->d_flm, sizeof(double2) * (size_t)B * B, hipMemcpyDeviceToHost, p.stream));
  S2DW_HIP_CHECK(hipStreamSynchronize(p.stream));
  return kErrNone;
}

int s2dw_hip_synthesis_wav2flm(double *flm, const double *wav, const double *scoeff, const double *K_gamma,
                               const double *Phi2, const double *Slm, int J, int B, int N, int bl_scoeff,
                               double alpha) {
  if (!wav) return kErrArgInvalid;
  return synthesis_common(flm, nullptr, wav, scoeff, K_gamma, Phi2, Slm, J, B, N, bl_scoeff, alpha);
}

int s2dw_hip_synthesis_wav2flm_dynamic(double *flm, const double *const *wavdyn, const double *scoeff,
                                       const double *K_gamma, const double *Phi2, const double *Slm, int J,
                                       int B, int N, int bl_scoeff, double alpha) {
  if (!wavdyn) return kErrArgInvalid;
  for (int j = 0; j <= J; ++j)
    if (!wavdyn[j]) return kErrArgInvalid;
  return synthesis_common(flm, wavdyn, nullptr, scoeff, K_gamma, Phi2, Slm, J, B, N, bl_scoeff, alpha);
}

int s2dw_hip_plan_create(s2dw_hip_plan_t *plan, int B, int N, int J, double alpha, const double *K_gamma,
                         const double *Phi2, const double *Slm, int precision, void *stream) {
  if (!plan || !K_gamma || !Phi2 || !Slm) return kErrArgInvalid;
  const int prec = precision & ~S2DW_ROUNDING_REFERENCE;
  if (prec != S2DW_PREC_FP64 && prec != S2DW_PREC_SPLIT) return kErrArgInvalid;
  auto *h = new (std::nothrow) s2dw_hip_plan;
  if (!h) return kErrMemAlloc;
  const int e = plan_build(&h->p, B, N, J, alpha, K_gamma, Phi2, Slm, precision, (hipStream_t)stream, false);
  if (e) {
    plan_free(&h->p);
    delete h;
    *plan = nullptr;
    return e;
  }
  *plan = h;
  return kErrNone;
}

int s2dw_hip_plan_destroy(s2dw_hip_plan_t plan) {
  if (!plan) return kErrArgInvalid;
  plan_free(&plan->p);
  delete plan;
  return kErrNone;
}

int s2dw_hip_plan_sizes(s2dw_hip_plan_t plan, long long *wav_doubles, int *bl_scoeff, long long *wav_offsets,
                        int *bl_hi, int *bl_lo) {
  if (!plan) return kErrArgInvalid;
  const Plan &p = plan->p;
  if (wav_doubles) *wav_doubles = p.wav_doubles;
  if (bl_scoeff) *bl_scoeff = p.bl_scoeff;
  for (int j = 0; j <= p.J; ++j) {
    if (wav_offsets) wav_offsets[j] = p.sc[j].wav_off;
    if (bl_hi) bl_hi[j] = p.sc[j].hi;
    if (bl_lo) bl_lo[j] = p.sc[j].lo;
  }
  return kErrNone;
}

long long s2dw_hip_plan_device_bytes(s2dw_hip_plan_t plan) { return plan ? plan->p.device_bytes : -1; }

int s2dw_hip_analysis_device(s2dw_hip_plan_t plan, double *d_wav, double *d_scoeff, const double *d_flm) {
  if (!plan || !d_wav || !d_scoeff || !d_flm) return kErrArgInvalid;
  return run_graphed(&plan->p, 0, d_wav, d_scoeff, d_flm);
}

int s2dw_hip_synthesis_device(s2dw_hip_plan_t plan, double *d_flm, const double *d_wav, const double *d_scoeff) {
  if (!plan || !d_wav || !d_scoeff || !d_flm) return kErrArgInvalid;
  return run_graphed(&plan->p, 1, d_flm, const_cast<double *>(d_wav), d_scoeff);
}

int s2dw_hip_plan_sync(s2dw_hip_plan_t plan) {
  if (!plan) return kErrArgInvalid;
  S2DW_HIP_CHECK(hipStreamSynchronize(plan->p.stream));
  return kErrNone;
}

void *s2dw_hip_plan_stream(s2dw_hip_plan_t plan) { return plan ? (void *)plan->p.stream : nullptr; }

int s2dw_hip_wigner_plane(int l, double *out, double *out2) { return wigner_plane(l, out, out2); }

int s2dw_hip_plan_wigner_form(s2dw_hip_plan_t plan) {
  if (!plan) return -1;
  return plan_form(plan->p);
}

int s2dw_hip_plan_wigner_table(s2dw_hip_plan_t plan, int l, double *out) {
  if (!plan || !out) return kErrArgInvalid;
  const Plan &p = plan->p;
  if (!p.d_Sa || l < 0 || l >= p.B) {
    set_error("the plan has no Wigner table, or l=%d is outside [0, %d)", l, p.B);
    return kErrArgInvalid;
  }
  const long long off = (long long)l * (l + 1) * (2LL * l + 1) / 6;
  S2DW_HIP_CHECK(hipStreamSynchronize(p.stream));
  S2DW_HIP_CHECK(hipMemcpy(out, p.d_Sa + off, sizeof(double) * (size_t)(l + 1) * (l + 1), hipMemcpyDeviceToHost));
  return kErrNone;
}

int s2dw_hip_plan_set_timing(s2dw_hip_plan_t plan, int enable) {
  if (!plan) return kErrArgInvalid;
  plan->p.timing = enable != 0;
  return kErrNone;
}

int s2dw_hip_plan_autotune(s2dw_hip_plan_t plan, int reps, char *report, int report_len) {
  if (!plan) return kErrArgInvalid;
  std::string r;
  const int e = plan_autotune(&plan->p, reps, &r);
  if (report && report_len > 0) {
    std::strncpy(report, r.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = '\0';
  }
  return e;
}

int s2dw_hip_plan_set_graphs(s2dw_hip_plan_t plan, int enable) {
  if (!plan) return kErrArgInvalid;
  plan->p.graphs = enable != 0;
  if (!plan->p.graphs) plan_drop_graphs(&plan->p);
  return kErrNone;
}

int s2dw_hip_plan_stage_times(s2dw_hip_plan_t plan, double *ms, int *count) {
  if (!plan || !ms || !count) return kErrArgInvalid;
  return stage_times(&plan->p, ms, count);
}

// ---------------------------------------------------------------------------
// Moments of the wavelet coefficients (s2dw_stat_moments,
// s2dw_stat_mod.f90:81-275).
// ---------------------------------------------------------------------------

int s2dw_hip_stat_moments_device(s2dw_hip_plan_t plan, double *d_wav, const double *d_mask, double *d_out) {
  if (!plan || !d_wav || !d_out) return kErrArgInvalid;
  if (((uintptr_t)d_wav & 15) || ((uintptr_t)d_mask & 15)) {
    set_error("s2dw_hip_stat_moments_device: d_wav and d_mask must be 16-byte aligned");
    return kErrArgInvalid;
  }
  Plan &p = plan->p;
  const int n = p.J + 1;
  std::vector<long long> off(n), cnt(n);
  for (int j = 0; j < n; ++j) {
    off[j] = p.sc[j].wav_off;
    cnt[j] = (long long)(2 * p.sc[j].hi - 1) * 2 * p.sc[j].hi * p.N;
  }
  return moments_run(p.stream, off.data(), cnt.data(), n, d_wav, d_mask, p.d_mom, d_out);
}

int s2dw_hip_analysis_moments_device(s2dw_hip_plan_t plan, double *d_out, double *d_scoeff, const double *d_flm) {
  if (!plan || !d_out || !d_scoeff || !d_flm) return kErrArgInvalid;
  if (!plan->p.momf_ok) {
    set_error("s2dw_hip_analysis_moments_device: this plan runs a scale outside the grouped launches; use "
              "s2dw_hip_analysis_device and s2dw_hip_stat_moments_device");
    return kErrArgInvalid;
  }
  return run_graphed(&plan->p, 2, d_out, d_scoeff, d_flm);
}

int s2dw_hip_plan_moments_fusable(s2dw_hip_plan_t plan) { return plan && plan->p.momf_ok ? 1 : 0; }

int s2dw_hip_stat_moments_device_noalloc(double *d_wav, const double *d_mask, int J, int B, int N, double alpha,
                                         double *d_out, void *stream) {
  return moments_noalloc(d_wav, d_mask, J, B, N, alpha, d_out, static_cast<hipStream_t>(stream));
}

namespace {
struct DevBufs {
  std::vector<void *> v;
  hipStream_t st = nullptr;
  ~DevBufs() {
    if (st) (void)hipStreamSynchronize(st);
    for (void *q : v) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  }
  int get(double **out, size_t doubles) {
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, sizeof(double) * (doubles > 0 ? doubles : 1));
    if (e != hipSuccess) return hip_error(e, "hipMalloc");
    v.push_back(q);
    *out = static_cast<double *>(q);
    return kErrNone;
  }
};
}  // namespace

int s2dw_hip_stat_moments(double *const *wavdyn, int J, int B, int N, double alpha, double *mean, double *var,
                          double *skew, double *kur, const double *const *mask, double *five, double *six) {
  if (!wavdyn || !mean || !var || !skew || !kur || J < 0 || B < 1 || N < 1 || !(alpha > 1.0))
    return kErrArgInvalid;
  const int n = J + 1;
  for (int j = 0; j < n; ++j)
    if (!wavdyn[j] || (mask && !mask[j])) return kErrArgInvalid;
  std::vector<long long> off(n), cnt(n);
  long long total = 0;
  for (int j = 0; j < n; ++j) {
    int hi = 0, lo = 0;
    band_limits(j, B, alpha, &hi, &lo);
    cnt[j] = (long long)(2 * hi - 1) * 2 * hi * N;
    off[j] = total;
    total += cnt[j];
  }
  DevBufs d;
  S2DW_HIP_CHECK(hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking));
  double *d_w = nullptr, *d_m = nullptr, *d_part = nullptr, *d_out = nullptr;
  int err;
  if ((err = d.get(&d_w, (size_t)total))) return err;
  if (mask && (err = d.get(&d_m, (size_t)total))) return err;
  if ((err = d.get(&d_part, (size_t)moments_partial_doubles(cnt.data(), n)))) return err;
  if ((err = d.get(&d_out, 8 * (size_t)n))) return err;
  std::vector<double> out(8 * (size_t)n);
  {
    std::lock_guard<std::mutex> lk(g_mu);  // the staging buffers
    for (int j = 0; j < n; ++j) {
      if ((err = copy_h2d(d_w + off[j], wavdyn[j], sizeof(double) * cnt[j], d.st))) return err;
      if (mask && (err = copy_h2d(d_m + off[j], mask[j], sizeof(double) * cnt[j], d.st))) return err;
    }
    if ((err = moments_run(d.st, off.data(), cnt.data(), n, d_w, d_m, d_part, d_out))) return err;
    S2DW_HIP_CHECK(hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, d.st));
    // with a mask the coefficients are returned masked (s2dw_stat_mod.f90:136-140)
    if (mask)
      for (int j = 0; j < n; ++j)
        if ((err = copy_d2h(wavdyn[j], d_w + off[j], sizeof(double) * cnt[j], d.st))) return err;
    S2DW_HIP_CHECK(hipStreamSynchronize(d.st));
  }
  int warn = -1;
  for (int j = 0; j < n; ++j) {
    if (out[7 * (size_t)n + j] == 2) {
      // the scales before it are filled; its variance is NaN (the Fortran
      // module warns for the earlier zero-variance scales, then halts here)
      var[j] = std::numeric_limits<double>::quiet_NaN();
      set_error("s2dw_stat_moments: scale %d has %.0f samples; at least two are required", j, out[6 * (size_t)n + j]);
      return kErrSizeInvalid;
    }
    if (out[7 * (size_t)n + j] == 1 && warn < 0) warn = j;
    mean[j] = out[j];
    var[j] = out[(size_t)n + j];
    skew[j] = out[2 * (size_t)n + j];
    kur[j] = out[3 * (size_t)n + j];
    if (five) five[j] = out[4 * (size_t)n + j];
    if (six) six[j] = out[5 * (size_t)n + j];
  }
  if (warn >= 0) {
    set_error("s2dw_stat_moments: skewness/kurtosis & higher undefined when variance zero (scale %d)", warn);
    return kErrArthWarning;
  }
  return kErrNone;
}

}  // extern "C"
