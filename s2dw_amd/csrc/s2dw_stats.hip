// s2dw_stats.hip -- moments of the wavelet coefficients of each scale on the
// device: mean, variance, skewness, excess kurtosis and the fifth and excess
// sixth standardised moments of s2dw_stat_moments (s2dw_stat_mod.f90:81-275),
// with its optional mask.
//
// One pass over HBM.  Each workgroup reduces an 8192-value chunk of one scale
// to a record (count, mean, central sums M_2..M_6, sum of mask values) by an
// exact two-pass over the values its lanes hold in registers: a workgroup sum
// gives the chunk mean, a second workgroup sum the powers of the deviations
// from it.  A second kernel combines the records of each scale the same way:
// the scale mean from the records' counts and means, then each record's
// central sums moved to that mean with the binomial theorem and summed (the
// pairwise update of P. Pebay, "Formulas for robust, one-pass parallel
// computation of covariances and arbitrary-order statistical moments",
// Sandia report SAND2008-6212 (2008), written for many parts at once), in a
// fixed order, so the result does not depend on scheduling.  It then applies
// the reference's finishing formulas: nj = nint(sum of mask),
// mean = (sum over included values) / nj, central sums re-centred on that
// mean, var = (sum s^2 - (sum s)^2 / nj) / (nj - 1), and zero higher moments
// when var < TOL_ZERO (s2dw_stat_mod.f90:143-150, 229-268).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "s2dw_internal.h"
#include "s2dw_moments.h"

namespace s2dw {
namespace {

constexpr int kMomThreads = 256;
constexpr int kMomPairs = 16;  // double2 loads per lane: 32 values
constexpr long long kMomChunk = (long long)kMomThreads * kMomPairs * 2;
constexpr int kMomBatch = 32;  // scales per launch
constexpr double kTolZero = kMomTolZero;

struct MomTab {
  long long off[kMomBatch];    // scale offset in the packed coefficients (doubles)
  long long count[kMomBatch];  // values of the scale
  int blk0[kMomBatch + 1];     // first workgroup (and partial) of each scale
  int nscale;
  int masked;
};

template <bool MASKED>
__global__ __launch_bounds__(kMomThreads) void k_mom_partial(const MomTab t, double *__restrict__ wav,
                                                             const double *__restrict__ mask,
                                                             double *__restrict__ part, int part_base) {
  const int b = blockIdx.x;
  int s = 0;
  while (s + 1 < t.nscale && b >= t.blk0[s + 1]) ++s;
  const long long cnt = t.count[s];
  const long long base = (long long)(b - t.blk0[s]) * kMomChunk;
  double *w = wav + t.off[s];
  const double *mk = MASKED ? mask + t.off[s] : nullptr;
  double2 x[kMomPairs], mv[kMomPairs];
#pragma unroll
  for (int e = 0; e < kMomPairs; ++e) {
    // pairs never straddle the end: (2hi-1) 2hi N is even
    const long long i = base + 2LL * (e * kMomThreads + threadIdx.x);
    x[e] = make_double2(0.0, 0.0);
    mv[e] = make_double2(0.0, 0.0);
    if (i < cnt) {
      x[e] = *reinterpret_cast<const double2 *>(w + i);
      mv[e] = MASKED ? *reinterpret_cast<const double2 *>(mk + i) : make_double2(1.0, 1.0);
    }
  }
  // the masked coefficients replace the input (s2dw_stat_mod.f90:136-140);
  // values outside the mask (or past the end) take part with weight 0
  double v[2 * kMomPairs], wt[2 * kMomPairs];
  double ms = 0;
#pragma unroll
  for (int e = 0; e < kMomPairs; ++e) {
    const long long i = base + 2LL * (e * kMomThreads + threadIdx.x);
    const bool in = i < cnt;
    const double2 y = make_double2(x[e].x * mv[e].x, x[e].y * mv[e].y);
    if (MASKED && in) *reinterpret_cast<double2 *>(w + i) = y;
    v[2 * e] = y.x;
    v[2 * e + 1] = y.y;
    wt[2 * e] = in && fabs(mv[e].x - 1.0) < kTolZero ? 1.0 : 0.0;
    wt[2 * e + 1] = in && fabs(mv[e].y - 1.0) < kTolZero ? 1.0 : 0.0;
    if (in) ms += mv[e].x + mv[e].y;
  }
  mom_record<2 * kMomPairs, kMomThreads>(v, wt, ms, part + 8LL * (part_base + b));
}

// One workgroup per scale: combine its records, then the reference's
// formulas.  out[k * nout + jj], k = 0 mean, 1 var, 2 skew, 3 kur, 4 five,
// 5 six, 6 nj, 7 flag (0 ok, 1 var < TOL_ZERO: higher moments set to zero,
// 2 nj < 2).
__global__ __launch_bounds__(kMomThreads) void k_mom_final(const MomTab t, const double *__restrict__ part,
                                                           int part_base, double *__restrict__ out, int nout,
                                                           int jj0) {
  __shared__ double sh3[kMomThreads / 64][3];
  __shared__ double sh5[kMomThreads / 64][5];
  const int s = blockIdx.x;
  const int p0 = t.blk0[s], p1 = t.blk0[s + 1];
  // records (n_i, mu_i, M_{2..6,i}, ms_i): the scale's count and mean from
  // sum n_i and sum n_i mu_i, then each record's central sums moved to that
  // mean, T_p = sum_k C(p,k) (mu_i - mu)^(p-k) M_{k,i} with M_0 = n_i, M_1 = 0,
  // and summed
  double c3[3] = {0, 0, 0};
  for (int p = p0 + (int)threadIdx.x; p < p1; p += kMomThreads) {
    const double *o = part + 8LL * (part_base + p);
    c3[0] += o[0];
    c3[1] += o[0] * o[1];
    c3[2] += o[7];
  }
  block_sum<3, kMomThreads>(c3, sh3);
  const double c = c3[0], mu = c > 0 ? c3[1] / c : 0.0;
  double T[5] = {0, 0, 0, 0, 0};  // T_2..T_6 about mu
  for (int p = p0 + (int)threadIdx.x; p < p1; p += kMomThreads) {
    const double *o = part + 8LL * (part_base + p);
    const double n = o[0], d = o[1] - mu, M2 = o[2], M3 = o[3], M4 = o[4], M5 = o[5];
    const double d2 = d * d, d3 = d2 * d, d4 = d2 * d2;
    T[0] += M2 + n * d2;
    T[1] += M3 + 3 * d * M2 + n * d3;
    T[2] += M4 + 4 * d * M3 + 6 * d2 * M2 + n * d4;
    T[3] += M5 + 5 * d * M4 + 10 * d2 * M3 + 10 * d3 * M2 + n * d4 * d;
    T[4] += o[6] + 6 * d * M5 + 15 * d2 * M4 + 20 * d3 * M3 + 15 * d4 * M2 + n * d3 * d3;
  }
  block_sum<5, kMomThreads>(T, sh5);
  if (threadIdx.x != 0) return;
  // nj = nint(sum(mask)) with a mask, the number of values without
  // (s2dw_stat_mod.f90:143-148); Fortran nint rounds halves away from zero
  const double nj = t.masked ? round(c3[2]) : (double)t.count[s];
  double r[8];
  if (nj < 2) {
    for (int k = 0; k < 6; ++k) r[k] = NAN;
    r[6] = nj;
    r[7] = 2;
  } else {
    const double mean = c * mu / nj;
    // sums of s = x - mean over the included values, from the central sums
    // about mu: S_p = sum_k C(p,k) (mu - mean)^(p-k) T_k, T_0 = c, T_1 = 0
    const double d = mu - mean;
    const double Tk[7] = {c, 0, T[0], T[1], T[2], T[3], T[4]};
    double S[7];
    constexpr double Cb[7][7] = {{1},
                                 {1, 1},
                                 {1, 2, 1},
                                 {1, 3, 3, 1},
                                 {1, 4, 6, 4, 1},
                                 {1, 5, 10, 10, 5, 1},
                                 {1, 6, 15, 20, 15, 6, 1}};
    for (int p = 1; p <= 6; ++p) {
      double acc = 0, dp = 1;  // d^(p-k), k from p down to 0
      for (int k = p; k >= 0; --k) {
        acc += Cb[p][k] * dp * Tk[k];
        dp *= d;
      }
      S[p] = acc;
    }
    const double ep = S[1];
    const double var = (S[2] - ep * ep / nj) / (nj - 1);
    r[0] = mean;
    r[1] = var;
    r[6] = nj;
    if (var < kTolZero) {
      r[2] = r[3] = r[4] = r[5] = 0;
      r[7] = 1;
    } else {
      const double sd = sqrt(var);
      r[2] = S[3] / (nj * sd * sd * sd);
      r[3] = S[4] / (nj * var * var) - 3.0;
      r[4] = S[5] / (nj * sd * sd * sd * sd * sd);
      r[5] = S[6] / (nj * var * var * var) - 15.0;
      r[7] = 0;
    }
  }
  for (int k = 0; k < 8; ++k) out[(long long)k * nout + jj0 + s] = r[k];
}

}  // namespace

long long moments_partial_doubles(const long long *count, int nscales) {
  long long blocks = 0;
  for (int s = 0; s < nscales; ++s) blocks += (count[s] + kMomChunk - 1) / kMomChunk;
  return 8 * blocks;
}

int moments_run(hipStream_t st, const long long *off, const long long *count, int nscales, double *d_wav,
                const double *d_mask, double *d_part, double *d_out) {
  if (nscales <= 0 || !d_wav || !d_part || !d_out) return kErrArgInvalid;
  int part_base = 0;
  for (int s0 = 0; s0 < nscales; s0 += kMomBatch) {
    MomTab t{};
    t.nscale = nscales - s0 < kMomBatch ? nscales - s0 : kMomBatch;
    t.masked = d_mask != nullptr;
    int nb = 0;
    for (int s = 0; s < t.nscale; ++s) {
      if (count[s0 + s] < 0 || count[s0 + s] % 2) return kErrArgInvalid;
      t.off[s] = off[s0 + s];
      t.count[s] = count[s0 + s];
      t.blk0[s] = nb;
      nb += (int)((count[s0 + s] + kMomChunk - 1) / kMomChunk);
    }
    t.blk0[t.nscale] = nb;
    if (nb > 0) {
      if (d_mask)
        hipLaunchKernelGGL(k_mom_partial<true>, dim3(nb), dim3(kMomThreads), 0, st, t, d_wav, d_mask, d_part,
                           part_base);
      else
        hipLaunchKernelGGL(k_mom_partial<false>, dim3(nb), dim3(kMomThreads), 0, st, t, d_wav, d_mask, d_part,
                           part_base);
      S2DW_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mom_final, dim3(t.nscale), dim3(kMomThreads), 0, st, t, d_part, part_base, d_out,
                       nscales, s0);
    S2DW_HIP_CHECK(hipGetLastError());
    part_base += nb;
  }
  return kErrNone;
}

int moments_finish(hipStream_t st, const long long *count, const long long *rec0, int nscales, const double *d_part,
                   double *d_out) {
  if (nscales <= 0 || !d_part || !d_out) return kErrArgInvalid;
  for (int s0 = 0; s0 < nscales; s0 += kMomBatch) {
    MomTab t{};
    t.nscale = nscales - s0 < kMomBatch ? nscales - s0 : kMomBatch;
    t.masked = 0;
    for (int s = 0; s <= t.nscale; ++s) {
      if (s < t.nscale) t.count[s] = count[s0 + s];
      t.blk0[s] = (int)(rec0[s0 + s] - rec0[s0]);
    }
    hipLaunchKernelGGL(k_mom_final, dim3(t.nscale), dim3(kMomThreads), 0, st, t, d_part, (int)rec0[s0], d_out,
                       nscales, s0);
    S2DW_HIP_CHECK(hipGetLastError());
  }
  return kErrNone;
}

// ---------------------------------------------------------------------------
// Moments without allocation (s2dw_hip_stat_moments_device_noalloc): the
// chunk records live in a static device array, `scratch`, of kScratchDoubles
// doubles (1 MiB) that every GPU the library runs on holds until the process
// ends.  Host code never passes the array's name to a kernel: on the first
// call on a device (that device current) hipGetSymbolAddress gives the
// array's address on that device, kept per device; calls use that address.
// The records are written before they are read, so nothing is cleared; calls
// that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
namespace {
constexpr long long kScratchDoubles = 131072;
constexpr int kScratchDevices = 64;
constexpr int kScratchScales = 32;
static __device__ __align__(256) double scratch[kScratchDoubles];
double *g_scratch[kScratchDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

int moments_noalloc(double *d_wav, const double *d_mask, int J, int B, int N, double alpha, double *d_out,
                    hipStream_t st) {
  if (!d_wav || !d_out || J < 0 || B < 1 || N < 1 || !(alpha > 1.0)) return kErrArgInvalid;
  if (J + 1 > kScratchScales) {
    set_error("s2dw_hip_stat_moments_device_noalloc: J=%d exceeds the %d scales it handles", J, kScratchScales - 1);
    return kErrSizeInvalid;
  }
  if (((uintptr_t)d_wav & 15) || ((uintptr_t)d_mask & 15)) {
    set_error("s2dw_hip_stat_moments_device_noalloc: d_wav and d_mask must be 16-byte aligned");
    return kErrArgInvalid;
  }
  const int n = J + 1;
  long long off[kScratchScales], cnt[kScratchScales], total = 0;
  for (int j = 0; j < n; ++j) {
    int hi = 0, lo = 0;
    band_limits(j, B, alpha, &hi, &lo);
    cnt[j] = (long long)(2 * hi - 1) * 2 * hi * N;
    off[j] = total;
    total += cnt[j];
  }
  const long long need = moments_partial_doubles(cnt, n);
  if (need > kScratchDoubles) {
    set_error("s2dw_hip_stat_moments_device_noalloc: B=%d N=%d J=%d needs %lld doubles of scratch, more than the "
              "%lld it holds",
              B, N, J, need, kScratchDoubles);
    return kErrSizeInvalid;
  }
  int dev = 0;
  S2DW_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return kErrInitFail;
  double *ptr = nullptr;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
      void *q = nullptr;
      S2DW_HIP_CHECK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
      g_scratch[dev] = static_cast<double *>(q);
    }
    ptr = g_scratch[dev];
  }
  return moments_run(st, off, cnt, n, d_wav, d_mask, ptr, d_out);
}

}  // namespace s2dw
