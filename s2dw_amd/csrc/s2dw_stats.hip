// s2dw_stats.hip -- moments of the wavelet coefficients of each scale on the
// device: mean, variance, skewness, excess kurtosis and the fifth and excess
// sixth standardised moments of s2dw_stat_moments (s2dw_stat_mod.f90:81-275),
// with its optional mask.
//
// One pass over HBM.  Each workgroup reduces an 8192-value chunk of one scale
// to (count, mean, central sums M_2..M_6, sum of mask values) by an exact
// two-pass over the values its lanes hold in registers: a workgroup sum gives
// the chunk mean, a second workgroup sum the powers of the deviations from
// it.  A second kernel merges the chunk results of each scale in a fixed
// order (the result does not depend on scheduling) with the update formula
// for central sums of P. Pebay, "Formulas for robust, one-pass parallel
// computation of covariances and arbitrary-order statistical moments",
// Sandia report SAND2008-6212 (2008), eq. 3.1, and applies the reference's
// finishing formulas: nj = nint(sum of mask),
// mean = (sum over included values) / nj, central sums re-centred on that
// mean, var = (sum s^2 - (sum s)^2 / nj) / (nj - 1), and zero higher moments
// when var < TOL_ZERO (s2dw_stat_mod.f90:143-150, 229-268).
#include <hip/hip_runtime.h>

#include <cmath>

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

// count, mean, M_2..M_6 about the mean, sum of mask values
struct Mom {
  double n, mu, m2, m3, m4, m5, m6, ms;
};

__device__ __forceinline__ Mom mom_zero() { return Mom{0, 0, 0, 0, 0, 0, 0, 0}; }

// a <- a merged with b (Pebay 2008, eq. 3.1, orders 2..6):
//   M_p = A_p + B_p + sum_{k=1}^{p-2} C(p,k) d^k [(-nb/n)^k A_{p-k} + (na/n)^k B_{p-k}]
//         + q [(d na/n)^{p-1} - (-d nb/n)^{p-1}],   d = mu_b - mu_a, q = na nb d / n.
__device__ __forceinline__ void mom_merge(Mom &a, const Mom &b) {
  if (b.n == 0) {
    a.ms += b.ms;
    return;
  }
  if (a.n == 0) {
    const double ms = a.ms + b.ms;
    a = b;
    a.ms = ms;
    return;
  }
  const double n = a.n + b.n, d = b.mu - a.mu, ra = a.n / n, rb = b.n / n;
  const double q = a.n * rb * d;
  const double A[7] = {0, 0, a.m2, a.m3, a.m4, a.m5, a.m6};
  const double Bm[7] = {0, 0, b.m2, b.m3, b.m4, b.m5, b.m6};
  double dk[6], pa[6], pb[6], ea[6], eb[6];  // d^k, (-rb)^k, ra^k, (ra d)^k, (-rb d)^k
  dk[0] = pa[0] = pb[0] = ea[0] = eb[0] = 1.0;
#pragma unroll
  for (int k = 1; k < 6; ++k) {
    dk[k] = dk[k - 1] * d;
    pa[k] = pa[k - 1] * -rb;
    pb[k] = pb[k - 1] * ra;
    ea[k] = ea[k - 1] * (ra * d);
    eb[k] = eb[k - 1] * (-rb * d);
  }
  constexpr double C[7][5] = {{0}, {0}, {0}, {0, 3}, {0, 4, 6}, {0, 5, 10, 10}, {0, 6, 15, 20, 15}};
  double M[7];
#pragma unroll
  for (int p = 2; p <= 6; ++p) {
    double s = A[p] + Bm[p];
#pragma unroll
    for (int k = 1; k <= p - 2; ++k) s += C[p][k] * dk[k] * (pa[k] * A[p - k] + pb[k] * Bm[p - k]);
    M[p] = s + q * (ea[p - 1] - eb[p - 1]);
  }
  a.n = n;
  a.mu += d * rb;
  a.m2 = M[2];
  a.m3 = M[3];
  a.m4 = M[4];
  a.m5 = M[5];
  a.m6 = M[6];
  a.ms += b.ms;
}

__device__ __forceinline__ Mom mom_shfl_xor(const Mom &a, int lane_mask) {
  Mom b;
  b.n = __shfl_xor(a.n, lane_mask);
  b.mu = __shfl_xor(a.mu, lane_mask);
  b.m2 = __shfl_xor(a.m2, lane_mask);
  b.m3 = __shfl_xor(a.m3, lane_mask);
  b.m4 = __shfl_xor(a.m4, lane_mask);
  b.m5 = __shfl_xor(a.m5, lane_mask);
  b.m6 = __shfl_xor(a.m6, lane_mask);
  b.ms = __shfl_xor(a.ms, lane_mask);
  return b;
}

// Merge the Mom of every lane of the workgroup; lane 0 of wave 0 returns
// the result.  The pairing is fixed (xor tree in the wave, waves in order).
__device__ __forceinline__ Mom mom_block_reduce(Mom m, Mom *sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const Mom o2 = mom_shfl_xor(m, o);
    // both partners compute the same merge in the same operand order
    if (lane & o) {
      Mom t = o2;
      mom_merge(t, m);
      m = t;
    } else {
      mom_merge(m, o2);
    }
  }
  if (lane == 0) sh[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = sh[0];
    for (int w = 1; w < kMomThreads / 64; ++w) mom_merge(m, sh[w]);
  }
  return m;
}

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

// One workgroup per scale: merge its partials, then the reference's formulas.
// out[k * nout + jj], k = 0 mean, 1 var, 2 skew, 3 kur, 4 five, 5 six, 6 nj,
// 7 flag (0 ok, 1 var < TOL_ZERO: higher moments set to zero, 2 nj < 2).
__global__ __launch_bounds__(kMomThreads) void k_mom_final(const MomTab t, const double *__restrict__ part,
                                                           int part_base, double *__restrict__ out, int nout,
                                                           int jj0) {
  __shared__ Mom sh[kMomThreads / 64];
  const int s = blockIdx.x;
  const int p0 = t.blk0[s], p1 = t.blk0[s + 1];
  Mom m = mom_zero();
  for (int p = p0 + (int)threadIdx.x; p < p1; p += kMomThreads) {
    const double *o = part + 8LL * (part_base + p);
    const Mom b = Mom{o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7]};
    mom_merge(m, b);
  }
  m = mom_block_reduce(m, sh);
  if (threadIdx.x != 0) return;
  const double c = m.n;
  // nj = nint(sum(mask)) with a mask, the number of values without
  // (s2dw_stat_mod.f90:143-148); Fortran nint rounds halves away from zero
  const double nj = t.masked ? round(m.ms) : (double)t.count[s];
  double r[8];
  if (nj < 2) {
    for (int k = 0; k < 6; ++k) r[k] = NAN;
    r[6] = nj;
    r[7] = 2;
  } else {
    const double mean = c * m.mu / nj;
    // sums of s = x - mean over the included values, from the central sums
    // about their own mean mu: T_p = sum_k C(p,k) (mu - mean)^(p-k) M_k,
    // M_0 = c, M_1 = 0
    const double d = m.mu - mean;
    const double Mk[7] = {c, 0, m.m2, m.m3, m.m4, m.m5, m.m6};
    double T[7];
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
        acc += Cb[p][k] * dp * Mk[k];
        dp *= d;
      }
      T[p] = acc;
    }
    const double ep = T[1];
    const double var = (T[2] - ep * ep / nj) / (nj - 1);
    r[0] = mean;
    r[1] = var;
    r[6] = nj;
    if (var < kTolZero) {
      r[2] = r[3] = r[4] = r[5] = 0;
      r[7] = 1;
    } else {
      const double sd = sqrt(var);
      r[2] = T[3] / (nj * sd * sd * sd);
      r[3] = T[4] / (nj * var * var) - 3.0;
      r[4] = T[5] / (nj * sd * sd * sd * sd * sd);
      r[5] = T[6] / (nj * var * var * var) - 15.0;
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

}  // namespace s2dw
