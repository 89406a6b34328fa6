// s2dw_moments.h -- workgroup reduction of coefficient values to a moments
// record (count, mean, M_2..M_6 about that mean, sum of mask values), used by
// the moments kernels (s2dw_stats.hip) and by the analysis epilogue that
// reduces instead of storing (s2dw_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace s2dw {

// TOL_ZERO of s2dw_types_mod.f90:51, a default-real literal
constexpr double kMomTolZero = (double)1e-12f;

// Sum of K values over a workgroup of NT lanes, returned to every lane in the
// same bits (xor butterflies within each wave, then the waves in order).
template <int K, int NT>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*sh)[K]) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], o);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[wave][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = sh[0][k];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) t += sh[w][k];
    v[k] = t;
  }
  __syncthreads();
}

// Exact two-pass over the values v[e] (weight wt[e] = 1 to include, 0 to
// skip) held by the lanes of the workgroup: a workgroup sum gives the mean,
// a second one the powers 2..6 of the deviations from it.  Lane 0 writes
// rec[0..7] = count, mean, M_2, ..., M_6, msum (msum: the workgroup sum of
// the lanes' ms).  Every lane of the workgroup must call it.
template <int E, int NT>
__device__ __forceinline__ void mom_record(const double (&v)[E], const double (&wt)[E], double ms,
                                           double *__restrict__ rec) {
  __shared__ double sh3[NT / 64][3];
  __shared__ double sh5[NT / 64][5];
  double c3[3] = {0, 0, ms};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    c3[0] += wt[e];
    c3[1] += wt[e] * v[e];
  }
  block_sum<3, NT>(c3, sh3);
  const double mu = c3[0] > 0 ? c3[1] / c3[0] : 0.0;
  double p5[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const double s1 = wt[e] * (v[e] - mu), s2 = s1 * s1, s3 = s2 * s1;
    p5[0] += s2;
    p5[1] += s3;
    p5[2] += s2 * s2;
    p5[3] += s3 * s2;
    p5[4] += s3 * s3;
  }
  block_sum<5, NT>(p5, sh5);
  if (threadIdx.x == 0) {
    rec[0] = c3[0];
    rec[1] = mu;
#pragma unroll
    for (int k = 0; k < 5; ++k) rec[2 + k] = p5[k];
    rec[7] = c3[2];
  }
}

}  // namespace s2dw
