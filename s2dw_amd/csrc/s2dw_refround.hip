// s2dw_refround.hip -- the analysis beta sum with the reference's own
// rounding (reference-rounding form of the plain transform).
//
// The reference forms Tmbg by a direct sum over m' (s2dw_core_mod.f90:707-743,
// 471-506 in the static routine):
//   Tb(m, b, g) = sum_{m'=0}^{hi-1} Tg(m, m', g) * eta(m', b),
//   eta = 1 (m' = 0), 2 cos(m' beta_b) (m + N - 1 even), 2 i sin(m' beta_b) (odd),
//   beta_b = PI (2b + 1) / (4 hi),
// with cos and sin evaluated at the double-precision product m' * beta_b.  The
// rounding of that product (up to ~1e-13 in the argument at m' ~ 256) makes
// the reference's basis differ from the exact DCT/DST the FFT path computes;
// at L = 256 with white input that moves the wavelet coefficients by ~2.8e-12.
// This form reproduces the sum with the reference's eta values (computed on
// the host with the same libm calls) as two FP64 GEMMs per scale on the
// device (rocBLAS, strided-batched over g), one with the cosine table for all
// m and one with the sine table, and a kernel that keeps the cosine result for
// the even class and forms Tg(m, 0) + i * (sine result) for the odd class.
// Layouts (one scale, row-major): Tg[g][m'][m], Tb[g][b][m] (complex),
// Ec[b][m'], Es[b][m'] (real, Ec[b][0] = 1, Es[b][0] = 0).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "s2dw_internal.h"

namespace s2dw {

namespace {

__global__ void k_refbeta_combine(int hi, int N, const double2 *__restrict__ Tg, const double2 *__restrict__ Cs,
                                  double2 *__restrict__ Tb) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per_g = 2LL * hi * hi;
  if (id >= per_g * N) return;
  const int g = (int)(id / per_g);
  const long long r = id - g * per_g;
  const int m = (int)(r % hi);
  if (((m + N - 1) & 1) == 0) return;  // even class: the cosine GEMM wrote Tb already
  const double2 t0 = Tg[(long long)g * hi * hi + m];  // m' = 0
  const double2 s = Cs[id];
  Tb[id] = make_double2(t0.x - s.y, t0.y + s.x);
}

int blas_error(rocblas_status s, const char *what) {
  set_error("rocBLAS status %d in %s", (int)s, what);
  return s == rocblas_status_memory_error ? kErrMemAlloc : kErrInitFail;
}

}  // namespace

int refbeta_build(Plan *p) {
  const double PI = 3.141592653589793238462643383279502884197;
  const int Jp1 = p->J + 1;
  p->eta_off.assign(Jp1, -1);
  long long tot = 0, tmp = 0;
  for (int j = 0; j < Jp1; ++j) {
    const Scale &s = p->sc[j];
    if (s.lo + 1 > s.hi - 1) continue;
    p->eta_off[j] = tot;
    tot += 2LL * (2LL * s.hi * s.hi);  // Ec and Es, 2hi x hi each
    tmp = std::max(tmp, (long long)p->N * 2 * s.hi * s.hi);
  }
  std::vector<double> eta((size_t)std::max(1LL, tot));
  for (int j = 0; j < Jp1; ++j) {
    if (p->eta_off[j] < 0) continue;
    const int hi = p->sc[j].hi;
    double *Ec = eta.data() + p->eta_off[j], *Es = Ec + 2LL * hi * hi;
    for (int bb = 0; bb <= 2 * hi - 1; ++bb) {
      const double beta_b = PI * (2.0 * bb + 1.0) / (4.0 * hi);  // s2dw_core_mod.f90:720
      Ec[(long long)bb * hi] = 1.0;
      Es[(long long)bb * hi] = 0.0;
      for (int mm = 1; mm <= hi - 1; ++mm) {
        Ec[(long long)bb * hi + mm] = 2.0 * std::cos(mm * beta_b);
        Es[(long long)bb * hi + mm] = 2.0 * std::sin(mm * beta_b);
      }
    }
  }
  hipError_t e;
  if ((e = hipMalloc(&p->d_eta, eta.size() * sizeof(double))) != hipSuccess) return hip_error(e, "hipMalloc");
  p->allocs.push_back(p->d_eta);
  p->device_bytes += (long long)(eta.size() * sizeof(double));
  if ((e = hipMalloc(&p->d_cs, (size_t)std::max(1LL, tmp) * sizeof(double2))) != hipSuccess)
    return hip_error(e, "hipMalloc");
  p->allocs.push_back(p->d_cs);
  p->device_bytes += (long long)(std::max(1LL, tmp) * sizeof(double2));
  if ((e = hipMemcpyAsync(p->d_eta, eta.data(), eta.size() * sizeof(double), hipMemcpyHostToDevice, p->stream)) !=
      hipSuccess)
    return hip_error(e, "eta upload");
  rocblas_handle h = nullptr;
  if (rocblas_status st = rocblas_create_handle(&h); st != rocblas_status_success) return blas_error(st, "create");
  p->blas = h;
  // one call at plan time, so that rocBLAS loads its kernels and sets up its
  // workspace here rather than inside a transform
  if (int err = refbeta_run(p, p->stream)) return err;
  if ((e = hipStreamSynchronize(p->stream)) != hipSuccess) return hip_error(e, "eta upload");
  return kErrNone;
}

void refbeta_free(Plan *p) {
  if (p->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(p->blas));
  p->blas = nullptr;
}

// Tb_j for every non-empty scale, on stream st.
int refbeta_run(Plan *p, hipStream_t st) {
  rocblas_handle h = static_cast<rocblas_handle>(p->blas);
  if (rocblas_status s = rocblas_set_stream(h, st); s != rocblas_status_success) return blas_error(s, "set_stream");
  const double one = 1.0, zero = 0.0;
  const int N = p->N;
  for (int j = 0; j <= p->J; ++j) {
    if (p->eta_off[j] < 0) continue;
    const Scale &s = p->sc[j];
    const int hi = s.hi;
    const double *Ec = p->d_eta + p->eta_off[j], *Es = Ec + 2LL * hi * hi;
    const double *X = reinterpret_cast<const double *>(p->d_tg + s.tg_off);
    double *Tb = reinterpret_cast<double *>(p->d_tb + s.tb_off);
    double *Cs = reinterpret_cast<double *>(p->d_cs);
    // column-major view: C'(2hi x 2hi) = X'(2hi x hi) * E'(hi x 2hi), X' = Tg_g
    // as [m'][2m + re/im], E' = E as [b][m'], C' = out as [b][2m + re/im]
    const rocblas_int M = 2 * hi, Nn = 2 * hi, K = hi;
    const rocblas_stride sx = 2LL * hi * hi, sc = 4LL * hi * hi;
    for (int which = 0; which < 2; ++which) {
      rocblas_status r = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, M, Nn, K,
                                                       &one, X, 2 * hi, sx, which ? Es : Ec, hi, 0, &zero,
                                                       which ? Cs : Tb, 2 * hi, sc, N);
      if (r != rocblas_status_success) return blas_error(r, "dgemm_strided_batched");
    }
    const long long n = (long long)N * 2 * hi * hi;
    hipLaunchKernelGGL(k_refbeta_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hi, N,
                       p->d_tg + s.tg_off, p->d_cs, p->d_tb + s.tb_off);
    S2DW_HIP_CHECK(hipGetLastError());
  }
  return kErrNone;
}

}  // namespace s2dw
