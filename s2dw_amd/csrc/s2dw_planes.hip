// s2dw_planes.hip -- Wigner planes Delta^l_{m'm}(pi/2) with the reference's
// own rounding, generated on the device once per plan.
//
// The reference builds each plane with s2dw_dl_beta_operator
// (src/mod/s2dw_dl_mod.f90:56-68): a three-term recursion along the second
// index of every row, started at 1 (_core, 92-238), with a rescaling by 1e-150
// whenever a value exceeds 1e150, then a row factor sign(i) *
// exp(log_first_row(i) - lrenorm(i)), log_first_row a running sum of
// logarithms over the rows, and the rest of the plane filled by symmetry
// (_fill, 258-304).  The row factors carry a relative rounding of up to ~1e-13
// and the recursions their own; at L = 256 with white input these move the
// wavelet coefficients by ~5e-12 from the same transform with exact Wigner
// values.  To land within 1e-12 of the reference's output, the plain form
// reproduces these values bit for bit:
//  - the host computes every input that needs a libm function exactly as the
//    reference does (cos, sin, tan, log, exp, sqrt and the divisions of the
//    coefficient arrays: cpi, cp2 per l; the row factors for 0..3 rescalings);
//  - the device runs the recursions with the same IEEE multiplications and
//    subtractions in the same order (no contraction into FMAs), one thread
//    per row, once to find the rescaling points and once to write.
//
// Only the rows the transform reads are built: for 0 <= m', m <= l the fill
// maps d(a, b) to the computed element (-min(a,b), -max(a,b)) of the first
// loop's rows (a <= 0), with the sign (-1)^(b-a) when b > a; the analysis
// also reads d(m', n) for the negative n of its columns (|n| < N), which come
// from the second loop's rows a = 1..N-1 (and row 0).
//
// Outputs (all indices zero-based):
//  Sa[off(l) + a*(l+1) + b] = d^l_{ab}, 0 <= a, b <= l, off(l) = sum_{l'<l} (l'+1)^2
//  thin[rowA(m',l)*nk + n]    = d^l_{m'n},  0 <= n < nk, n <= l
//  thinneg[rowA(m',l)*nk + n] = d^l_{m',-n}, 1 <= n < nk, n <= l
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "s2dw_internal.h"

namespace s2dw {

namespace {

__host__ __device__ __forceinline__ long long tri0p(long long l) { return l * (l + 1) / 2; }
__host__ __device__ __forceinline__ long long rowAp(long long mp, long long l, long long L) {
  return mp * L - mp * (mp - 1) / 2 + (l - mp);
}

constexpr int kMaxRescale = 3;  // rescalings per row (at most 2 occur for l < 1024)

struct RefPlaneArgs {
  int L, nk;
  double omc, c, s, si, big, bigi;
  const double *cpi, *cp2;  // [l(l-1)/2 + m - 1], m = 1..l
  const double *rn1;        // first-loop rows: [(tri0(l) + k) * 4 + e], row i = l + 1 - k, e rescalings
  const double *rn2;        // second-loop rows a = 1..nk-1: [(l * (nk - 1) + a - 1) * 4 + e]
  double *Sa, *thin, *thinneg;
  int *status;              // set to 1 when a row needs more than kMaxRescale rescalings
};

// The recursion of one row i of plane l (1-based columns j), as
// s2dw_dl_beta_operator_core runs it (s2dw_dl_mod.f90:157-206).  The column 2
// factor is `* si` in the first loop and `/ s` in the second (both exact at
// s = 1, kept as written).  emit(j, u, ev) receives the value u of column j
// as first computed and ev, the number of rescalings the row made before
// column j was computed; a rescaling at column p multiplies columns 1..p by
// bigi, so column j ends up rescaled (nev - ev) times, nev the row's total.
template <class Emit>
__device__ __forceinline__ int ref_row(const RefPlaneArgs &a, int l, int i, int ncol, bool first_loop, Emit emit) {
#pragma clang fp contract(off)
  const double *cpi = a.cpi + (long long)l * (l - 1) / 2 - 1;  // cpi[m], m >= 1
  const double *cp2 = a.cp2 + (long long)l * (l - 1) / 2 - 1;
  const double lp = (double)(l + 1), di = (double)i;
  int nev = 0;
  double vm1 = 1.0;  // column 1
  emit(1, vm1, 0);
  if (ncol < 2) return 0;
  double lam = first_loop ? (lp * a.omc - di + a.c) * a.si : (lp * a.omc - di + a.c) / a.s;
  double v = lam * vm1 * cpi[1];  // column 2
  emit(2, v, 0);
  for (int m = 2; m + 1 <= ncol; ++m) {
    lam = (lp * a.omc - di + (double)m * a.c) / a.s;
    const double vn = lam * cpi[m] * v - cp2[m] * vm1;
    emit(m + 1, vn, nev);
    if (vn > a.big) {  // rescale the row so far, the new value included
      ++nev;
      vm1 = v * a.bigi;
      v = vn * a.bigi;
    } else {
      vm1 = v;
      v = vn;
    }
  }
  return nev;
}

// final value of a column: u * bigi for each rescaling at or after it, then
// the row factor (the reference multiplies in that order)
__device__ __forceinline__ double ref_final(double u, int ev, int nev, double bigi, double rn) {
#pragma clang fp contract(off)
  for (int e = ev; e < nev; ++e) u = u * bigi;
  return u * rn;
}

__global__ void k_ref_rows_first(const RefPlaneArgs a) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int L = a.L, nk = a.nk;
  if (id >= tri0p(L)) return;
  int l = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
  while (tri0p(l + 1) <= id) ++l;
  while (tri0p(l) > id) --l;
  const int k = (int)(id - tri0p(l));  // row a = -k, 1-based i = l + 1 - k
  const int i = l + 1 - k, ncol = i;
  // pass 1 counts the rescalings; pass 2 reruns the same recursion and writes
  const int nev = ref_row(a, l, i, ncol, true, [](int, double, int) {});
  if (nev > kMaxRescale) {
    *a.status = 1;
    return;
  }
  const double rn = a.rn1[(tri0p(l) + k) * 4 + nev];
  const long long off = (long long)l * (l + 1) * (2 * l + 1) / 6;
  const int lp = l + 1;
  ref_row(a, l, i, ncol, true, [&](int j, double u, int ev) {
    const double v = ref_final(u, ev, nev, a.bigi, rn);
    const int M = l + 1 - j;  // element (-k, -M)
    a.Sa[off + (long long)M * lp + k] = v;
    const double vs = ((M - k) & 1) ? -v : v;
    if (M > k) a.Sa[off + (long long)k * lp + M] = vs;
    if (k < nk) a.thin[rowAp(M, l, L) * nk + k] = v;
    if (M < nk && M > k) a.thin[rowAp(k, l, L) * nk + M] = vs;
    if (k == 0 && M >= 1 && M < nk) a.thinneg[rowAp(0, l, L) * nk + M] = v;
  });
}

__global__ void k_ref_rows_second(const RefPlaneArgs a) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int L = a.L, nk = a.nk;
  if (nk < 2 || id >= L * (nk - 1)) return;
  const int l = id / (nk - 1), ar = id % (nk - 1) + 1;  // row a = ar > 0
  if (ar > l) return;
  const int i = l + 1 + ar;
  const int ncol = 2 * l + 2 - i;  // = l + 1 - a
  const bool rec = i <= 2 * l;     // row 2l+1 holds column 1 only
  const int nev = rec ? ref_row(a, l, i, ncol, false, [](int, double, int) {}) : 0;
  if (nev > kMaxRescale) {
    *a.status = 1;
    return;
  }
  const double rn = a.rn2[((long long)l * (nk - 1) + ar - 1) * 4 + nev];
  auto put = [&](int j, double u, int ev) {
    const double v = ref_final(u, ev, nev, a.bigi, rn);
    const int cc = l + 1 - j;  // element (a, -cc), cc in [a, l]
    if (cc < nk) a.thinneg[rowAp(ar, l, L) * nk + cc] = v;
    if (cc > ar) a.thinneg[rowAp(cc, l, L) * nk + ar] = v;
  };
  if (rec) ref_row(a, l, i, ncol, false, put);
  else put(1, 1.0, 0);
}

}  // namespace

long long ref_plane_doubles(int L) { return (long long)L * (L + 1) * (2LL * L + 1) / 6; }

int build_ref_planes(int L, int nk, double *d_Sa, double *d_thin, double *d_thinneg, hipStream_t st) {
  // host inputs, computed as s2dw_dl_beta_operator_core computes them
  // (s2dw_dl_mod.f90:118-173) with beta = PION2 (s2dw_types_mod.f90)
  const double beta = 1.570796326794896619231321691639751442099;
  const double c = std::cos(beta), s = std::sin(beta), si = 1.0 / s, t = std::tan(-beta / 2.0);
  const double c2 = std::cos(beta / 2.0), omc = 1.0 - c;
  const double big_const = 1.0e150, big = big_const, lbig = std::log(big), bigi = 1.0 / big_const;
  const long long ncoef = (long long)L * (L - 1) / 2;
  std::vector<double> cpi(std::max(1LL, ncoef)), cp2(std::max(1LL, ncoef));
  const int nneg = std::max(1, nk - 1);
  std::vector<double> rn1((size_t)tri0p(L) * 4), rn2((size_t)L * nneg * 4, 0.0);
  std::vector<double> lfr(2 * (size_t)L + 2), sgn(2 * (size_t)L + 2);
  for (int l = 0; l < L; ++l) {
    double *cpl = cpi.data() + (long long)l * (l - 1) / 2 - 1, *cp2l = cp2.data() + (long long)l * (l - 1) / 2 - 1;
    const double xllp1 = (double)(l * (l + 1));
    std::vector<double> cp(l + 1);
    for (int m = 1; m <= l; ++m) {
      const double xm = (double)(l - m);
      cpl[m] = 2.0 / std::sqrt(xllp1 - xm * (xm + 1.0));
      cp[m] = 1.0 / cpl[m];
    }
    for (int m = 2; m <= l; ++m) cp2l[m] = cpl[m] * cp[m - 1];
    if (l >= 1) cp2l[1] = 0.0;
    // row factors
    const int n = 2 * l + 1;
    lfr[1] = (2.0 * (double)l) * std::log(std::fabs(c2));
    sgn[1] = 1.0;
    for (int i = 2; i <= n; ++i) {
      const int m = l + 1 - i;
      const double ratio = std::sqrt((double)(l + m + 1) / (double)(l - m));
      lfr[i] = lfr[i - 1] + std::log(ratio) + std::log(std::fabs(t));
      sgn[i] = sgn[i - 1] * t / std::fabs(t);
    }
    auto factors = [&](int i, double *out) {
      double lr = 0.0;
      for (int e = 0; e < 4; ++e) {
        out[e] = sgn[i] * std::exp(lfr[i] - lr);
        lr = lr - lbig;
      }
    };
    for (int k = 0; k <= l; ++k) factors(l + 1 - k, &rn1[(size_t)(tri0p(l) + k) * 4]);
    for (int ar = 1; ar < nk && ar <= l; ++ar) factors(l + 1 + ar, &rn2[((size_t)l * nneg + ar - 1) * 4]);
  }
  double *d_cpi = nullptr, *d_cp2 = nullptr, *d_rn1 = nullptr, *d_rn2 = nullptr;
  int *d_status = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(d_cpi);
    (void)hipFree(d_cp2);
    (void)hipFree(d_rn1);
    (void)hipFree(d_rn2);
    (void)hipFree(d_status);
  };
  auto fail = [&](hipError_t e, const char *what) {
    cleanup();
    return hip_error(e, what);
  };
  hipError_t e;
  if ((e = hipMalloc(&d_cpi, cpi.size() * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d_cp2, cp2.size() * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d_rn1, rn1.size() * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d_rn2, rn2.size() * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&d_status, sizeof(int))) != hipSuccess) return fail(e, "hipMalloc");
  const long long nrowsA = tri0p(L);
  if ((e = hipMemcpyAsync(d_cpi, cpi.data(), cpi.size() * sizeof(double), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(d_cp2, cp2.data(), cp2.size() * sizeof(double), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(d_rn1, rn1.data(), rn1.size() * sizeof(double), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(d_rn2, rn2.data(), rn2.size() * sizeof(double), hipMemcpyHostToDevice, st)) ||
      (e = hipMemsetAsync(d_status, 0, sizeof(int), st)) ||
      (e = hipMemsetAsync(d_thin, 0, (size_t)nrowsA * nk * sizeof(double), st)) ||
      (e = hipMemsetAsync(d_thinneg, 0, (size_t)nrowsA * nk * sizeof(double), st)))
    return fail(e, "plane table upload");
  RefPlaneArgs a{};
  a.L = L;
  a.nk = nk;
  a.omc = omc;
  a.c = c;
  a.s = s;
  a.si = si;
  a.big = big;
  a.bigi = bigi;
  a.cpi = d_cpi;
  a.cp2 = d_cp2;
  a.rn1 = d_rn1;
  a.rn2 = d_rn2;
  a.Sa = d_Sa;
  a.thin = d_thin;
  a.thinneg = d_thinneg;
  a.status = d_status;
  hipLaunchKernelGGL(k_ref_rows_first, dim3((unsigned)((nrowsA + 127) / 128)), dim3(128), 0, st, a);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "k_ref_rows_first");
  if (nk > 1) {
    const int n2 = L * (nk - 1);
    hipLaunchKernelGGL(k_ref_rows_second, dim3((n2 + 127) / 128), dim3(128), 0, st, a);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "k_ref_rows_second");
  }
  int status = 0;
  if ((e = hipMemcpyAsync(&status, d_status, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return fail(e, "plane status");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "plane build");
  cleanup();
  if (status) {
    set_error("a Wigner plane row needs more than %d rescalings (B=%d)", kMaxRescale, L);
    return kErrArth;
  }
  return kErrNone;
}

}  // namespace s2dw
