// s2dw_setup.cpp -- host-side setup of the S2DW transform: band limits,
// admissible kernels K_j(l) and scaling function Phi^2(l), directionality
// S(l,n), admissibility check.  FP64 on the CPU: these are once-per-
// configuration quadratures whose outputs enter the transform as inputs, and
// they must reproduce the reference bit-for-bit (K_j feeds every wavelet
// coefficient).  Each routine follows the reference routine it names
// (src/mod/s2dw_core_mod.f90).
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "s2dw_internal.h"

namespace s2dw {

// s2dw_types_mod.f90:48-60; TOL_ADMISS is a default-real literal there.
static const double kTolAdmiss = static_cast<double>(1e-12f);
static const double kTolQuad = 1e-12;
static const double kTolLimit = 1e-12;
static const double kTolCeil = 1e-5;

// real**integer as the Fortran compilers emit it (binary powering).
double powi(double a, int b) {
  const bool recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

// s2dw_core_assimilate_int, s2dw_core_mod.f90:1398-1415.
bool assimilate_int(double x, int *y) {
  const double f = std::floor(x), c = std::ceil(x);
  if (std::fabs(x - f) < kTolCeil) {
    *y = static_cast<int>(f);
    return true;
  }
  if (std::fabs(x - c) < kTolCeil) {
    *y = static_cast<int>(c);
    return true;
  }
  *y = 0;
  return false;
}

// s2dw_core_comp_Jmax, s2dw_core_mod.f90:1365-1376.
int comp_Jmax(int B, double alpha) {
  const double x = std::log(static_cast<double>(B)) / std::log(alpha);
  int j;
  if (!assimilate_int(x, &j)) j = static_cast<int>(std::ceil(x));
  return j;
}

// s2dw_core_params_valid, s2dw_core_mod.f90:1327-1345.
bool params_valid(int J, int B, double alpha, int N) {
  const int jmax = comp_Jmax(B, alpha);
  if (J > jmax || J < 1) return false;
  if (N >= 0 && N > B) return false;
  return true;
}

// Band limits of scale jj, s2dw_core_mod.f90:606-616.
void band_limits(int jj, int B, double alpha, int *hi, int *lo) {
  int h, l;
  const double xh = B / powi(alpha, jj - 1), xl = B / powi(alpha, jj + 1);
  if (!assimilate_int(xh, &h)) h = static_cast<int>(std::ceil(xh));
  if (!assimilate_int(xl, &l)) l = static_cast<int>(std::floor(xl));
  *hi = h < B ? h : B;
  *lo = l > 0 ? l : 0;
}

namespace {

// s2dw_core_phi2_integrand, s2dw_core_mod.f90:182-199.
double phi2_integrand(double k, int bw, double alpha) {
  int bw_lo;
  const double q = static_cast<double>(bw) / alpha;
  if (!assimilate_int(q, &bw_lo)) bw_lo = static_cast<int>(std::floor(q));
  const double t = 2.0 * ((k - bw_lo) / static_cast<double>(bw - bw_lo)) - 1;
  return std::exp(-2.0 / (1.0 - t * t)) / k;
}

// Extended trapezoid refinement (trapzd, s2dw_core_mod.f90:1586-1621) driven
// by Simpson's rule with the same stopping test (qsimp, 1724-1759).
int simpson(int bw, double alpha, double a, double b, double *out) {
  double s = 0.0, os = 0.0, ost = 0.0;
  for (int j = 1; j <= 40; ++j) {
    if (j == 1) {
      s = 0.5 * (b - a) * (phi2_integrand(a, bw, alpha) + phi2_integrand(b, bw, alpha));
    } else {
      const int it = 1 << (j - 2);
      const double del = (b - a) / static_cast<double>(it);
      double x = a + 0.5 * del, sum = 0.0;
      for (int k = 0; k < it; ++k) {
        sum += phi2_integrand(x, bw, alpha);
        x += del;
      }
      s = 0.5 * (s + (b - a) * sum / static_cast<double>(it));
    }
    const double v = (4.0 * s - ost) / 3.0;
    *out = v;
    if (j > 5 && (std::fabs(v - os) < kTolQuad * std::fabs(os) || (v == 0.0 && os == 0.0))) return kErrNone;
    os = v;
    ost = s;
  }
  return kErrQuadStepExceed;
}

// logfact, s2dw_core_mod.f90:1497-1544: Lanczos series whose constants are
// default-real (single precision) literals in the reference.
double logfact(int n) {
  static const double c[6] = {double(76.18009172947146f),  double(-86.50532032941677f),
                              double(24.01409824083091f),  double(-1.231739572450155f),
                              double(0.1208650973866179e-2f), double(-0.5395239384953e-5f)};
  const double x = static_cast<double>(n) + 1.0;
  const double temp = x + 5.5 - (x + 0.5) * std::log(x + 5.5);
  double sum = double(1.000000000190015f), y = x;
  for (int k = 0; k < 6; ++k) {
    y += 1.0;
    sum += c[k] / y;
  }
  return -temp + std::log(double(2.5066282746310005f) * sum / x);
}

// binomial, s2dw_core_mod.f90:1470-1478.
int binomial(int n, int r) {
  return static_cast<int>(std::floor(0.5 + std::exp(logfact(n) - logfact(r) - logfact(n - r))));
}

}  // namespace

// s2dw_core_init_kernels, s2dw_core_mod.f90:97-157.
int init_kernels(double *K, double *Phi2, int *bl_scoeff, int J, int B, double alpha) {
  if (!params_valid(J, B, alpha, -1)) return kErrSizeInvalid;
  std::vector<double> prev(B), cur(B);
  int bw = 0, err = kErrNone;
  for (int jj = 0; jj <= J + 1; ++jj) {
    int el_lo;
    const double xb = B / powi(alpha, jj - 1);
    if (!assimilate_int(xb, &bw)) bw = static_cast<int>(std::ceil(xb));
    const double xl = static_cast<double>(bw) / alpha;
    if (!assimilate_int(xl, &el_lo)) el_lo = static_cast<int>(std::floor(xl));
    double cphi = 0.0;
    int e = simpson(bw, alpha, static_cast<double>(el_lo) + kTolLimit, static_cast<double>(bw) - kTolLimit, &cphi);
    if (e) err = e;
#pragma omp parallel for schedule(dynamic, 4)
    for (int el = 0; el < B; ++el) {
      double v;
      if (el <= el_lo) {
        v = 1.0;
      } else if (el >= bw) {
        v = 0.0;
      } else {
        double q = 0.0;
        const int e2 = simpson(bw, alpha, static_cast<double>(el), static_cast<double>(bw) - kTolLimit, &q);
        if (e2) {
#pragma omp critical
          err = e2;
        }
        v = q / cphi;
      }
      cur[el] = v;
      if (jj >= 1) {
        const double d = prev[el] - v;
        K[(jj - 1) + static_cast<size_t>(J + 1) * el] = d < 0.0 ? 0.0 : std::sqrt(d);
      }
    }
    prev.swap(cur);
  }
  for (int el = 0; el < B; ++el) Phi2[el] = prev[el];
  *bl_scoeff = bw;
  return err;
}

// s2dw_core_init_directionality, s2dw_core_mod.f90:219-250.
int init_directionality(double *S, int B, int N) {
  if (N > B) return kErrSizeInvalid;
  std::memset(S, 0, sizeof(double) * 2 * static_cast<size_t>(B) * N);
  const bool eta_real = ((N - 1) % 2) == 0;
  for (int el = 1; el < B; ++el) {
    const int mmax = el < N - 1 ? el : N - 1;
    for (int m = 0; m <= mmax; ++m) {
      if (((N + m) % 2) == 0) continue;  // betah = 0
      int g = el - (((N + el) % 2) == 0 ? 1 : 0);
      if (g > N - 1) g = N - 1;
      const int binom = binomial(g, (g - m) / 2);
      const double v = std::sqrt(static_cast<double>(binom) / powi(2.0, g));
      const size_t k = 2 * (el + static_cast<size_t>(B) * m);
      S[k + (eta_real ? 0 : 1)] = v;
    }
  }
  return kErrNone;
}

// s2dw_core_admiss, s2dw_core_mod.f90:285-304.
bool admiss(double *a, const double *K, const double *Phi2, int B, int J) {
  bool pass = true;
  for (int el = 0; el < B; ++el) {
    double v = Phi2[el];
    for (int jj = 0; jj <= J; ++jj) {
      const double k = K[jj + static_cast<size_t>(J + 1) * el];
      v += k * k;
    }
    a[el] = v;
    if (std::fabs(v - 1.0) > kTolAdmiss) pass = false;
  }
  return pass;
}

// quad_weights, s2dw_core_mod.f90:1436-1450 (Driscoll-Healy weights).
double quad_weight(double theta, int B) {
  double w = 0.0;
  for (int k = 0; k < B; ++k) w += std::sin((2.0 * k + 1.0) * theta) / (2.0 * k + 1.0);
  return (2.0 / static_cast<double>(B)) * std::sin(theta) * w;
}

}  // namespace s2dw
