/*
 * s2dw_oracle.c -- CPU restatement of the S2DW directional wavelet transform.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline of the project.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product (s2dw_amd/, libs2dw_hip.so) never links, loads or calls it.
 *
 * Every routine restates, stage for stage and loop for loop, the routine of
 * the reference it names (file:line relative to the reference root, Fortran
 * 90, column-major, first index fastest).  Arrays keep the reference layouts:
 *   flm(0:B-1,0:B-1)         complex, flm[l + B*m]
 *   K_gamma(0:J,0:B-1)       real,    K[j + (J+1)*l]
 *   Slm(0:B-1,0:N-1)         complex, S[l + B*n]
 *   wavdyn(jj)%coeff(0:2hi-2,0:2hi-1,0:N-1) real, w[a + (2hi-1)*(b + 2hi*g)]
 *   wav(0:J,0:2B-2,0:2B-1,0:N-1) real, jj fastest (static layout)
 *   scoeff(0:bl-1,0:bl-1)    complex, sc[l + bl*m]
 *
 * The one part of the reference path that is not the reference's own code is
 * FFTW (un-vendored, absent from the image).  Its legacy Fortran calls
 * dfftw_execute_dft_c2r / _r2c (s2dw_core_mod.f90:752-767, 1112-1128) are
 * restated here from FFTW's documented definition: unnormalised, backward
 * transform e^{+2 pi i k a/n} for c2r (imaginary part of X_0 ignored),
 * forward e^{-2 pi i k a/n} for r2c, computed with a mixed-radix FFT of our
 * own (orc_fft_*), exact to rounding.
 *
 * Parity pins (tests/): setup routines and Wigner planes are checked against
 * the reference's own Fortran, compiled from its sources (oracle/_ref) into
 * golden fixtures; the transform against the defining sum (brute force with
 * mpmath Wigner matrices at L<=12), the N=1 reduction to scipy's spherical
 * harmonics, and the exact round trip synthesis(analysis(f)) = f.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* s2dw_types_mod.f90:48-60 (TOL_ADMISS and TOL_ZERO are default-real
 * literals, so they carry single-precision rounding). */
#define ORC_TOL_ADMISS ((double)1e-12f)
#define ORC_TOL_ZERO ((double)1e-12f)
#define ORC_TOL_QUAD 1e-12
#define ORC_TOL_LIMIT 1e-12
#define ORC_TOL_CEIL 1e-5
/* s2dw_types_mod.f90:71-74 */
static const double ORC_PI = 3.141592653589793238462643383279502884197;
static const double ORC_PION2 = 1.570796326794896619231321691639751442099;

/* error codes, s2dw_error_mod.f90:37-52 */
enum {
  ORC_ERR_NONE = 0,
  ORC_ERR_MEM_ALLOC_FAIL = 4,
  ORC_ERR_ARTH = 5,
  ORC_ERR_ARTH_WARNING = 6,
  ORC_ERR_SIZE_INVALID = 8,
  ORC_ERR_QUAD_STEP_EXCEED = 12,
  ORC_ERR_ADMISS_FAIL = 13
};

static int orc_threads = 0; /* 0: OpenMP default */

void orc_set_threads(int n) { orc_threads = n; }

int orc_get_threads(void) {
#ifdef _OPENMP
  return orc_threads > 0 ? orc_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* real**integer as compiled by flang/gfortran (compiler-rt __powidf2). */
static double orc_powi(double a, int b) {
  const int recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

/* s2dw_core_assimilate_int, s2dw_core_mod.f90:1398-1415 */
int orc_assimilate_int(double x, int *y) {
  if (fabs(x - floor(x)) < ORC_TOL_CEIL) {
    *y = (int)floor(x);
    return 1;
  } else if (fabs(x - ceil(x)) < ORC_TOL_CEIL) {
    *y = (int)ceil(x);
    return 1;
  }
  *y = 0;
  return 0;
}

/* s2dw_core_comp_Jmax, s2dw_core_mod.f90:1365-1376 */
int orc_comp_Jmax(int B, double alpha) {
  int J_max;
  double x = log((double)B) / log(alpha);
  if (!orc_assimilate_int(x, &J_max)) J_max = (int)ceil(x);
  return J_max;
}

/* s2dw_core_params_valid, s2dw_core_mod.f90:1327-1345 (N < 0: absent) */
int orc_params_valid(int J, int B, double alpha, int N) {
  int valid = 1;
  int J_max = orc_comp_Jmax(B, alpha);
  if (J > J_max || J < 1) valid = 0;
  if (N >= 0 && N > B) valid = 0;
  return valid;
}

/* Band limits of scale jj, s2dw_core_mod.f90:606-616 (identical at 372-381,
 * 850-859, 1093-1102). */
void orc_band_limits(int jj, int B, double alpha, int *bl_hi, int *bl_lo) {
  int hi, lo;
  if (orc_assimilate_int(B / orc_powi(alpha, jj - 1), &hi)) {
    hi = hi < B ? hi : B;
  } else {
    hi = (int)ceil(B / orc_powi(alpha, jj - 1));
    hi = hi < B ? hi : B;
  }
  if (orc_assimilate_int(B / orc_powi(alpha, jj + 1), &lo)) {
    lo = lo > 0 ? lo : 0;
  } else {
    lo = (int)floor(B / orc_powi(alpha, jj + 1));
    lo = lo > 0 ? lo : 0;
  }
  *bl_hi = hi;
  *bl_lo = lo;
}

/* s2dw_core_phi2_integrand, s2dw_core_mod.f90:182-199 */
static double orc_phi2_integrand(double k, int bw, double alpha) {
  int bw_lo;
  if (!orc_assimilate_int((double)bw / alpha, &bw_lo)) bw_lo = (int)floor((double)bw / alpha);
  double t = 2.0 * ((k - bw_lo) / (double)(bw - bw_lo)) - 1;
  return exp(-2.0 / (1.0 - t * t)) / k;
}

/* trapzd, s2dw_core_mod.f90:1586-1621 */
static void orc_trapzd(int bw, double alpha, double a_lim, double b_lim, double *s, int n) {
  if (n == 1) {
    *s = 0.5 * (b_lim - a_lim) *
         (orc_phi2_integrand(a_lim, bw, alpha) + orc_phi2_integrand(b_lim, bw, alpha));
  } else {
    int it = 1 << (n - 2);
    double del = (b_lim - a_lim) / (double)it;
    double x = a_lim + 0.5 * del;
    double fsum = 0.0;
    for (int j = 1; j <= it; j++) {
      fsum = fsum + orc_phi2_integrand(x, bw, alpha);
      x = x + del;
    }
    *s = 0.5 * (*s + (b_lim - a_lim) * fsum / (double)it);
  }
}

/* qsimp, s2dw_core_mod.f90:1724-1759.  Returns ORC_ERR_QUAD_STEP_EXCEED on
 * non-convergence (the reference halts there). */
static int orc_qsimp(int bw, double alpha, double a_lim, double b_lim, double *integral) {
  double os = 0.0, ost = 0.0, st = 0.0;
  for (int j = 1; j <= 40; j++) {
    orc_trapzd(bw, alpha, a_lim, b_lim, &st, j);
    *integral = (4.0 * st - ost) / 3.0;
    if (j > 5) {
      if (fabs(*integral - os) < ORC_TOL_QUAD * fabs(os) || (*integral == 0.0 && os == 0.0))
        return ORC_ERR_NONE;
    }
    os = *integral;
    ost = st;
  }
  return ORC_ERR_QUAD_STEP_EXCEED;
}

/* s2dw_core_init_kernels, s2dw_core_mod.f90:97-157 */
int orc_init_kernels(double *K_gamma, double *Phi2, int *bl_scoeff, int J, int B, double alpha) {
  if (!orc_params_valid(J, B, alpha, -1)) return ORC_ERR_SIZE_INVALID;
  double *Phi2_all = (double *)malloc(sizeof(double) * (size_t)(J + 2) * B);
  if (!Phi2_all) return ORC_ERR_MEM_ALLOC_FAIL;
  int bw = 0, el_lo, err = ORC_ERR_NONE;
#define P2A(jj, el) Phi2_all[(jj) + (size_t)(J + 2) * (el)]
  for (int jj = 0; jj <= J + 1; jj++) {
    if (!orc_assimilate_int(B / orc_powi(alpha, jj - 1), &bw)) bw = (int)ceil(B / orc_powi(alpha, jj - 1));
    if (!orc_assimilate_int((double)bw / alpha, &el_lo)) el_lo = (int)floor((double)bw / alpha);
    double C_Phi;
    int e = orc_qsimp(bw, alpha, (double)el_lo + ORC_TOL_LIMIT, (double)bw - ORC_TOL_LIMIT, &C_Phi);
    if (e) err = e;
    /* the el loop is independent per el; parallel like the reference's
     * loop-level parallelism, results are identical (no reductions). */
#pragma omp parallel for schedule(dynamic, 4) num_threads(orc_get_threads())
    for (int el = 0; el <= B - 1; el++) {
      if (el <= el_lo) {
        P2A(jj, el) = 1;
      } else if (el >= bw) {
        P2A(jj, el) = 0;
      } else {
        double q;
        int e2 = orc_qsimp(bw, alpha, (double)el, (double)bw - ORC_TOL_LIMIT, &q);
        if (e2) {
#pragma omp critical
          err = e2;
        }
        P2A(jj, el) = q / C_Phi;
      }
      if (jj >= 1) {
        double Phi2_diff = P2A(jj - 1, el) - P2A(jj, el);
        K_gamma[(jj - 1) + (size_t)(J + 1) * el] = Phi2_diff < 0.0 ? 0.0 : sqrt(Phi2_diff);
      }
    }
  }
  for (int el = 0; el < B; el++) Phi2[el] = P2A(J + 1, el);
  *bl_scoeff = bw;
#undef P2A
  free(Phi2_all);
  return err;
}

/* logfact, s2dw_core_mod.f90:1497-1544.  The fitting constants are
 * default-real literals in the reference, hence the float casts. */
static double orc_logfact(int n) {
  double c[6] = {(double)76.18009172947146f, (double)-86.50532032941677f,
                 (double)24.01409824083091f, (double)-1.231739572450155f,
                 (double)0.1208650973866179e-2f, (double)-0.5395239384953e-5f};
  double x = (double)n + 1.0;
  double temp = x + 5.5 - (x + 0.5) * log(x + 5.5);
  double sum = (double)1.000000000190015f;
  double y = x;
  for (int nn = 0; nn < 6; nn++) {
    y = y + 1.0;
    sum = sum + c[nn] / y;
  }
  return -temp + log((double)2.5066282746310005f * sum / x);
}

/* binomial, s2dw_core_mod.f90:1470-1478 */
static int orc_binomial(int n, int r) {
  return (int)floor(0.5 + exp(orc_logfact(n) - orc_logfact(r) - orc_logfact(n - r)));
}

/* s2dw_core_init_directionality, s2dw_core_mod.f90:219-250 */
int orc_init_directionality(cplx *Slm, int B, int N) {
  if (N > B) return ORC_ERR_SIZE_INVALID;
  for (size_t i = 0; i < (size_t)B * N; i++) Slm[i] = 0;
  for (int el = 1; el <= B - 1; el++) {
    int mmax = el < N - 1 ? el : N - 1;
    for (int m = 0; m <= mmax; m++) {
      int betah = (1 - (((N + m) % 2) ? -1 : 1)) / 2;
      if (betah > 0) {
        int g = el - (1 + (((N + el) % 2) ? -1 : 1)) / 2;
        if (N - 1 < g) g = N - 1;
        int binom = orc_binomial(g, (g - m) / 2);
        cplx eta = ((N - 1) % 2 == 0) ? 1.0 : I;
        Slm[el + (size_t)B * m] = sqrt((double)binom / orc_powi(2.0, g)) * eta;
      }
    }
  }
  return ORC_ERR_NONE;
}

/* s2dw_core_admiss, s2dw_core_mod.f90:285-304 */
int orc_admiss(double *admiss, const double *K_gamma, const double *Phi2, int B, int J) {
  int pass = 1;
  for (int el = 0; el < B; el++) {
    admiss[el] = Phi2[el];
    for (int jj = 0; jj <= J; jj++) {
      double k = K_gamma[jj + (size_t)(J + 1) * el];
      admiss[el] = admiss[el] + k * k;
    }
    if (fabs(admiss[el] - 1.0) > ORC_TOL_ADMISS) pass = 0;
  }
  return pass;
}

/* quad_weights, s2dw_core_mod.f90:1436-1450 */
double orc_quad_weights(double theta_t, int B) {
  double w = 0.0;
  for (int k = 0; k <= B - 1; k++) w = w + sin((2.0 * k + 1.0) * theta_t) / (2.0 * k + 1.0);
  return (2.0 / (double)B) * sin(theta_t) * w;
}

/* ------------------------------------------------------------------------
 * Wigner plane d^l_{m'm}(beta): s2dw_dl_beta_operator, s2dw_dl_mod.f90:56-68,
 * with _core (92-238) and _fill (258-304).  dl(a,b), a,b in [-l,l], stored
 * column-major: dl[(a+l) + (2l+1)*(b+l)].
 * ---------------------------------------------------------------------- */
#define DLI(i, j) (((i)-1) + (size_t)(2 * l + 1) * ((j)-1)) /* 1-based (i,j) */

/* The operator recursion of s2dw_dl_beta_operator_core (s2dw_dl_mod.f90:
 * 92-238), written once for a floating type R (double: the restatement;
 * long double: the extended-precision diagnostic).  Output rounded to double. */
#define ORC_DL_CORE(NAME, R, COS, SIN, TAN, LOG, EXP, SQRT, FABS)                                   \
  static void NAME(double *out, double beta_in, int l) {                                           \
    const int n = 2 * l + 1;                                                                       \
    R *dl = (R *)calloc((size_t)n * n, sizeof(R));                                                 \
    const R beta = (R)beta_in;                                                                     \
    if (FABS(beta) < (R)1e-5) {                                                                    \
      for (int i = 1; i <= n; i++) dl[DLI(i, i)] = 1;                                              \
    } else {                                                                                       \
      const R big_const = (R)1.0e150;                                                              \
      R *cp = (R *)calloc(n + 1, sizeof(R)), *cpi = (R *)calloc(n + 1, sizeof(R));                 \
      R *cp2 = (R *)calloc(n + 1, sizeof(R)), *lfr = (R *)calloc(n + 1, sizeof(R));                \
      R *sgn = (R *)calloc(n + 1, sizeof(R)), *lrenorm = (R *)calloc(n + 1, sizeof(R));            \
      R ang = beta;                                                                                \
      R c = COS(ang), s = SIN(ang), si = 1 / s, t = TAN(-ang / 2), c2 = COS(ang / 2);              \
      R omc = 1 - c;                                                                               \
      lfr[1] = (2 * (R)l) * LOG(FABS(c2));                                                         \
      sgn[1] = 1;                                                                                  \
      for (int i = 2; i <= n; i++) {                                                               \
        int m = l + 1 - i;                                                                         \
        R ratio = SQRT((R)(l + m + 1) / (R)(l - m));                                               \
        lfr[i] = lfr[i - 1] + LOG(ratio) + LOG(FABS(t));                                           \
        sgn[i] = sgn[i - 1] * t / FABS(t);                                                         \
      }                                                                                            \
      R big = big_const, lbig = LOG(big), bigi = 1 / big_const;                                    \
      R xllp1 = (R)(l * (l + 1));                                                                  \
      for (int m = 1; m <= l; m++) {                                                               \
        R xm = (R)(l - m);                                                                         \
        cpi[m] = 2 / SQRT(xllp1 - xm * (xm + 1));                                                  \
        cp[m] = 1 / cpi[m];                                                                        \
      }                                                                                            \
      for (int m = 2; m <= l; m++) cp2[m] = cpi[m] * cp[m - 1];                                    \
      dl[DLI(1, 1)] = 1;                                                                           \
      dl[DLI(2 * l + 1, 1)] = 1;                                                                   \
      for (int index = 2; index <= l + 1; index++) {                                               \
        dl[DLI(index, 1)] = 1;                                                                     \
        R lambda = ((R)(l + 1) * omc - (R)index + c) * si;                                         \
        dl[DLI(index, 2)] = lambda * dl[DLI(index, 1)] * cpi[1];                                   \
        if (index > 2) {                                                                           \
          for (int m = 2; m <= index - 1; m++) {                                                   \
            lambda = ((R)(l + 1) * omc - (R)index + (R)m * c) / s;                                 \
            dl[DLI(index, m + 1)] = lambda * cpi[m] * dl[DLI(index, m)] - cp2[m] * dl[DLI(index, m - 1)]; \
            if (dl[DLI(index, m + 1)] > big) {                                                     \
              lrenorm[index] = lrenorm[index] - lbig;                                              \
              for (int im = 1; im <= m + 1; im++) dl[DLI(index, im)] = dl[DLI(index, im)] * bigi;  \
            }                                                                                      \
          }                                                                                        \
        }                                                                                          \
      }                                                                                            \
      for (int index = l + 2; index <= 2 * l; index++) {                                           \
        dl[DLI(index, 1)] = 1;                                                                     \
        R lambda = ((R)(l + 1) * omc - (R)index + c) / s;                                          \
        dl[DLI(index, 2)] = lambda * dl[DLI(index, 1)] * cpi[1];                                   \
        if (index < 2 * l) {                                                                       \
          for (int m = 2; m <= 2 * l - index + 1; m++) {                                           \
            lambda = ((R)(l + 1) * omc - (R)index + (R)m * c) / s;                                 \
            dl[DLI(index, m + 1)] = lambda * cpi[m] * dl[DLI(index, m)] - cp2[m] * dl[DLI(index, m - 1)]; \
            if (dl[DLI(index, m + 1)] > big) {                                                     \
              for (int im = 1; im <= m + 1; im++) dl[DLI(index, im)] = dl[DLI(index, im)] * bigi;  \
              lrenorm[index] = lrenorm[index] - lbig;                                              \
            }                                                                                      \
          }                                                                                        \
        }                                                                                          \
      }                                                                                            \
      for (int i = 1; i <= l + 1; i++) {                                                           \
        R renorm = sgn[i] * EXP(lfr[i] - lrenorm[i]);                                              \
        for (int j = 1; j <= i; j++) dl[DLI(i, j)] = dl[DLI(i, j)] * renorm;                       \
      }                                                                                            \
      for (int i = l + 2; i <= 2 * l + 1; i++) {                                                   \
        R renorm = sgn[i] * EXP(lfr[i] - lrenorm[i]);                                              \
        for (int j = 1; j <= 2 * l + 2 - i; j++) dl[DLI(i, j)] = dl[DLI(i, j)] * renorm;           \
      }                                                                                            \
      free(cp); free(cpi); free(cp2); free(lfr); free(sgn); free(lrenorm);                        \
    }                                                                                              \
    for (size_t k = 0; k < (size_t)n * n; k++) out[k] = (double)dl[k];                             \
    free(dl);                                                                                      \
  }

ORC_DL_CORE(orc_dl_core, double, cos, sin, tan, log, exp, sqrt, fabs)
ORC_DL_CORE(orc_dl_core_ld, long double, cosl, sinl, tanl, logl, expl, sqrtl, fabsl)

static void orc_dl_fill(double *dl, int l) {
  for (int i = 1; i <= l; i++)
    for (int j = l + 1; j <= 2 * l + 1 - i; j++) dl[DLI(2 * l + 2 - i, 2 * l + 2 - j)] = dl[DLI(j, i)];
  for (int i = 1; i <= l + 1; i++) {
    int sgn = -1;
    for (int j = i + 1; j <= l + 1; j++) {
      dl[DLI(i, j)] = dl[DLI(j, i)] * sgn;
      sgn = -sgn;
    }
  }
  for (int i = l + 2; i <= 2 * l + 1; i++) {
    int sgn = ((i + 1) % 2) ? -1 : 1;
    for (int j = 1; j <= 2 * l + 2 - i; j++) {
      dl[DLI(j, i)] = dl[DLI(i, j)] * sgn;
      sgn = -sgn;
    }
    for (int j = i; j <= 2 * l + 1; j++) dl[DLI(j, i)] = dl[DLI(2 * l + 2 - i, 2 * l + 2 - j)];
  }
  for (int i = l + 2; i <= 2 * l + 1; i++)
    for (int j = 2 * l + 3 - i; j <= i - 1; j++) dl[DLI(j, i)] = dl[DLI(2 * l + 2 - i, 2 * l + 2 - j)];
}
#undef DLI

/* s2dw_dl_beta_operator, s2dw_dl_mod.f90:56-68 */
static int orc_dl_precise = 0;

/* Diagnostic switch: 1 computes the Wigner planes with the same recursion in
 * long double (x87 80-bit) arithmetic, rounded to double at the end.  It shows
 * how much of a difference against another implementation comes from the
 * rounding of the reference's double-precision planes.  Default 0. */
void orc_set_dl_precise(int on) { orc_dl_precise = on; }

void orc_dl_beta_operator(double *dl, double beta, int l) {
  if (orc_dl_precise)
    orc_dl_core_ld(dl, beta, l);
  else
    orc_dl_core(dl, beta, l);
  orc_dl_fill(dl, l);
}

/* Diagnostic switches (default 0, the reference's arithmetic): 1 replaces
 * the reference's phase factor exp(I*(n-m)*PION2) (core_mod.f90:666, 975)
 * by the exact power of i, resp. its beta-sum basis 2cos(m' beta_b),
 * 2 sin(m' beta_b) (core_mod.f90:720-731, evaluated at the rounded product
 * m' * beta_b) by the exact values (argument reduced in integers, long
 * double).  They measure how far each of the reference's roundings moves its
 * output. */
static int orc_exact_phase = 0, orc_exact_beta = 0;
void orc_set_exact_phase(int on) { orc_exact_phase = on; }
void orc_set_exact_beta(int on) { orc_exact_beta = on; }
static cplx orc_ipow(int k) {
  const int r = ((k % 4) + 4) % 4;
  return r == 0 ? 1.0 : r == 1 ? I : r == 2 ? -1.0 : -I;
}

/* ------------------------------------------------------------------------
 * FFTW substitute (FFTW's documented transform definitions).
 * ---------------------------------------------------------------------- */
typedef struct {
  int n;
  int nf;
  int f[32];
  cplx *tw; /* tw[k] = exp(+2 pi i k / n) */
} orc_fftplan;

static void orc_fftplan_init(orc_fftplan *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  for (int r = 4; r >= 2 && m > 1;) { /* radix-4 first, then 2, then odd */
    if (m % r == 0) {
      p->f[p->nf++] = r;
      m /= r;
    } else {
      r -= 2;
    }
  }
  for (int r = 3; m > 1; r += 2) {
    while (m % r == 0) {
      p->f[p->nf++] = r;
      m /= r;
    }
  }
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  for (int k = 0; k < n; k++) {
    /* angle reduced to [0, pi/4]-ish accuracy via exact integer symmetry */
    double a = 2.0 * ORC_PI * (double)k / (double)n;
    p->tw[k] = cos(a) + I * sin(a);
  }
}

static void orc_fftplan_free(orc_fftplan *p) { free(p->tw); }

/* DIT mixed radix, out-of-place.  sign=+1: backward (e^{+}), -1: forward. */
static void orc_fft_rec(const orc_fftplan *p, int n, const cplx *in, size_t istride, cplx *out, int twstride,
                        int fi, int sign) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int r = p->f[fi], m = n / r;
  for (int q = 0; q < r; q++) orc_fft_rec(p, m, in + q * istride, istride * r, out + (size_t)q * m, twstride * r, fi + 1, sign);
  cplx tmp[64];
  cplx *t = r <= 64 ? tmp : (cplx *)malloc(sizeof(cplx) * r);
  const int N0 = p->n;
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < r; q++) t[q] = out[k + (size_t)q * m];
    for (int q = 0; q < r; q++) {
      cplx s = 0;
      for (int u = 0; u < r; u++) {
        long idx = ((long)u * (k + (long)m * q)) % n;
        cplx w = p->tw[(size_t)idx * twstride % N0];
        if (sign < 0) w = conj(w);
        s += t[u] * w;
      }
      out[k + (size_t)q * m] = s;
    }
  }
  if (t != tmp) free(t);
}

/* dfftw_execute_dft_c2r on n points: in[0..n/2], out[0..n-1] (Im in[0] and,
 * for even n, Im in[n/2] are ignored, as FFTW assumes Hermitian input). */
static void orc_c2r(const orc_fftplan *p, const cplx *in, double *out, cplx *work) {
  int n = p->n;
  cplx *h = work, *o = work + n;
  h[0] = creal(in[0]);
  for (int k = 1; k <= n / 2; k++) {
    h[k] = in[k];
    h[n - k] = conj(in[k]);
  }
  if (n % 2 == 0) h[n / 2] = creal(in[n / 2]);
  orc_fft_rec(p, n, h, 1, o, 1, 0, +1);
  for (int a = 0; a < n; a++) out[a] = creal(o[a]);
}

/* dfftw_execute_dft_r2c on n points: in[0..n-1] real, out[0..n/2]. */
static void orc_r2c(const orc_fftplan *p, const double *in, cplx *out, cplx *work) {
  int n = p->n;
  cplx *h = work, *o = work + n;
  for (int a = 0; a < n; a++) h[a] = in[a];
  orc_fft_rec(p, n, h, 1, o, 1, 0, -1);
  for (int k = 0; k <= n / 2; k++) out[k] = o[k];
}

/* exposed for tests: complex DFT of length n (sign +1 backward, -1 forward) */
void orc_fft(int n, const cplx *in, cplx *out, int sign) {
  orc_fftplan p;
  orc_fftplan_init(&p, n);
  orc_fft_rec(&p, n, in, 1, out, 1, 0, sign);
  orc_fftplan_free(&p);
}

/* ------------------------------------------------------------------------
 * Analysis, one scale: s2dw_core_analysis_flm2wav_dynamic body,
 * s2dw_core_mod.f90:604-768 (identical numerics in the static routine,
 * 369-527).  wj = wavdyn(jj)%coeff.
 * ---------------------------------------------------------------------- */
static int orc_analysis_scale(double *wj, const cplx *flm, const double *K_gamma, const cplx *Slm, int jj, int J,
                              int B, int N, double alpha) {
  int bl_hi, bl_lo;
  orc_band_limits(jj, B, alpha, &bl_hi, &bl_lo);
  const int hi = bl_hi, nN = 2 * N - 1;
  const size_t tsz = (size_t)hi * hi * nN;
  cplx *Tmmm = (cplx *)calloc(tsz, sizeof(cplx));
  if (!Tmmm) return ORC_ERR_MEM_ALLOC_FAIL;
#define FLM(l, m) flm[(l) + (size_t)B * (m)]
#define KG(j, l) K_gamma[(j) + (size_t)(J + 1) * (l)]
#define SLM(l, m) Slm[(l) + (size_t)B * (m)]
  /* Tmmm, 618-674: the reference keeps a private Tmmm per OpenMP thread and
   * sums them (reduction(+:Tmmm)).  Here each thread owns a private copy for a
   * static block of el values and the copies are summed in thread order, so
   * the result is deterministic for a given thread count. */
  int nthr = orc_get_threads();
  int ntot = bl_hi - 1 - bl_lo; /* el = bl_lo+1 .. bl_hi-1 */
  if (nthr > ntot) nthr = ntot > 0 ? ntot : 1;
  cplx **priv = (cplx **)calloc(nthr, sizeof(cplx *));
  double **dlb = (double **)calloc(nthr, sizeof(double *));
  int err = (priv && dlb) ? 0 : ORC_ERR_MEM_ALLOC_FAIL;
  for (int t = 0; t < nthr && !err; t++) {
    priv[t] = t == 0 ? Tmmm : (cplx *)calloc(tsz, sizeof(cplx));
    dlb[t] = (double *)malloc(sizeof(double) * (size_t)(2 * hi + 1) * (2 * hi + 1));
    if (!priv[t] || !dlb[t]) err = ORC_ERR_MEM_ALLOC_FAIL;
  }
  if (!err) {
#pragma omp parallel num_threads(nthr)
    {
#ifdef _OPENMP
      int tid = omp_get_thread_num();
#else
      int tid = 0;
#endif
      cplx *T = priv[tid];
      double *dlbuf = dlb[tid];
#pragma omp for schedule(dynamic, 10)
      for (int el = bl_lo + 1; el <= bl_hi - 1; el++) {
        orc_dl_beta_operator(dlbuf, ORC_PION2, el);
        const int ld = 2 * el + 1;
#define DL(a, b) dlbuf[((a) + el) + (size_t)ld * ((b) + el)]
        double mmmsign = ((N - 1 < el ? N - 1 : el) % 2 == 0) ? -1.0 : 1.0;
        int nmax = N - 1 < el ? N - 1 : el;
        for (int mmm = -nmax; mmm <= nmax; mmm++) {
          mmmsign = -mmmsign;
          for (int mm = 0; mm <= el; mm++) {
            double msign = -1.0;
            for (int m = 0; m <= el; m++) {
              msign = -msign;
              cplx sb;
              if (mmm < 0) {
                sb = mmmsign * KG(jj, el) * SLM(el, -mmm) * FLM(el, m);
              } else {
                sb = conj(KG(jj, el) * SLM(el, mmm)) * FLM(el, m);
              }
              T[m + (size_t)hi * (mm + (size_t)hi * (mmm + N - 1))] +=
                  (orc_exact_phase ? orc_ipow(mmm - m) : cexp(I * (double)(mmm - m) * ORC_PION2)) * DL(mm, m) *
                  DL(mm, mmm) * sb;
            }
          }
        }
#undef DL
      }
    }
  }
  for (int t = 0; t < nthr; t++) {
    if (t > 0 && priv && priv[t]) {
      if (!err)
        for (size_t k = 0; k < tsz; k++) Tmmm[k] += priv[t][k];
      free(priv[t]);
    }
    if (dlb && dlb[t]) free(dlb[t]);
  }
  free(priv);
  free(dlb);
  if (err) {
    free(Tmmm);
    return err;
  }
  /* Tmmg, 676-705 */
  cplx *Tmmg = (cplx *)calloc((size_t)hi * hi * N, sizeof(cplx));
#pragma omp parallel for collapse(2) schedule(dynamic, 10) num_threads(orc_get_threads())
  for (int gg = 0; gg <= N - 1; gg++) {
    for (int mm = 0; mm <= hi - 1; mm++) {
      double gamma_g = ORC_PI * gg / (double)N;
      for (int m = 0; m <= hi - 1; m++) {
        for (int mmm = -(N - 1); mmm <= N - 1; mmm++) {
          int k_indicator = (1 - (((N + mmm) % 2 != 0) ? -1 : 1)) / 2;
          if (k_indicator > 0)
            Tmmg[m + (size_t)hi * (mm + (size_t)hi * gg)] +=
                Tmmm[m + (size_t)hi * (mm + (size_t)hi * (mmm + N - 1))] * cexp(I * (mmm * gamma_g));
        }
      }
    }
  }
  free(Tmmm);
  /* Tmbg, 707-743 */
  cplx *Tmbg = (cplx *)calloc((size_t)hi * 2 * hi * N, sizeof(cplx));
#pragma omp parallel for collapse(2) schedule(dynamic, 10) num_threads(orc_get_threads())
  for (int gg = 0; gg <= N - 1; gg++) {
    for (int bb = 0; bb <= 2 * hi - 1; bb++) {
      double beta_b = ORC_PI * (2.0 * bb + 1.0) / (4.0 * hi);
      for (int mm = 0; mm <= hi - 1; mm++) {
        for (int m = 0; m <= hi - 1; m++) {
          cplx eta;
          if (mm == 0) {
            eta = 1.0;
          } else if (orc_exact_beta) {
            const long double ang = 3.141592653589793238462643383279502884L *
                                    (long double)((long)mm * (2 * bb + 1) % (8L * hi)) / (4.0L * hi);
            eta = ((m + N - 1) % 2 == 0) ? (double)(2.0L * cosl(ang)) : I * (double)(2.0L * sinl(ang));
          } else if ((m + N - 1) % 2 == 0)
            eta = 2.0 * cos(mm * beta_b);
          else
            eta = I * 2.0 * sin(mm * beta_b);
          Tmbg[m + (size_t)hi * (bb + (size_t)2 * hi * gg)] += Tmmg[m + (size_t)hi * (mm + (size_t)hi * gg)] * eta;
        }
      }
    }
  }
  free(Tmmg);
  /* alpha c2r FFT, 745-768 */
  const int na = 2 * hi - 1;
  orc_fftplan p;
  orc_fftplan_init(&p, na);
#pragma omp parallel num_threads(orc_get_threads())
  {
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)na);
#pragma omp for collapse(2) schedule(dynamic, 10)
    for (int gg = 0; gg <= N - 1; gg++) {
      for (int bb = 0; bb <= 2 * hi - 1; bb++) {
        orc_c2r(&p, &Tmbg[(size_t)hi * (bb + (size_t)2 * hi * gg)], &wj[(size_t)na * (bb + (size_t)2 * hi * gg)], work);
      }
    }
    free(work);
  }
  orc_fftplan_free(&p);
  free(Tmbg);
  return ORC_ERR_NONE;
}

/* Scale sizes for callers: hi and lo per scale jj. */
void orc_scale_limits(int J, int B, double alpha, int *hi, int *lo) {
  for (int jj = 0; jj <= J; jj++) orc_band_limits(jj, B, alpha, &hi[jj], &lo[jj]);
}

/* s2dw_core_analysis_flm2wav_dynamic, s2dw_core_mod.f90:572-775.
 * wavdyn[jj] points to caller storage of (2hi-1)*2hi*N doubles. */
int orc_analysis_dynamic(double **wavdyn, cplx *scoeff, const cplx *flm, const double *K_gamma, const cplx *Slm, int J,
                         int B, int N, int bl_scoeff, double alpha) {
  for (int jj = 0; jj <= J; jj++) {
    int e = orc_analysis_scale(wavdyn[jj], flm, K_gamma, Slm, jj, J, B, N, alpha);
    if (e) return e;
  }
  /* scaling coefficients, 772-773 */
  for (int m = 0; m < bl_scoeff; m++)
    for (int l = 0; l < bl_scoeff; l++) scoeff[l + (size_t)bl_scoeff * m] = FLM(l, m);
  return ORC_ERR_NONE;
}

/* s2dw_core_analysis_flm2wav (static layout), s2dw_core_mod.f90:342-534:
 * wav(0:J,0:2B-2,0:2B-1,0:N-1), jj fastest; only the per-scale sub-block
 * (0:2hi-2, 0:2hi-1, :) of each jj is written (509). */
int orc_analysis_static(double *wav, cplx *scoeff, const cplx *flm, const double *K_gamma, const cplx *Slm, int J, int B,
                        int N, int bl_scoeff, double alpha) {
  for (int jj = 0; jj <= J; jj++) {
    int hi, lo;
    orc_band_limits(jj, B, alpha, &hi, &lo);
    size_t n = (size_t)(2 * hi - 1) * 2 * hi * N;
    double *wj = (double *)malloc(sizeof(double) * n);
    if (!wj) return ORC_ERR_MEM_ALLOC_FAIL;
    int e = orc_analysis_scale(wj, flm, K_gamma, Slm, jj, J, B, N, alpha);
    if (e) {
      free(wj);
      return e;
    }
    for (int g = 0; g < N; g++)
      for (int b = 0; b < 2 * hi; b++)
        for (int a = 0; a < 2 * hi - 1; a++)
          wav[jj + (size_t)(J + 1) * (a + (size_t)(2 * B - 1) * (b + (size_t)2 * B * g))] =
              wj[a + (size_t)(2 * hi - 1) * (b + (size_t)2 * hi * g)];
    free(wj);
  }
  for (int m = 0; m < bl_scoeff; m++)
    for (int l = 0; l < bl_scoeff; l++) scoeff[l + (size_t)bl_scoeff * m] = FLM(l, m);
  return ORC_ERR_NONE;
}

/* ------------------------------------------------------------------------
 * Synthesis, one scale: s2dw_core_synthesis_wav2flm_dynamic body,
 * s2dw_core_mod.f90:1090-1253 (same numerics as the static 847-1009).
 * Adds the scale's contribution into flm.
 * ---------------------------------------------------------------------- */
static int orc_synthesis_scale(cplx *flm, const double *wj, const double *K_gamma, const cplx *Slm, int jj, int J, int B,
                               int N, double alpha) {
  int bl_hi, bl_lo;
  orc_band_limits(jj, B, alpha, &bl_hi, &bl_lo);
  const int hi = bl_hi, na = 2 * hi - 1;
  /* V(0:hi-1, 0:2hi-1, 0:N-1) */
  cplx *V = (cplx *)calloc((size_t)hi * 2 * hi * N, sizeof(cplx));
  if (!V) return ORC_ERR_MEM_ALLOC_FAIL;
#define VV(m, x, g) V[(m) + (size_t)hi * ((x) + (size_t)2 * hi * (g))]
  /* Step 1, 1104-1130 */
  orc_fftplan p;
  orc_fftplan_init(&p, na);
#pragma omp parallel num_threads(orc_get_threads())
  {
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)na);
#pragma omp for collapse(2) schedule(dynamic, 10)
    for (int gg = 0; gg <= N - 1; gg++) {
      for (int bb = 0; bb <= 2 * hi - 1; bb++) {
        orc_r2c(&p, &wj[(size_t)na * (bb + (size_t)2 * hi * gg)], &VV(0, bb, gg), work);
        for (int m = 0; m < hi; m++) VV(m, bb, gg) = VV(m, bb, gg) * (4.0 * (ORC_PI * ORC_PI) / (double)N) / (2.0 * hi - 1.0);
      }
    }
    free(work);
  }
  orc_fftplan_free(&p);
  /* Step 2a, 1132-1155: Umnb(0:hi-1, 0:N-1, 0:2hi-1).  The reference sums
   * over gg with an OpenMP reduction; here each (mmm) row sums gg in order. */
  cplx *Umnb = (cplx *)calloc((size_t)hi * N * 2 * hi, sizeof(cplx));
#define UU(m, n, b) Umnb[(m) + (size_t)hi * ((n) + (size_t)N * (b))]
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
  for (int mmm = 0; mmm <= N - 1; mmm++) {
    int k_indicator = (1 - (((N + mmm) % 2 != 0) ? -1 : 1)) / 2;
    if (k_indicator > 0) {
      for (int gg = 0; gg <= N - 1; gg++) {
        double gamma_g = ORC_PI * gg / (double)N;
        cplx e = cexp(-I * (double)mmm * gamma_g);
        for (int b = 0; b < 2 * hi; b++)
          for (int m = 0; m < hi; m++) UU(m, mmm, b) += VV(m, b, gg) * e;
      }
    }
  }
  /* Step 2b, 1157-1179: V(:, mm+hi-1, mmm) = sum_b U * w * exp(-i mm beta_b),
   * written back into V reinterpreted as (0:hi-1, 0:2hi-2, 0:N-1).  The
   * reference's reduction(+:V) over bb is replaced by a per-(mmm, mm) sum over
   * bb in order. */
  for (int mmm = 0; mmm < N; mmm++)
    for (int x = 0; x < 2 * hi - 1; x++)
      for (int m = 0; m < hi; m++) VV(m, x, mmm) = 0;
  double *wb = (double *)malloc(sizeof(double) * 2 * hi);
  for (int bb = 0; bb < 2 * hi; bb++) wb[bb] = orc_quad_weights(ORC_PI * (2.0 * bb + 1.0) / (4.0 * hi), hi);
#pragma omp parallel for collapse(2) schedule(dynamic, 10) num_threads(orc_get_threads())
  for (int mmm = 0; mmm <= N - 1; mmm++) {
    for (int mm = -(hi - 1); mm <= hi - 1; mm++) {
      int k_indicator = (1 - (((N + mmm) % 2 != 0) ? -1 : 1)) / 2;
      if (k_indicator > 0) {
        for (int bb = 0; bb <= 2 * hi - 1; bb++) {
          double beta_b = ORC_PI * (2.0 * bb + 1.0) / (4.0 * hi);
          double w = wb[bb];
          cplx e = cexp(-I * (double)mm * beta_b);
          for (int m = 0; m < hi; m++) VV(m, mm + hi - 1, mmm) += UU(m, mmm, bb) * w * e;
        }
      }
    }
  }
  free(wb);
  free(Umnb);
  /* Step 3a, 1181-1198 */
  for (int mmm = 0; mmm <= N - 1; mmm++) {
    int k_indicator = (1 - (((N + mmm) % 2 != 0) ? -1 : 1)) / 2;
    if (k_indicator <= 0) continue;
    for (int mm = 1; mm <= hi - 1; mm++)
      for (int m = 0; m <= hi - 1; m++) {
        double sg = ((m + mmm) % 2) ? -1.0 : 1.0;
        VV(m, mm + hi - 1, mmm) = VV(m, mm + hi - 1, mmm) + sg * VV(m, -mm + hi - 1, mmm);
      }
  }
  /* Step 3b, 1200-1226: wig(0:hi-1, 0:hi-1, 0:N-1) */
  cplx *wig = (cplx *)calloc((size_t)hi * hi * N, sizeof(cplx));
#define WIG(l, m, n) wig[(l) + (size_t)hi * ((m) + (size_t)hi * (n))]
#pragma omp parallel num_threads(orc_get_threads())
  {
    double *dlbuf = (double *)malloc(sizeof(double) * (size_t)(2 * hi + 1) * (2 * hi + 1));
#pragma omp for schedule(dynamic, 10)
    for (int el = bl_lo + 1; el <= bl_hi - 1; el++) {
      orc_dl_beta_operator(dlbuf, ORC_PION2, el);
      const int ld = 2 * el + 1;
#define DL(a, b) dlbuf[((a) + el) + (size_t)ld * ((b) + el)]
      int nmax = N - 1 < el ? N - 1 : el;
      for (int mmm = 0; mmm <= nmax; mmm++) {
        for (int m = 0; m <= el; m++) {
          int k_indicator = (1 - (((N + mmm) % 2 != 0) ? -1 : 1)) / 2;
          if (k_indicator > 0) {
            cplx s = 0;
            for (int mm = 0; mm <= el; mm++) s += VV(m, mm + hi - 1, mmm) * DL(mm, m) * DL(mm, mmm);
            WIG(el, m, mmm) = s * (orc_exact_phase ? conj(orc_ipow(mmm - m)) : cexp(-I * (double)(mmm - m) * ORC_PION2));
          }
        }
      }
#undef DL
    }
    free(dlbuf);
  }
  free(V);
  /* flm accumulate, 1228-1253 */
#pragma omp parallel for schedule(dynamic, 10) num_threads(orc_get_threads())
  for (int el = bl_lo + 1; el <= bl_hi - 1; el++) {
    int nmax = N - 1 < el ? N - 1 : el;
    for (int mmm = 0; mmm <= nmax; mmm++) {
      for (int m = 0; m <= el; m++) {
        if (mmm == 0) {
          FLM(el, m) = FLM(el, m) + ((double)(2.0 * el + 1) / (8.0 * (ORC_PI * ORC_PI))) * WIG(el, m, 0) * KG(jj, el) * SLM(el, 0);
        } else {
          int k_indicator = (1 - (((N + mmm) % 2 != 0) ? -1 : 1)) / 2;
          if (k_indicator > 0)
            FLM(el, m) = FLM(el, m) + 2.0 * ((double)(2.0 * el + 1) / (8 * (ORC_PI * ORC_PI))) * WIG(el, m, mmm) *
                                          KG(jj, el) * SLM(el, mmm);
        }
      }
    }
  }
  free(wig);
#undef VV
#undef UU
#undef WIG
  return ORC_ERR_NONE;
}

/* s2dw_core_synthesis_wav2flm_dynamic, s2dw_core_mod.f90:1060-1266 (the
 * caller frees wavdyn afterwards, as the Fortran routine does at 1130/1264). */
int orc_synthesis_dynamic(cplx *flm, double *const *wavdyn, const cplx *scoeff, const double *K_gamma, const double *Phi2,
                          const cplx *Slm, int J, int B, int N, int bl_scoeff, double alpha) {
  for (size_t k = 0; k < (size_t)B * B; k++) flm[k] = 0;
  for (int jj = 0; jj <= J; jj++) {
    int e = orc_synthesis_scale(flm, wavdyn[jj], K_gamma, Slm, jj, J, B, N, alpha);
    if (e) return e;
  }
  /* scaling part, 1257-1261 */
  for (int el = 0; el <= bl_scoeff - 1; el++)
    for (int m = 0; m <= bl_scoeff - 1; m++) FLM(el, m) = FLM(el, m) + Phi2[el] * scoeff[el + (size_t)bl_scoeff * m];
  return ORC_ERR_NONE;
}

/* s2dw_core_synthesis_wav2flm (static layout), s2dw_core_mod.f90:816-1019 */
int orc_synthesis_static(cplx *flm, const double *wav, const cplx *scoeff, const double *K_gamma, const double *Phi2,
                         const cplx *Slm, int J, int B, int N, int bl_scoeff, double alpha) {
  for (size_t k = 0; k < (size_t)B * B; k++) flm[k] = 0;
  for (int jj = 0; jj <= J; jj++) {
    int hi, lo;
    orc_band_limits(jj, B, alpha, &hi, &lo);
    double *wj = (double *)malloc(sizeof(double) * (size_t)(2 * hi - 1) * 2 * hi * N);
    if (!wj) return ORC_ERR_MEM_ALLOC_FAIL;
    for (int g = 0; g < N; g++)
      for (int b = 0; b < 2 * hi; b++)
        for (int a = 0; a < 2 * hi - 1; a++)
          wj[a + (size_t)(2 * hi - 1) * (b + (size_t)2 * hi * g)] =
              wav[jj + (size_t)(J + 1) * (a + (size_t)(2 * B - 1) * (b + (size_t)2 * B * g))];
    int e = orc_synthesis_scale(flm, wj, K_gamma, Slm, jj, J, B, N, alpha);
    free(wj);
    if (e) return e;
  }
  for (int el = 0; el <= bl_scoeff - 1; el++)
    for (int m = 0; m <= bl_scoeff - 1; m++) FLM(el, m) = FLM(el, m) + Phi2[el] * scoeff[el + (size_t)bl_scoeff * m];
  return ORC_ERR_NONE;
}
#undef FLM
#undef KG
#undef SLM

/* ran2_dp, s2dw_test.f90:247-285 (Numerical Recipes ran2 as compiled there:
 * AM, EPS, RNMX are default-real literals).  State is explicit here. */
typedef struct {
  int idum, idum2, iy;
  int iv[32];
} orc_ran2_state;

void orc_ran2_init(orc_ran2_state *st, int seed) {
  memset(st, 0, sizeof(*st));
  st->idum = seed;
  st->idum2 = 123456789;
}

double orc_ran2_dp(orc_ran2_state *st) {
  const int IM1 = 2147483563, IM2 = 2147483399, IMM1 = IM1 - 1, IA1 = 40014, IA2 = 40692, IQ1 = 53668, IQ2 = 52774,
            IR1 = 12211, IR2 = 3791, NTAB = 32, NDIV = 1 + IMM1 / NTAB;
  const double AM = (double)(1.f / (float)IM1), EPS = (double)1.2e-7f, RNMX = (double)(1.f - 1.2e-7f);
  int k, j;
  (void)EPS;
  if (st->idum <= 0) {
    st->idum = -st->idum > 1 ? -st->idum : 1;
    st->idum2 = st->idum;
    for (j = NTAB + 8; j >= 1; j--) {
      k = st->idum / IQ1;
      st->idum = IA1 * (st->idum - k * IQ1) - k * IR1;
      if (st->idum < 0) st->idum += IM1;
      if (j <= NTAB) st->iv[j - 1] = st->idum;
    }
    st->iy = st->iv[0];
  }
  k = st->idum / IQ1;
  st->idum = IA1 * (st->idum - k * IQ1) - k * IR1;
  if (st->idum < 0) st->idum += IM1;
  k = st->idum2 / IQ2;
  st->idum2 = IA2 * (st->idum2 - k * IQ2) - k * IR2;
  if (st->idum2 < 0) st->idum2 += IM2;
  j = 1 + st->iy / NDIV;
  st->iy = st->iv[j - 1] - st->idum2;
  st->iv[j - 1] = st->idum;
  if (st->iy < 1) st->iy += IMM1;
  double r = AM * st->iy;
  return r < RNMX ? r : RNMX;
}

/* s2dw_test_gen_flm, s2dw_test.f90:192-224 (real part drawn first). */
void orc_test_gen_flm(int B, cplx *flm, int seed) {
  orc_ran2_state st;
  orc_ran2_init(&st, seed);
  for (size_t k = 0; k < (size_t)B * B; k++) flm[k] = 0;
  for (int el = 0; el <= B - 1; el++) {
    flm[el] = 2.0 * orc_ran2_dp(&st) - 1.0;
    for (int m = 1; m <= el; m++) {
      double re = 2.0 * orc_ran2_dp(&st) - 1.0;
      double im = 2.0 * orc_ran2_dp(&st) - 1.0;
      flm[el + (size_t)B * m] = re + I * im;
    }
  }
}

/* s2dw_stat_moments, s2dw_stat_mod.f90:81-275, one scale after another (the
 * reference splits the scales over OpenMP threads).  wavdyn[jj] holds
 * (2hi-1)*2hi*N values (band limits as 108-119) and is multiplied by mask[jj]
 * in place when mask is not NULL (136-140).  Every sum runs in array order
 * over the entries whose mask is 1 within TOL_ZERO (153-197), or over all
 * entries without a mask (201-242).  five and six may be NULL (the optional
 * outputs).  Returns ORC_ERR_SIZE_INVALID (a halting error in the reference)
 * when nj < 2 (143-150), ORC_ERR_ARTH_WARNING when some variance is below
 * TOL_ZERO (results complete, 251-260). */
int orc_stat_moments(double *const *wavdyn, int J, int B, int N, double alpha, double *mean, double *var,
                     double *skew, double *kur, double *const *mask, double *five, double *six) {
  int warn = 0;
  for (int jj = 0; jj <= J; jj++) {
    int hi, lo;
    orc_band_limits(jj, B, alpha, &hi, &lo);
    const size_t cnt = (size_t)(2 * hi - 1) * (size_t)(2 * hi) * (size_t)N;
    double *w = wavdyn[jj];
    const double *mk = mask ? mask[jj] : NULL;
    if (mk)
      for (size_t i = 0; i < cnt; i++) w[i] = w[i] * mk[i];
    long nj;
    if (mk) {
      double sm = 0;
      for (size_t i = 0; i < cnt; i++) sm += mk[i];
      nj = lround(sm); /* nint: halves away from zero */
    } else {
      nj = (long)cnt;
    }
    if (nj < 2) return ORC_ERR_SIZE_INVALID;
#define ORC_IN(i) (!mk || fabs(mk[i] - 1.0) < ORC_TOL_ZERO)
    double sum = 0;
    for (size_t i = 0; i < cnt; i++)
      if (ORC_IN(i)) sum += w[i];
    const double mu = sum / nj;
    double ep = 0, v = 0, sk = 0, ku = 0, f5 = 0, s6 = 0;
    for (size_t i = 0; i < cnt; i++)
      if (ORC_IN(i)) ep += w[i] - mu;
    for (size_t i = 0; i < cnt; i++)
      if (ORC_IN(i)) {
        const double s1 = w[i] - mu;
        v += s1 * s1;
      }
    for (size_t i = 0; i < cnt; i++)
      if (ORC_IN(i)) {
        const double s1 = w[i] - mu;
        sk += s1 * s1 * s1;
      }
    for (size_t i = 0; i < cnt; i++)
      if (ORC_IN(i)) {
        const double s1 = w[i] - mu;
        ku += s1 * s1 * s1 * s1;
      }
    for (size_t i = 0; i < cnt; i++)
      if (ORC_IN(i)) {
        const double s1 = w[i] - mu;
        f5 += s1 * s1 * s1 * s1 * s1;
      }
    for (size_t i = 0; i < cnt; i++)
      if (ORC_IN(i)) {
        const double s1 = w[i] - mu;
        s6 += s1 * s1 * s1 * s1 * s1 * s1;
      }
#undef ORC_IN
    v = (v - ep * ep / (double)nj) / (double)(nj - 1);
    const double sd = sqrt(v);
    mean[jj] = mu;
    var[jj] = v;
    if (v < ORC_TOL_ZERO) {
      warn = 1;
      skew[jj] = 0;
      kur[jj] = 0;
      if (five) five[jj] = 0;
      if (six) six[jj] = 0;
    } else {
      skew[jj] = sk / (nj * (sd * sd * sd));
      kur[jj] = ku / (nj * (v * v)) - 3.0;
      if (five) five[jj] = f5 / (nj * (sd * sd * sd * sd * sd));
      if (six) six[jj] = s6 / (nj * (v * v * v)) - 15.0;
    }
  }
  return warn ? ORC_ERR_ARTH_WARNING : ORC_ERR_NONE;
}
