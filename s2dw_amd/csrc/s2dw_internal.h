// s2dw_internal.h -- declarations shared by the host and device sources of
// libs2dw_hip.so.  Not part of the public C ABI (include/s2dw_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace s2dw {

// Error codes, numbering of s2dw_error_mod.f90:37-52.
enum {
  kErrNone = 0,
  kErrInit = 1,
  kErrNotInit = 2,
  kErrInitFail = 3,
  kErrMemAlloc = 4,
  kErrArth = 5,
  kErrArthWarning = 6,
  kErrSizeInvalid = 8,
  kErrArgInvalid = 10,
  kErrQuadStepExceed = 12,
  kErrAdmissFail = 13,
};

// Host setup (s2dw_setup.cpp).
double powi(double a, int b);
bool assimilate_int(double x, int *y);
int comp_Jmax(int B, double alpha);
bool params_valid(int J, int B, double alpha, int N);
void band_limits(int jj, int B, double alpha, int *hi, int *lo);
int init_kernels(double *K, double *Phi2, int *bl_scoeff, int J, int B, double alpha);
int init_directionality(double *S, int B, int N);
bool admiss(double *a, const double *K, const double *Phi2, int B, int J);
double quad_weight(double theta, int B);

// Error text for the calling thread (s2dw_capi.cpp).
void set_error(const char *fmt, ...);
int hip_error(hipError_t e, const char *what);
void host_copy_mt(void *dst, const void *src, size_t n);  // memcpy over OpenMP threads (large copies)

#define S2DW_HIP_CHECK(call)                                  \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return ::s2dw::hip_error(e_, #call); \
  } while (0)

// ---------------------------------------------------------------------------
// FFT description passed by value to kernels.  A length-n transform is run
// either directly as mixed-radix Stockham stages (n = product of radices in
// {2,3,4,5,7,8}) or, when n has a larger prime factor, by Bluestein's chirp-z
// convolution on a power-of-two length M >= 2n-1.
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 16;
struct FftSpec {
  int n;          // logical length
  int M;          // work length (n, or the Bluestein power of two)
  int bluestein;  // 0/1
  int nst;        // number of Stockham stages for M
  int radix[kMaxStages];
  const double2 *tw;     // M entries, e^{+2 pi i k / M}
  const double2 *chirp;  // M entries, e^{+i pi (k^2 mod 2n) / n} for k < n, 0 beyond (Bluestein)
  const double2 *chirp_c; // M entries, conj(chirp)                         (Bluestein)
  const double2 *filt_b; // M entries, FFT_M(conj chirp) / M  (backward transforms)
  const double2 *filt_f; // M entries, FFT_M(chirp) / M       (forward transforms)
};

// Per-scale geometry.
struct Scale {
  int hi, lo;            // bl_hi, bl_lo
  long long wav_off;     // offset (doubles) of this scale in the packed wav
  long long tg_off;      // offset (double2) of Tg_j / Vt_j in the contraction buffer
  long long tb_off;      // offset (double2) of Tb_j / V_j in the transform buffer
  int fa, fb;            // FFT slots: alpha (2hi-1) and beta (2hi)
};

// Columns of the real B matrices of the l-contraction: one per (n, part of
// S(l,n)) that is not identically zero.
struct Column {
  int n;     // azimuthal index (analysis: |n| <= N-1 with N+n odd; synthesis n >= 0)
  int part;  // 0: real part of S, 1: imaginary part
};

// Optional per-stage HIP-event timing (stage ids in include/s2dw_hip.h).
constexpr int kNumStages = 11;
struct StageEvents {
  int stage;
  hipEvent_t a, b;
};

struct Plan {
  int B = 0, N = 0, J = 0, bl_scoeff = 0, precision = 0, device = 0;
  double alpha = 0;
  std::vector<Scale> sc;
  std::vector<Column> cols_a, cols_s;
  int nk = 0;  // number of thin Delta^l_{m'k} chains per m' (k = 0..nk-1)
  long long wav_doubles = 0;
  std::vector<double> K, Phi2, S;  // host copies (reference layouts)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipStream_t aux = nullptr;          // second stream for small concurrent launches
  bool single_stream = false;         // keep every launch on `stream` (callers that run many plans at once)
  hipEvent_t fork = nullptr, join = nullptr;
  // device tables
  double *d_S0 = nullptr;     // chain start values, triangle [l0][k]
  double *d_thin = nullptr;   // Delta^l_{m'k}, [tri(m',l)][k]
  double *d_thinneg = nullptr;  // Delta^l_{m',-k}, [tri(m',l)][k] (table form only)
  // table form of the plain transform: Wigner planes with the reference's
  // rounding, [off(l) + m'(l+1) + m] (s2dw_planes.hip), read by the MFMA
  // contraction kernels instead of the device recursions
  double *d_Sa = nullptr;
  long long sa_doubles = 0;
  // split-precision analysis contraction (S2DW_PREC_SPLIT): digits of the
  // analysis columns [rowA(m',l)][16] and the per-transform row exponents
  uint4 *d_Bq = nullptr;
  long long bq_bytes = 0;
  int *d_erow = nullptr;
  // reference-rounding form (table planes plus the reference's own phase
  // factors and beta-sum basis values, s2dw_refround.hip)
  bool refround = false;
  int form_req = -1;           // form asked for by the host-buffer entry points' cache
  bool tab_syn = false;        // the synthesis contraction reads the Wigner table too
  double2 *d_Gref = nullptr;   // analysis gamma matrix as the reference rounds it, [g][c]
  double2 *d_phref = nullptr;  // e^{i k pi/2} as the reference rounds it, k = n - m + phoff
  int phoff = 0;
  double *d_eta = nullptr;     // per scale: Ec[b][m'], Es[b][m'] (2hi x hi each)
  std::vector<long long> eta_off;
  double2 *d_cs = nullptr;     // sine-GEMM result of one scale
  void *blas = nullptr;        // rocblas_handle
  double *d_Ba = nullptr;     // analysis rows [tri(m',l)][2 + 2*NCa]
  long long ba_bytes = 0;     // its size in bytes
  double *d_Bs = nullptr;     // synthesis rows [tri(l,m')][2 + 2*NCs]
  int *d_lmeta = nullptr;     // per l: flush info (int2), then per group of 4 l: end bits
  int *d_gev = nullptr;       // inside d_lmeta
  double *d_lcoef = nullptr;  // l-recursion coefficients [m'][l] (A, Bc)
  double *d_uinv = nullptr;   // [l][m] 1/sqrt((l+1)^2 - m^2) (0 where not positive)
  int4 *d_sblk = nullptr;     // synthesis l-blocks (l_first, count, slot-0 scale, slot-1 scale)
  int n_sblk = 0;
  double2 *d_G = nullptr;     // analysis gamma matrix [g][c]
  double *d_wq = nullptr;     // Driscoll-Healy weights, per scale 2hi (offset = tb-like)
  long long *d_wq_off = nullptr;
  std::vector<long long> wq_off;
  std::vector<FftSpec> ffts;  // distinct lengths
  std::vector<void *> allocs; // every device allocation, freed on destroy
  // scratch
  double2 *d_fT = nullptr;    // [l][m]
  double2 *d_tg = nullptr;    // contraction output / synthesis Vt
  double2 *d_tb = nullptr;    // beta/alpha intermediate
  double2 *d_fo = nullptr;    // synthesis output [l][m]
  double2 *d_tw8 = nullptr;   // per scale e^{2 pi i k/(8hi)} tables
  std::vector<long long> tw8_off;
  double2 *eg = nullptr;      // synthesis gamma table [n-index][g]
  std::vector<int> nlist;     // synthesis n values (n >= 0, N+n odd)
  unsigned char *d_cov = nullptr;  // per l: covered by some non-empty scale
  double *d_Phi2 = nullptr;
  int *d_his = nullptr;       // per scale bl_hi
  long long *d_woffs = nullptr;  // per scale packed wav offset
  int max_lds = 65536;         // LDS bytes a workgroup may use on this device
  long long tg_doubles = 0;    // size of d_tg in doubles
  long long device_bytes = 0;
  bool timing = false;
  std::vector<StageEvents> ev_live;  // recorded, not yet read
  std::vector<hipEvent_t> ev_pool;   // free events
  int syn_q = 4;              // chains per lane in the VALU synthesis contraction
  bool syn_mfma = true;       // synthesis contraction on FP64 MFMA
  bool syn_wave = true;       // ... with independent waves (else a V tile staged in LDS)
  int smf_NCp = 0, smf_CW = 0, smf_nch = 0, n_swork = 0;
  int smf_wcount[3] = {0, 0, 0};  // work items per k-step class (<=2, <=4, <=8)
  bool smf_side_small = false;     // classes 1, 2 read only scales j >= 2
  std::vector<int> smf_lb, smf_ub;
  std::vector<long long> smf_btoff;
  double *d_BT = nullptr;     // per-scale synthesis B tables
  long long bt_doubles = 0;
  double2 *d_cc = nullptr;    // m'-recursion coefficients [m'][l]
  double *d_cp = nullptr;     // checkpoint planes [chunk][2][m][l]
  double2 *d_fpart = nullptr; // chunk partial sums [chunk][m][l]
  int4 *d_swork = nullptr;    // MFMA synthesis work list
  int ana_p = 2;              // chains per lane in the VALU analysis contraction
  bool ana_mfma = true;       // analysis contraction on FP64 MFMA
  bool ana_wave_indep = true; // ... with independent waves (else workgroup-staged operands)
  bool ana_mfma_ok = true, syn_mfma_ok = true;  // the MFMA contractions are available on this plan
  int fft_engine = 0;         // 0 workgroup engine, 1 wave engine, 2 generic Stockham
  bool odd_alpha = true;      // full-band alpha stages of 2hi - 1 = 1023 on the odd-length engine
  bool beta_one = true;       // full-band beta stages with one FFT per column (k_beta_*_h2)
  bool syn_xcd = true;        // synthesis contraction: consecutive work items share an XCD's L2
  int syn_cw_min = 8;         // synthesis contraction: smallest m' chunk width
  double2 *d_otw = nullptr;   // its twiddles e^{+2 pi i k/1023} (set when that engine runs)
  // HIP graphs of whole transforms for the device API: one executable graph
  // per (direction, buffers), captured on `cap` at first use and replayed
  // on `stream` afterwards
  struct GraphEntry {
    int kind;
    const void *b0, *b1, *b2;
    hipGraphExec_t exec;
    unsigned long long last;
  };
  void *groups = nullptr;  // grouped launches of the small scales (kernels' own state)
  bool graphs = false;
  hipStream_t cap = nullptr;
  double *d_mom = nullptr;    // moments partials of s2dw_hip_stat_moments_device
  long long mom_doubles = 0;
  // analysis reducing to moments (s2dw_hip_analysis_moments_device): one
  // record per alpha-c2r workgroup, scale j owning records momf_rec0[j] ..
  // momf_rec0[j+1]-1; available when every non-empty scale runs in the
  // grouped or paired launches
  bool momf_ok = false;
  std::vector<long long> momf_rec0;
  double *d_momf = nullptr;
  std::vector<GraphEntry> gcache;
  unsigned long long gtick = 0;
};

int plan_build(Plan *p, int B, int N, int J, double alpha, const double *K, const double *Phi2, const double *S,
               int precision, hipStream_t stream, bool own_stream);
void plan_free(Plan *p);
// d_mom != NULL: the alpha-c2r stage reduces to moments records instead of
// writing d_wav (which may be NULL), and d_mom receives the per-scale moments
// (rows as moments_run's d_out)
int run_analysis(Plan *p, double *d_wav, double2 *d_scoeff, const double2 *d_flm, double *d_mom = nullptr);
int run_synthesis(Plan *p, double2 *d_flm, const double *d_wav, const double2 *d_scoeff);
// run_analysis (kind 0; kind 2: reducing to moments, out0 = d_mom) or
// run_synthesis (kind 1) through the plan's graph
// cache when graphs are on and stage timing is off, eagerly otherwise
int run_graphed(Plan *p, int kind, void *out0, void *out1, const void *in);
int plan_autotune(Plan *p, int reps, std::string *report);
void plan_drop_graphs(Plan *p);
int stage_times(Plan *p, double *ms, int *count);
int pack_static(Plan *p, const double *d_packed, double *d_static, bool to_static);
// Moments of packed per-scale coefficients (s2dw_stats.hip): scale s holds
// count[s] (even) doubles at d_wav + off[s]; d_mask (same packing) or NULL;
// d_part holds moments_partial_doubles(count, nscales) doubles; d_out gets
// 8 rows of nscales values (mean, var, skew, kur, five, six, nj, flag).
long long moments_partial_doubles(const long long *count, int nscales);
int moments_run(hipStream_t st, const long long *off, const long long *count, int nscales, double *d_wav,
                const double *d_mask, double *d_part, double *d_out);
// Final merge of moments records already in d_part: scale s (count[s]
// values) owns records rec0[s] .. rec0[s+1]-1 (8 doubles each).
int moments_finish(hipStream_t st, const long long *count, const long long *rec0, int nscales, const double *d_part,
                   double *d_out);
// Moments of packed coefficients of (J, B, N, alpha) with the records in the
// library's static device scratch (no allocation).
int moments_noalloc(double *d_wav, const double *d_mask, int J, int B, int N, double alpha, double *d_out,
                    hipStream_t st);
int wigner_plane(int l, double *out, double *out2);
// Reference-rounding Wigner planes (s2dw_planes.hip): doubles of the table for
// band limit L, and its construction into d_Sa with the exact thin columns.
long long ref_plane_doubles(int L);
int build_ref_planes(int L, int nk, double *d_Sa, double *d_thin, double *d_thinneg, hipStream_t st);
// Reference-rounding beta sum of the analysis (s2dw_refround.hip).
int refbeta_build(Plan *p);
void refbeta_free(Plan *p);
int refbeta_run(Plan *p, hipStream_t st);

}  // namespace s2dw
