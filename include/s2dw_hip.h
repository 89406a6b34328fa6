/*
 * s2dw_hip.h -- C ABI of libs2dw_hip.so, the MI355X (gfx950) implementation of
 * the S2DW scale-discretised directional wavelet transform on the sphere.
 *
 * Every entry point replaces one public procedure of the reference's Fortran
 * module s2dw_core_mod (src/mod/s2dw_core_mod.f90, public list at 29-39) and
 * keeps its argument order, meaning and array layout, so that a Fortran
 * module with the same procedure names (s2dw_amd/fortran/s2dw_core_mod.f90)
 * can forward to it through ISO_C_BINDING.
 *
 * Conventions (all identical to the reference):
 *  - B is the harmonic band limit (0 <= l < B), N the azimuthal band limit,
 *    J the number of scales minus one (scales jj = 0..J), alpha the dilation.
 *  - Arrays are Fortran column-major, first index fastest.  complex(dpc)
 *    arrays are interleaved (re, im) doubles:
 *      flm(0:B-1,0:B-1)        flm[2*(l + B*m) + {0,1}]
 *      K_gamma(0:J,0:B-1)      K[j + (J+1)*l]
 *      Phi2(0:B-1)             Phi2[l]
 *      Slm(0:B-1,0:N-1)        S[2*(l + B*n) + {0,1}]
 *      scoeff(0:bl-1,0:bl-1)   sc[2*(l + bl*m) + {0,1}]
 *      wav(0:J,0:2B-2,0:2B-1,0:N-1) (static)   wav[j + (J+1)*(a + (2B-1)*(b + 2B*g))]
 *      wavdyn(jj)%coeff(0:2hi-2,0:2hi-1,0:N-1) (dynamic, one array per scale)
 *                              w[a + (2hi-1)*(b + 2hi*g)]
 *  - Integers are Fortran default integers (int32), alpha is real(dp).
 *  - Every function returns an S2DW error code (numbering of
 *    s2dw_error_mod.f90:37-52); 0 means success.  The reference halts inside
 *    the failing routine instead (s2dw_error, error_mod.f90:127-191); the
 *    Fortran module re-raises the returned code through s2dw_error, which
 *    reproduces the message and the halt.
 *  - The transform entry points take host pointers, copy to the GPU, run and
 *    copy back before returning (synchronous, like the Fortran routines).
 *    The *_device entry points take device pointers and a HIP stream and
 *    return after enqueueing.
 *  - Thread safety: calls on distinct plans may run concurrently from
 *    different host threads; calls on one plan must not overlap.
 */
#ifndef S2DW_HIP_H
#define S2DW_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes, s2dw_error_mod.f90:37-52. */
#define S2DW_ERROR_NONE 0
#define S2DW_ERROR_INIT 1
#define S2DW_ERROR_NOT_INIT 2
#define S2DW_ERROR_INIT_FAIL 3
#define S2DW_ERROR_MEM_ALLOC_FAIL 4
#define S2DW_ERROR_ARTH 5
#define S2DW_ERROR_ARTH_WARNING 6
#define S2DW_ERROR_SIZE_WARNING 7
#define S2DW_ERROR_SIZE_INVALID 8
#define S2DW_ERROR_SIZE_NOT_DEF 9
#define S2DW_ERROR_ARG_INVALID 10
#define S2DW_ERROR_ARG_WARNING 11
#define S2DW_ERROR_QUAD_STEP_EXCEED 12
#define S2DW_ERROR_ADMISS_FAIL 13
#define S2DW_ERROR_FILEIO 14

/* Precision of the l-contraction (the rest of the path is always FP64). */
#define S2DW_PREC_FP64 0  /* plain form: everything in FP64 */
#define S2DW_PREC_SPLIT 1 /* split-precision contraction on MFMA (see plan docs) */

/* Rounding of the FP64 plain form (OR-ed into the precision argument of
 * s2dw_hip_plan_create).  The reference's own arithmetic rounds a few
 * quantities coarsely: its Wigner planes (row factors from running sums of
 * logarithms, s2dw_dl_mod.f90:139-146, 216-228), the phase e^{i(n-m)pi/2}
 * (core_mod.f90:666, exp of a rounded product) and the basis of its direct
 * beta sum (cos/sin of the rounded product m' beta_b, core_mod.f90:720-731).
 * At B = 256 with white N(0,1) input each of these moves the wavelet
 * coefficients by 1e-12 to 6e-12 (values up to ~70); with a red spectrum the
 * moves are ~1e-14.
 * S2DW_WIGNER_STREAMED (default, 0): Wigner values from exact-start
 * recursions in registers, no plane stored; exact phases; beta sums as FFTs.
 * The fastest form.
 * S2DW_WIGNER_TABLE (16): Wigner planes with the reference's rounding --
 * the values s2dw_dl_beta_operator produces, bit for bit -- generated on the
 * device once per plan (8 * B(B+1)(2B+1)/6 bytes, 359 MB at B = 512) and read
 * by the contraction kernels.
 * S2DW_ROUNDING_REFERENCE (48): the table planes plus the reference's rounded
 * phases and beta-sum basis (the beta sums as FP64 GEMMs with its cos/sin
 * values instead of FFTs); the wavelet coefficients land within the FP64
 * rounding of the sums of the reference's output.  The host-buffer entry
 * points (the Fortran drop-in) use this form.  Runs eagerly (no graph replay).
 * The table and reference forms need N <= 8 (at most 8 analysis columns) and
 * the MFMA contraction kernels; other plans run the streamed form
 * (s2dw_hip_plan_wigner_form tells which). */
#define S2DW_WIGNER_STREAMED 0
#define S2DW_WIGNER_TABLE 16
#define S2DW_ROUNDING_REFERENCE 48

/* Text of the last error raised on the calling thread ("" if none). */
const char *s2dw_hip_last_error(void);

/* Library version string. */
const char *s2dw_hip_version(void);

/* ---------------------------------------------------------------------
 * Setup routines (host, FP64).  These run on the CPU: they are once-per-
 * configuration quadratures whose results are inputs at the boundary.
 * ------------------------------------------------------------------- */

/* s2dw_core_init_kernels, s2dw_core_mod.f90:97-157.  Writes K_gamma
 * ((J+1)*B), Phi2 (B) and *bl_scoeff. */
int s2dw_hip_init_kernels(double *K_gamma, double *Phi2, int *bl_scoeff, int J, int B, double alpha);

/* s2dw_core_init_directionality, s2dw_core_mod.f90:219-250.  Writes Slm
 * (B*N complex). */
int s2dw_hip_init_directionality(double *Slm, int B, int N);

/* s2dw_core_comp_Jmax, s2dw_core_mod.f90:1365-1376.  Returns J_max. */
int s2dw_hip_comp_Jmax(int B, double alpha);

/* s2dw_core_admiss, s2dw_core_mod.f90:285-304.  Writes admiss (B) and
 * *pass (1 if |admiss-1| <= TOL_ADMISS for all l, else 0). */
int s2dw_hip_admiss(double *admiss, const double *K_gamma, const double *Phi2, int B, int J, int *pass);

/* s2dw_core_assimilate_int, s2dw_core_mod.f90:1398-1415.  Returns 1 and
 * sets *y when x is within TOL_CEIL of an integer, else returns 0, *y = 0. */
int s2dw_hip_assimilate_int(double x, int *y);

/* Band limits (bl_hi, bl_lo) of every scale jj = 0..J, exactly as the
 * transform routines compute them (s2dw_core_mod.f90:606-616). */
int s2dw_hip_band_limits(int J, int B, double alpha, int *bl_hi, int *bl_lo);

/* ---------------------------------------------------------------------
 * Transforms, host pointers, synchronous (drop-in for the Fortran API).
 * ------------------------------------------------------------------- */

/* s2dw_core_analysis_flm2wav, s2dw_core_mod.f90:342-534 (static wav). */
int s2dw_hip_analysis_flm2wav(double *wav, double *scoeff, const double *flm, const double *K_gamma,
                              const double *Slm, int J, int B, int N, int bl_scoeff, double alpha);

/* s2dw_core_analysis_flm2wav_dynamic, s2dw_core_mod.f90:572-775.
 * wavdyn[jj] points to caller-allocated (2hi-1)*2hi*N doubles for scale jj
 * (the Fortran module allocates wavdyn(jj)%coeff and passes c_loc). */
int s2dw_hip_analysis_flm2wav_dynamic(double *const *wavdyn, double *scoeff, const double *flm,
                                      const double *K_gamma, const double *Slm, int J, int B, int N,
                                      int bl_scoeff, double alpha);

/* s2dw_core_synthesis_wav2flm, s2dw_core_mod.f90:816-1019 (static wav). */
int s2dw_hip_synthesis_wav2flm(double *flm, const double *wav, const double *scoeff, const double *K_gamma,
                               const double *Phi2, const double *Slm, int J, int B, int N, int bl_scoeff,
                               double alpha);

/* s2dw_core_synthesis_wav2flm_dynamic, s2dw_core_mod.f90:1060-1266.  Reads
 * wavdyn[jj]; freeing the Fortran allocations (1130, 1264) is left to the
 * Fortran module. */
int s2dw_hip_synthesis_wav2flm_dynamic(double *flm, const double *const *wavdyn, const double *scoeff,
                                       const double *K_gamma, const double *Phi2, const double *Slm, int J,
                                       int B, int N, int bl_scoeff, double alpha);

/* ---------------------------------------------------------------------
 * Plans: device-resident transforms.  A plan holds the per-configuration
 * device tables (Wigner chain starts and the thin Delta^l_{m'n} table, FFT
 * twiddles, kernel and directionality weights) and scratch, and runs on
 * one HIP stream.  The reference recomputes these per call; the plan
 * computes them once.
 * ------------------------------------------------------------------- */
typedef struct s2dw_hip_plan *s2dw_hip_plan_t;

/* Create a plan for (B, N, J, alpha) with the caller's K_gamma, Phi2, Slm
 * (layouts as above) and the chosen contraction precision.  All work of the
 * plan is enqueued on `stream` (a hipStream_t; NULL is the null stream).
 * Band limits up to B = 2048 are accepted (above B = 1024 the full-band
 * scales' FFTs run through Bluestein on 8192 points in place, and the plan
 * runs the streamed Wigner form whichever form the flags ask for; split
 * precision is refused with S2DW_ERROR_ARG_INVALID); above
 * that the alpha FFT (length 2B - 1) needs a Bluestein length of 16384,
 * which does not fit in the 160 KiB of LDS per workgroup, and the call
 * returns S2DW_ERROR_SIZE_INVALID. */
int s2dw_hip_plan_create(s2dw_hip_plan_t *plan, int B, int N, int J, double alpha, const double *K_gamma,
                         const double *Phi2, const double *Slm, int precision, void *stream);
int s2dw_hip_plan_destroy(s2dw_hip_plan_t plan);

/* Sizes: number of doubles of the packed wav (all scales, dynamic layout
 * concatenated in jj order), bl_scoeff, and per-scale offsets (J+1 entries,
 * in doubles) and band limits.  Any pointer may be NULL. */
int s2dw_hip_plan_sizes(s2dw_hip_plan_t plan, long long *wav_doubles, int *bl_scoeff, long long *wav_offsets,
                        int *bl_hi, int *bl_lo);

/* Device-memory bytes the plan holds (tables + scratch). */
long long s2dw_hip_plan_device_bytes(s2dw_hip_plan_t plan);

/* Analysis on device pointers: d_flm (B*B complex), d_wav (packed per-scale
 * dynamic layout, see plan_sizes), d_scoeff (bl*bl complex).  Enqueued on
 * the plan's stream; returns without synchronising. */
int s2dw_hip_analysis_device(s2dw_hip_plan_t plan, double *d_wav, double *d_scoeff, const double *d_flm);

/* Synthesis on device pointers (layouts as analysis_device). */
int s2dw_hip_synthesis_device(s2dw_hip_plan_t plan, double *d_flm, const double *d_wav, const double *d_scoeff);

/* Wait for the plan's stream. */
int s2dw_hip_plan_sync(s2dw_hip_plan_t plan);

/* The plan's HIP stream (hipStream_t as void*). */
void *s2dw_hip_plan_stream(s2dw_hip_plan_t plan);

/* Per-stage timing with HIP events on the plan's stream.  Stage ids:
 * 0 analysis f_lm relayout, 1 analysis l-contraction (incl. gamma-DFT),
 * 2 analysis beta FFTs, 3 analysis alpha c2r FFTs, 4 scaling coefficients,
 * 5 synthesis alpha r2c FFTs, 6 synthesis gamma + beta quadrature FFTs,
 * 7 synthesis Wigner contraction, 8 synthesis f_lm assembly, 9 analysis FFTs
 * of the grouped small scales (j >= 2), 10 synthesis FFTs of the grouped
 * small scales (stages 2, 3, 5, 6 then cover the other scales).
 * No counterpart in the reference (which prints wall-clock only,
 * s2dw_test.f90:97-135). */
#define S2DW_NUM_STAGES 11
int s2dw_hip_plan_set_timing(s2dw_hip_plan_t plan, int enable);
/* Graph replay for s2dw_hip_analysis_device / s2dw_hip_synthesis_device
 * (enable != 0): the first call with a given set of buffers captures the
 * whole transform (every kernel of both streams) into a HIP graph, later
 * calls with the same buffers replay it on the plan's stream.  Up to 8
 * graphs per plan (least recently used dropped); disabling drops them all.
 * Results are identical to eager launches.  Stage timing, while on, runs
 * eagerly.  No counterpart in the reference (which has no launch cost to
 * amortise). */
int s2dw_hip_plan_set_graphs(s2dw_hip_plan_t plan, int enable);

/* Autotuner (no counterpart in the reference, which has one code path).
 * Times, on the device, the alternatives a built plan can switch between:
 * analysis contraction (FP64 MFMA with independent waves, MFMA with
 * workgroup-staged operands, VALU with 2 or 1 chains per lane), synthesis
 * contraction (MFMA with independent waves or a staged V tile, VALU with 4 or
 * 8 chains per lane) and the helper stream
 * for the small scales (on/off), each candidate over `reps` transforms on
 * temporary zero-filled buffers, and keeps the fastest.  Results stay within
 * the FP64 parity tolerance whichever path is chosen.  If report is not
 * NULL, up to report_len - 1 characters of the measured times and the
 * choices are written there.  Blocks until done; drops captured graphs. */
int s2dw_hip_plan_autotune(s2dw_hip_plan_t plan, int reps, char *report, int report_len);
/* Sums (ms) and launch counts per stage of everything recorded since the
 * last call; waits for the recorded work.  ms and count hold 11 entries. */
int s2dw_hip_plan_stage_times(s2dw_hip_plan_t plan, double *ms, int *count);

/* ---------------------------------------------------------------------
 * Moments of the wavelet coefficients of each scale: s2dw_stat_moments,
 * s2dw_stat_mod.f90:81-275 (called after an analysis by s2dw_wstats.f90:
 * 181-190).  Per scale jj: mean, variance (n-1 normalised), skewness, excess
 * kurtosis, fifth and excess sixth standardised moments of the
 * (2hi-1)*2hi*N coefficients.  With a mask (same shapes): the coefficients
 * are multiplied by it in place, nj = nint(sum of the mask), and the sums run
 * over the entries whose mask is 1 within TOL_ZERO (stat_mod.f90:136-150).
 * A scale whose variance is below TOL_ZERO gets zero skewness and higher
 * moments (stat_mod.f90:251-260).
 * ------------------------------------------------------------------- */

/* Host pointers, synchronous; arguments in the order of s2dw_stat_moments.
 * wavdyn[jj]: (2hi-1)*2hi*N doubles (masked in place when mask != NULL);
 * mask: NULL or mask[jj] of the same shapes; mean..kur: J+1 doubles each;
 * five, six: J+1 doubles or NULL (the optional outputs).  Returns
 * S2DW_ERROR_SIZE_INVALID if a scale has nj < 2 (the reference halts there:
 * the outputs of the scales before it are filled and var of that scale is
 * NaN), S2DW_ERROR_ARTH_WARNING (results complete) if a variance is below
 * TOL_ZERO (the reference warns once per such scale: var tells which). */
int s2dw_hip_stat_moments(double *const *wavdyn, int J, int B, int N, double alpha, double *mean, double *var,
                          double *skew, double *kur, const double *const *mask, double *five, double *six);

/* Device pointers, enqueued on the plan's stream, no allocation: d_wav is
 * the plan's packed coefficients (as s2dw_hip_analysis_device writes them;
 * masked in place when d_mask != NULL), d_mask NULL or packed alike; both
 * 16-byte aligned.  d_out (device, 8*(J+1) doubles) receives rows
 * d_out[k*(J+1) + jj]: k = 0 mean, 1 var, 2 skew, 3 kur, 4 five, 5 six,
 * 6 nj, 7 flag (0 ok, 1 variance below TOL_ZERO, 2 nj < 2: values NaN). */
int s2dw_hip_stat_moments_device(s2dw_hip_plan_t plan, double *d_wav, const double *d_mask, double *d_out);

/* Analysis followed by the moments, without storing the wavelet coefficients:
 * the analysis of s2dw_core_analysis_flm2wav_dynamic (s2dw_core_mod.f90:
 * 572-775) whose alpha-transform stage reduces each workgroup's coefficients
 * to a moments record instead of writing them, then the per-scale moments of
 * s2dw_stat_moments without a mask (s2dw_stat_mod.f90:81-275), as the
 * s2dw_wstats pipeline computes them (s2dw_wstats.f90:181-190).  d_out as in
 * s2dw_hip_stat_moments_device, d_scoeff and d_flm as in
 * s2dw_hip_analysis_device.  Enqueued on the plan's stream (graph replay
 * when enabled), no allocation.  Returns S2DW_ERROR_ARG_INVALID when
 * s2dw_hip_plan_moments_fusable is 0 for the plan. */
int s2dw_hip_analysis_moments_device(s2dw_hip_plan_t plan, double *d_out, double *d_scoeff, const double *d_flm);
/* No counterpart in the reference: s2dw_hip_stat_moments_device without a
 * plan and without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  d_wav: the packed coefficients of (J, B, N, alpha) in
 * the dynamic per-scale layout (scale jj at the offsets s2dw_hip_plan_sizes
 * gives), masked in place when d_mask != NULL (packed alike); both 16-byte
 * aligned.  d_out: 8*(J+1) doubles, rows as s2dw_hip_stat_moments_device.
 * Enqueued on `stream` (a hipStream_t).  The per-chunk records live in a
 * static device array of 1 MiB (131072 doubles) per GPU, held until the
 * process ends: configurations needing more (e.g. B = 2048, N = 8) and J > 31
 * return S2DW_ERROR_SIZE_INVALID.  Calls that overlap on one GPU share that
 * array and race. */
int s2dw_hip_stat_moments_device_noalloc(double *d_wav, const double *d_mask, int J, int B, int N, double alpha,
                                         double *d_out, void *stream);

/* 1 when every non-empty scale of the plan runs in the grouped or paired
 * FFT launches (all configurations with B <= 1024 on the workgroup engine),
 * which s2dw_hip_analysis_moments_device needs; 0 otherwise. */
int s2dw_hip_plan_moments_fusable(s2dw_hip_plan_t plan);

/* ---------------------------------------------------------------------
 * Request loop (no counterpart in the reference, whose routines run one call
 * at a time from one host thread, s2dw_core_mod.f90).  A server owns a worker
 * thread, a pool of plans and an HBM budget.  Requests of any (B, N, J,
 * alpha) are queued FIFO and admitted while an execution slot is free and
 * their device footprint (a plan of their shape unless an idle one exists,
 * plus their buffers) fits the budget; idle plans are evicted LRU to make
 * room.  Up to min(max_inflight, 4) admitted requests run concurrently,
 * each on a stream (lane) of the server's.  Device buffers are pooled by
 * size; pooled bytes count against the budget and are freed first when room
 * is needed.  Results equal those of the single-call entry points bit for
 * bit.  Host arrays use the layouts of the dynamic entry points:
 * flm[l + B*m], wavelet coefficients packed per scale (jj = 0..J, each
 * (2hi-1)*2hi*N, alpha fastest), scoeff[l + bl*m].  Inputs are copied into
 * pinned staging buffers by the submit call; outputs are copied out by
 * s2dw_hip_server_wait, so output arrays must stay valid until it returns.
 * ------------------------------------------------------------------- */
typedef struct s2dw_hip_server *s2dw_hip_server_t;

int s2dw_hip_server_create(s2dw_hip_server_t *server, long long hbm_budget_bytes, int max_inflight);
/* As s2dw_hip_server_create, with the precision (and rounding form) of every
 * plan the server builds, as for s2dw_hip_plan_create: S2DW_PREC_FP64 (the
 * default of s2dw_hip_server_create) or S2DW_PREC_SPLIT, OR-ed with a
 * S2DW_WIGNER_* / S2DW_ROUNDING_REFERENCE form.  A request the precision
 * cannot serve (S2DW_PREC_SPLIT with N > 8) fails with the plan's error. */
int s2dw_hip_server_create_ex(s2dw_hip_server_t *server, long long hbm_budget_bytes, int max_inflight, int precision);
/* Finishes queued and running requests, then frees everything. */
int s2dw_hip_server_destroy(s2dw_hip_server_t server);

/* Analysis of flm.  wav (may be NULL) receives the packed coefficients,
 * scoeff the scaling coefficients (may be NULL only with keep_resident).
 * keep_resident != 0 keeps wav and scoeff on the device for a later
 * synthesis (wav_from = this ticket) until s2dw_hip_server_release. */
int s2dw_hip_server_submit_analysis(s2dw_hip_server_t server, int B, int N, int J, double alpha, const double *flm,
                                    double *wav, double *scoeff, int keep_resident, long long *ticket);

/* Synthesis into flm, from the resident result of analysis ticket wav_from
 * (>= 0; it may still be queued or running) or from host wav and scoeff. */
int s2dw_hip_server_submit_synthesis(s2dw_hip_server_t server, int B, int N, int J, double alpha, long long wav_from,
                                     const double *wav, const double *scoeff, double *flm, long long *ticket);

/* Blocks until the request has finished; returns its status.  latency_ms:
 * submit to completion; run_ms: admission to completion (either may be NULL). */
int s2dw_hip_server_wait(s2dw_hip_server_t server, long long ticket, double *latency_ms, double *run_ms);

/* Frees the resident result of an analysis ticket (refused while a queued or
 * running synthesis reads it). */
int s2dw_hip_server_release(s2dw_hip_server_t server, long long ticket);

int s2dw_hip_server_stats(s2dw_hip_server_t server, long long *in_use_bytes, int *inflight, int *queued,
                          long long *completed);

/* ---------------------------------------------------------------------
 * Diagnostics used by the tests: Wigner values Delta^l_{m'm}(pi/2) for
 * 0 <= m', m <= l generated by the device recursions of the transform
 * kernels.  out[m' + (l+1)*m] for the analysis (l-recursion) generator,
 * out2 (may be NULL) for the synthesis (m'-recursion) generator.
 * ------------------------------------------------------------------- */
int s2dw_hip_wigner_plane(int l, double *out, double *out2);

/* S2DW_WIGNER_STREAMED, S2DW_WIGNER_TABLE or S2DW_ROUNDING_REFERENCE: the
 * form the plan runs. */
int s2dw_hip_plan_wigner_form(s2dw_hip_plan_t plan);

/* Diagnostic: plane l of a table-form plan, out[m' * (l+1) + m] =
 * Delta^l_{m'm}(pi/2) for 0 <= m', m <= l (synchronous).  Returns
 * S2DW_ERROR_ARG_INVALID for a streamed plan. */
int s2dw_hip_plan_wigner_table(s2dw_hip_plan_t plan, int l, double *out);

#ifdef __cplusplus
}
#endif

#endif /* S2DW_HIP_H */
