// s2dw_fft.h -- LDS-resident batched FFT for gfx950.
//
// A workgroup transforms T sequences of length n held interleaved in LDS
// (element k of sequence t at k*T + t, so consecutive threads touch
// consecutive banks).  Lengths whose prime factors are in {2,3,5,7} run as
// mixed-radix Stockham autosort stages (radix 8/4/2/3/5/7 butterflies in
// registers, twiddles from a per-length table, ping-pong between two LDS
// buffers with one barrier per stage).  Other lengths run through
// Bluestein's chirp-z identity on a power-of-two work length M >= 2n-1: two
// power-of-two transforms and a pointwise product with a precomputed filter.
//
// Sign convention: DIR = +1 is the unnormalised backward transform
// sum_k x_k e^{+2 pi i jk/n}, DIR = -1 the forward transform e^{-...}.
#pragma once

#include <hip/hip_runtime.h>

#include "s2dw_internal.h"

namespace s2dw {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * w (DIR>0) or a * conj(w) (DIR<0)
template <int DIR>
__device__ __forceinline__ double2 cmulw(double2 a, double2 w) {
  if (DIR > 0) return make_double2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
  return make_double2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// multiply by DIR * i
template <int DIR>
__device__ __forceinline__ double2 cmuli(double2 a) {
  return DIR > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <int R>
struct RadixConst;
template <>
struct RadixConst<3> {
  static constexpr double c[2] = {-0.5, -0.5};
  static constexpr double s[2] = {0.86602540378443864676, -0.86602540378443864676};
};
template <>
struct RadixConst<5> {
  // cos/sin(2 pi r q / 5) for r q mod 5 = 1..4
  static constexpr double c[4] = {0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                                  0.30901699437494742410};
  static constexpr double s[4] = {0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                                  -0.95105651629515357212};
};
template <>
struct RadixConst<7> {
  static constexpr double c[6] = {0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624,
                                  -0.90096886790241912624, -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[6] = {0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048,
                                  -0.43388373911755812048, -0.97492791218182360702, -0.78183148246802980871};
};

// cos/sin(2 pi e / R), e = 1..R-1 (entry e-1), to 21 decimals
template <>
struct RadixConst<11> {
  static constexpr double c[10] = {0.841253532831181168862, 0.415415013001886425529, -0.142314838273285140444,
                                   -0.654860733945285064057, -0.959492973614497389890, -0.959492973614497389890,
                                   -0.654860733945285064057, -0.142314838273285140444, 0.415415013001886425529,
                                   0.841253532831181168862};
  static constexpr double s[10] = {0.540640817455597582108, 0.909631995354518371412, 0.989821441880932732376,
                                   0.755749574354258283774, 0.281732556841429697711, -0.281732556841429697711,
                                   -0.755749574354258283774, -0.989821441880932732376, -0.909631995354518371412,
                                   -0.540640817455597582108};
};
template <>
struct RadixConst<31> {
  static constexpr double c[30] = {
      0.979529941252494493938,  0.918957811620230629127,  0.820763441207276326364,  0.688966919075686567801,
      0.528964010326962457365,  0.347305252844820285542,  0.151427777504576663657,  -0.050649168838712712279,
      -0.250652532258720539315, -0.440394151557634309516, -0.612105982547662844147, -0.758758122692790901913,
      -0.874346616144582118827, -0.954139256400048851476, -0.994869323391895146321, -0.994869323391895146321,
      -0.954139256400048851476, -0.874346616144582118827, -0.758758122692790901913, -0.612105982547662844147,
      -0.440394151557634309516, -0.250652532258720539315, -0.050649168838712712279, 0.151427777504576663657,
      0.347305252844820285542,  0.528964010326962457365,  0.688966919075686567801,  0.820763441207276326364,
      0.918957811620230629127,  0.979529941252494493938};
  static constexpr double s[30] = {
      0.201298520088660079142,  0.394355855113318580102,  0.571268215094792279157,  0.724792787229119958865,
      0.848644257494750950464,  0.937752132147080458429,  0.988468324328111399162,  0.998716507171052807146,
      0.968077118866204305153,  0.897804539570741657137,  0.790775736937698582078,  0.651372482722222207454,
      0.485301962531081025215,  0.299363122973357954008,  0.101168321987432177786,  -0.101168321987432177786,
      -0.299363122973357954008, -0.485301962531081025215, -0.651372482722222207454, -0.790775736937698582078,
      -0.897804539570741657137, -0.968077118866204305153, -0.998716507171052807146, -0.988468324328111399162,
      -0.937752132147080458429, -0.848644257494750950464, -0.724792787229119958865, -0.571268215094792279157,
      -0.394355855113318580102, -0.201298520088660079142};
};

// In-register DFT of odd prime length R using the symmetric pairs
// (x_r + x_{R-r}, x_r - x_{R-r}).
template <int R, int DIR>
__device__ __forceinline__ void dft_odd(double2 *v) {
  constexpr int H = (R - 1) / 2;
  double2 ap[H], am[H];
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    ap[r - 1] = cadd(v[r], v[R - r]);
    am[r - 1] = csub(v[r], v[R - r]);
  }
  double2 y[R];
  double2 s0 = v[0];
#pragma unroll
  for (int r = 0; r < H; ++r) s0 = cadd(s0, ap[r]);
  y[0] = s0;
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    double re = v[0].x, im = v[0].y, ire = 0.0, iim = 0.0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int e = (r * q) % R;  // 1..R-1
      const double cc = RadixConst<R>::c[e - 1], ss = RadixConst<R>::s[e - 1];
      re += ap[r - 1].x * cc;
      im += ap[r - 1].y * cc;
      ire += am[r - 1].x * ss;
      iim += am[r - 1].y * ss;
    }
    // y_q = (re,im) + DIR*i*(ire,iim); y_{R-q} = (re,im) - DIR*i*(ire,iim)
    const double2 base = make_double2(re, im);
    const double2 rot = cmuli<DIR>(make_double2(ire, iim));
    y[q] = cadd(base, rot);
    y[R - q] = csub(base, rot);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = y[r];
}

template <int DIR>
__device__ __forceinline__ void dft4(double2 &x0, double2 &x1, double2 &x2, double2 &x3) {
  const double2 a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = cmuli<DIR>(csub(x1, x3));
  x0 = cadd(a, c);
  x2 = csub(a, c);
  x1 = cadd(b, d);
  x3 = csub(b, d);
}

template <int R, int DIR>
__device__ __forceinline__ void dft_small(double2 *v) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    constexpr double h = 0.70710678118654752440;
    double2 a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r] = cadd(v[r], v[r + 4]);
      b[r] = csub(v[r], v[r + 4]);
    }
    // b_r *= w^r, w = e^{DIR 2 pi i / 8}
    b[1] = make_double2(h * (b[1].x - DIR * b[1].y), h * (b[1].y + DIR * b[1].x));
    b[2] = cmuli<DIR>(b[2]);
    b[3] = make_double2(h * (-b[3].x - DIR * b[3].y), h * (-b[3].y + DIR * b[3].x));
    dft4<DIR>(a[0], a[1], a[2], a[3]);
    dft4<DIR>(b[0], b[1], b[2], b[3]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[2 * q] = a[q];
      v[2 * q + 1] = b[q];
    }
  } else {
    dft_odd<R, DIR>(v);
  }
}

// One Stockham stage of radix R over T interleaved sequences of length n.
template <int R, int DIR>
__device__ __forceinline__ void stockham_stage(const double2 *__restrict__ src, double2 *__restrict__ dst, int n, int logT,
                                               int Ns, const double2 *__restrict__ tw) {
  const int T = 1 << logT;
  const int nb = (n / R) << logT;
  const int twstep = n / (Ns * R);
  for (int idx = threadIdx.x; idx < nb; idx += blockDim.x) {
    const int t = idx & (T - 1);
    const int j = idx >> logT;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[((j + r * (n / R)) << logT) + t];
    const int jm = j % Ns;
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmulw<DIR>(v[r], tw[r * jm * twstep]);
    }
    dft_small<R, DIR>(v);
    const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[((idxD + r * Ns) << logT) + t] = v[r];
  }
}

// Runs the Stockham stages of length s.M; returns the buffer holding the result.
template <int DIR>
__device__ double2 *stockham(const FftSpec &s, double2 *a, double2 *b, int logT) {
  int Ns = 1;
  const int n = s.M;
  for (int st = 0; st < s.nst; ++st) {
    const int R = s.radix[st];
    switch (R) {
      case 8: stockham_stage<8, DIR>(a, b, n, logT, Ns, s.tw); break;
      case 4: stockham_stage<4, DIR>(a, b, n, logT, Ns, s.tw); break;
      case 2: stockham_stage<2, DIR>(a, b, n, logT, Ns, s.tw); break;
      case 3: stockham_stage<3, DIR>(a, b, n, logT, Ns, s.tw); break;
      case 5: stockham_stage<5, DIR>(a, b, n, logT, Ns, s.tw); break;
      default: stockham_stage<7, DIR>(a, b, n, logT, Ns, s.tw); break;
    }
    __syncthreads();
    double2 *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// Full length-n transform of the T sequences in A (k < n).  B is the second
// LDS buffer (both hold M*T elements).  Returns the buffer with the result.
// Caller must __syncthreads() after filling A.
template <int DIR>
__device__ double2 *fft_exec(const FftSpec &s, double2 *A, double2 *B, int logT) {
  if (!s.bluestein) return stockham<DIR>(s, A, B, logT);
  const int n = s.n, M = s.M, T = 1 << logT;
  for (int idx = threadIdx.x; idx < (M << logT); idx += blockDim.x) {
    const int k = idx >> logT;
    A[idx] = k < n ? cmulw<DIR>(A[idx], s.chirp[k]) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  double2 *X = stockham<-1>(s, A, B, logT);
  const double2 *F = DIR > 0 ? s.filt_b : s.filt_f;
  for (int idx = threadIdx.x; idx < (M << logT); idx += blockDim.x) X[idx] = cmul(X[idx], F[idx >> logT]);
  __syncthreads();
  double2 *Y = stockham<+1>(s, X, X == A ? B : A, logT);
  for (int idx = threadIdx.x; idx < (n << logT); idx += blockDim.x) Y[idx] = cmulw<DIR>(Y[idx], s.chirp[idx >> logT]);
  __syncthreads();
  (void)T;
  return Y;
}

// ---------------------------------------------------------------------------
// Wave-level power-of-two FFT.  A wave owns Tw = 64*E/n interleaved sequences
// of length n (element (k,t) at k*Tw + t of the wave's LDS region of 64*E
// elements).  Each Stockham stage: every lane loads its E/R radix-R
// butterflies into registers, twiddles, transforms, and writes them back in
// place.  A wave's LDS operations complete in program order, so no barrier is
// needed between stages; only the wave's own region is touched.
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void dft16(double2 *v) {
  // 16 = 4 x 4: radix-4 over stride-4 groups, twiddles w16^{q r}, radix-4 again
  double2 a[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int q = 0; q < 4; ++q) a[r][q] = v[r + 4 * q];
    dft4<DIR>(a[r][0], a[r][1], a[r][2], a[r][3]);
  }
  constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
  // a[r][q] *= w16^{r q}, w16 = e^{DIR 2 pi i/16}
#pragma unroll
  for (int r = 1; r < 4; ++r)
#pragma unroll
    for (int q = 1; q < 4; ++q) {
      const int e = r * q;  // 1..9
      double cr, ci;
      switch (e) {
        case 1: cr = c1; ci = s1; break;
        case 2: cr = h; ci = h; break;
        case 3: cr = s1; ci = c1; break;
        case 4: cr = 0.0; ci = 1.0; break;
        case 6: cr = -h; ci = h; break;
        default: cr = -c1; ci = -s1; break;  // e = 9: w^9 = (-c1, -s1)
      }
      const double2 x = a[r][q];
      a[r][q] = make_double2(x.x * cr - DIR * x.y * ci, x.y * cr + DIR * x.x * ci);
    }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    dft4<DIR>(a[0][q], a[1][q], a[2][q], a[3][q]);
#pragma unroll
    for (int r = 0; r < 4; ++r) v[q + 4 * r] = a[r][q];
  }
}

template <int R, int DIR>
__device__ __forceinline__ void dft_pow2(double2 *v) {
  if constexpr (R == 16) dft16<DIR>(v);
  else dft_small<R, DIR>(v);
}

template <int R, int E, int DIR>
__device__ __forceinline__ void wave_stage(double2 *__restrict__ buf, int n, int logTw, int Ns,
                                           const double2 *__restrict__ tw, int lane) {
  constexpr int NB = E / R;
  const int Tw = 1 << logTw;
  const int nR = n / R;
  const int twstep = n / (Ns * R);
  double2 v[NB][R];
  int jj[NB], tt[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = lane + 64 * q;
    tt[q] = b & (Tw - 1);
    jj[q] = b >> logTw;
#pragma unroll
    for (int r = 0; r < R; ++r) v[q][r] = buf[((jj[q] + r * nR) << logTw) + tt[q]];
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int jm = jj[q] & (Ns - 1);
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[q][r] = cmulw<DIR>(v[q][r], tw[r * jm * twstep]);
    }
    dft_pow2<R, DIR>(v[q]);
    const int idxD = (jj[q] - jm) * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[((idxD + r * Ns) << logTw) + tt[q]] = v[q][r];
  }
  __builtin_amdgcn_wave_barrier();
}

// Radix plan for a power of two n <= 64*E: as many 16s as possible, then 8/4/2.
template <int E, int DIR>
__device__ __forceinline__ void wave_fft(double2 *buf, int n, int logTw, const double2 *__restrict__ tw) {
  const int lane = threadIdx.x & 63;
  int Ns = 1, rem = n;
  while (rem > 1) {
    if (rem % 16 == 0 && E % 16 == 0) {
      wave_stage<16, E, DIR>(buf, n, logTw, Ns, tw, lane);
      Ns *= 16;
      rem /= 16;
    } else if (rem % 8 == 0 && E % 8 == 0) {
      wave_stage<8, E, DIR>(buf, n, logTw, Ns, tw, lane);
      Ns *= 8;
      rem /= 8;
    } else if (rem % 4 == 0) {
      wave_stage<4, E, DIR>(buf, n, logTw, Ns, tw, lane);
      Ns *= 4;
      rem /= 4;
    } else {
      wave_stage<2, E, DIR>(buf, n, logTw, Ns, tw, lane);
      Ns *= 2;
      rem /= 2;
    }
  }
}

// Length-n transform (direct power of two, or Bluestein on a power-of-two M)
// of the Tw sequences in the wave's region (k < n filled).  Result in place.
template <int E, int DIR>
__device__ __forceinline__ void wave_fft_exec(const FftSpec &s, double2 *buf, int logTw) {
  const int lane = threadIdx.x & 63;
  if (!s.bluestein) {
    wave_fft<E, DIR>(buf, s.M, logTw, s.tw);
    return;
  }
  const int n = s.n, M = s.M;
  for (int idx = lane; idx < (M << logTw); idx += 64) {
    const int k = idx >> logTw;
    buf[idx] = k < n ? cmulw<DIR>(buf[idx], s.chirp[k]) : make_double2(0.0, 0.0);
  }
  __builtin_amdgcn_wave_barrier();
  wave_fft<E, -1>(buf, M, logTw, s.tw);
  const double2 *F = DIR > 0 ? s.filt_b : s.filt_f;
  for (int idx = lane; idx < (M << logTw); idx += 64) buf[idx] = cmul(buf[idx], F[idx >> logTw]);
  __builtin_amdgcn_wave_barrier();
  wave_fft<E, +1>(buf, M, logTw, s.tw);
  for (int idx = lane; idx < (n << logTw); idx += 64) buf[idx] = cmulw<DIR>(buf[idx], s.chirp[idx >> logTw]);
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Workgroup FFT engine.  A workgroup of NT threads owns S = NT*E complex LDS
// slots holding Tw = S/M interleaved sequences of power-of-two length M
// (logical slot k*Tw + t holds element k of sequence t).  The physical LDS
// index of logical slot i is i ^ ((i >> 3) & 15), a permutation inside each
// aligned block of 16 slots: 16 consecutive lanes that read consecutive
// slots, and 16 consecutive lanes that write the slots of a radix-8 stage
// (8*Tw apart at Ns = 1, interleaved groups at larger Ns), all fall in
// distinct 16-byte bank groups.  Offsets that are multiples of 128 slots
// pass through unchanged, so compile-time strides stay immediates.
//
// Stage (radix R, Ns = product of the radices before it): the q-th
// butterfly b = threadIdx.x + NT*q, (jj, tt) = (b / Tw, b % Tw), reads
// elements jj + r*M/R of sequence tt, twiddles them by w^{r*(jj mod Ns)},
// w = e^{DIR 2 pi i/(Ns R)}, transforms, and writes them to
// (jj - jj mod Ns)*R + jj mod Ns + r*Ns.  Every LDS read and every global
// load (twiddles, optional post-multiplier) of a stage is issued before the
// barrier that separates the reads from the in-place writes.  Twiddles come
// from a table of e^{+2 pi i k/M} (k < M) as the powers w, w^2, w^4, w^8;
// the others are products of those (at most three roundings).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lpad(int i) { return i ^ ((i >> 3) & 15); }

// threadIdx.x as a value the compiler cannot trace back to the thread index:
// each stage then computes its slot indices where it uses them, instead of
// having them hoisted out of a tile loop around the transform and held in
// registers through every stage.
__device__ __forceinline__ int stage_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// Workgroup barrier that orders LDS only: the fences are restricted to the
// local address space, so outstanding global loads (prefetches) stay in
// flight across it.  Used where the barrier protects LDS data alone.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__host__ __device__ constexpr int lpad_size(int s) { return (s + 15) & ~15; }

template <int R, int DIR>
__device__ __forceinline__ void twiddle_powers(double2 *v, double2 w1, double2 w2, double2 w4, double2 w8) {
  // v[r] *= w^r for r = 1..R-1 (w^r with the sign of DIR)
  double2 p[R];
  p[1] = w1;
  if constexpr (R > 2) {
    p[2] = w2;
    p[3] = cmul(w1, w2);
  }
  if constexpr (R > 4) {
    p[4] = w4;
    p[5] = cmul(w1, w4);
    p[6] = cmul(w2, w4);
    p[7] = cmul(p[3], w4);
  }
  if constexpr (R > 8) {
    p[8] = w8;
#pragma unroll
    for (int r = 9; r < 16; ++r) p[r] = cmul(p[r - 8], w8);
  }
#pragma unroll
  for (int r = 1; r < R; ++r) v[r] = cmulw<DIR>(v[r], p[r]);
}

template <int R, int E, int NT, int DIR>
__device__ __forceinline__ void team_stage(double2 *buf, int M, int logTw, int Ns, const double2 *__restrict__ tw,
                                           const double2 *__restrict__ post) {
  constexpr int NB = E / R;
  const int Tw = 1 << logTw;
  const int nR = M / R;
  const int twstep = M / (Ns * R);
  double2 v[NB][R];
  double2 w1[NB], w2[NB], w4[NB], w8[NB];
  double2 pm[NB][R];
  int tt[NB], idxD[NB];
  const int tid = stage_tid();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = tid + NT * q;
    tt[q] = b & (Tw - 1);
    const int jj = b >> logTw;
    const int jm = jj & (Ns - 1);
    idxD[q] = (jj - jm) * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) v[q][r] = buf[lpad(((jj + r * nR) << logTw) + tt[q])];
    if (Ns > 1) {
      const int e = jm * twstep;
      w1[q] = tw[e];
      if (R > 2) w2[q] = tw[2 * e];
      if (R > 4) w4[q] = tw[4 * e];
      if (R > 8) w8[q] = tw[8 * e];
    }
    if (post) {
#pragma unroll
      for (int r = 0; r < R; ++r) pm[q][r] = post[idxD[q] + r * Ns];
    }
  }
  lds_barrier();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    if (Ns > 1) twiddle_powers<R, DIR>(v[q], w1[q], w2[q], w4[q], w8[q]);
    dft_pow2<R, DIR>(v[q]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 y = v[q][r];
      if (post) y = cmul(y, pm[q][r]);
      buf[lpad(((idxD[q] + r * Ns) << logTw) + tt[q])] = y;
    }
  }
  lds_barrier();
}

// Radix-8 stage with its twiddles (w, w^2, w^4 per butterfly) and, for the
// last stage of a Bluestein forward transform, the filter values already in
// registers.
template <int E, int NT, int DIR, bool PM>
__device__ __forceinline__ void team_stage8(double2 *buf, int M, int logTw, int Ns, const double2 (&w)[E / 8][3],
                                            const double2 (&pm)[E / 8][8]) {
  constexpr int R = 8, NB = E / R;
  const int Tw = 1 << logTw;
  const int nR = M / R;
  double2 v[NB][R];
  int tt[NB], idxD[NB];
  const int tid = stage_tid();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = tid + NT * q;
    tt[q] = b & (Tw - 1);
    const int jj = b >> logTw;
    const int jm = jj & (Ns - 1);
    idxD[q] = (jj - jm) * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) v[q][r] = buf[lpad(((jj + r * nR) << logTw) + tt[q])];
  }
  lds_barrier();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    // for Ns = 1 the twiddles are exactly 1 and the products exact
    twiddle_powers<R, DIR>(v[q], w[q][0], w[q][1], w[q][2], w[q][2]);
    dft_pow2<R, DIR>(v[q]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 y = v[q][r];
      if constexpr (PM) y = cmul(y, pm[q][r]);
      buf[lpad(((idxD[q] + r * Ns) << logTw) + tt[q])] = y;
    }
  }
  lds_barrier();
}

// Power-of-two transform of all Tw sequences in the workgroup's slots, in
// place: a first stage of radix R0 = 2^(log2 M mod 3) (no twiddles), then
// radix-8 stages whose twiddles (at most S = floor(log2(E NT) / 3) stages:
// 3 for the 2048 slots of 256 threads, 4 for 8192 slots) are loaded once, up
// front, by team_twiddles.
__host__ __device__ constexpr int team_ilog2(int x) { return x <= 1 ? 0 : 1 + team_ilog2(x / 2); }
template <int E, int NT>
constexpr int team_s() { return team_ilog2(E * NT) / 3; }
template <int E, int S>
struct TeamTw {
  double2 w[S][E / 8][3];
};

__device__ __forceinline__ int ilog2_dev(int x) { return 31 - __clz(x); }

// ns0: Ns of the first radix-8 stage (1 << (log2 M mod 3) when the radix-2/4
// stage comes first, 1 when it comes last)
template <int E, int NT>
__device__ __forceinline__ void team_twiddles(TeamTw<E, team_s<E, NT>()> &T, int M, int logTw,
                                              const double2 *__restrict__ tw, int ns0 = -1) {
  const int logM = ilog2_dev(M), nst = logM / 3;
  int Ns = ns0 > 0 ? ns0 : 1 << (logM % 3);
#pragma unroll
  for (int s = 0; s < team_s<E, NT>(); ++s) {
    if (s < nst) {
      const int step = M / (Ns * 8);
#pragma unroll
      for (int q = 0; q < E / 8; ++q) {
        const int jj = (threadIdx.x + NT * q) >> logTw;
        const int e = (jj & (Ns - 1)) * step;
        T.w[s][q][0] = tw[e];
        T.w[s][q][1] = tw[2 * e];
        T.w[s][q][2] = tw[4 * e];
      }
    }
    Ns *= 8;
  }
}

template <int E, int NT, int DIR>
__device__ __forceinline__ void team_fft(double2 *buf, int M, int logTw, const TeamTw<E, team_s<E, NT>()> &T,
                                         const double2 *__restrict__ tw, const double2 *__restrict__ post,
                                         const double2 (&pm)[E / 8][8], bool use_pm) {
  static_assert(E % 8 == 0, "radix-8 engine");
  const int logM = ilog2_dev(M), nst = logM / 3, r0 = logM % 3;
  if (r0 == 1) team_stage<2, E, NT, DIR>(buf, M, logTw, 1, tw, nst == 0 ? post : nullptr);
  else if (r0 == 2) team_stage<4, E, NT, DIR>(buf, M, logTw, 1, tw, nst == 0 ? post : nullptr);
  int Ns = 1 << r0;
#pragma unroll
  for (int s = 0; s < team_s<E, NT>(); ++s) {
    if (s < nst) {
      if (use_pm && s == nst - 1) team_stage8<E, NT, DIR, true>(buf, M, logTw, Ns, T.w[s], pm);
      else team_stage8<E, NT, DIR, false>(buf, M, logTw, Ns, T.w[s], pm);
      Ns *= 8;
    }
  }
}

// Bluestein with the forward transform's last stage handed to the backward
// transform's first stage in registers.  Forward: radix R0 = 2^(log2 M mod 3)
// first, then radix-8 stages; its last stage (Ns = M/8) leaves outputs
// jj + r*M/8 of butterfly jj in registers, multiplied by the filter.  The
// backward transform runs its radix-8 stages first: the first (Ns = 1)
// reads exactly those elements of butterfly jj, so it starts from the
// registers; its radix-R0 stage comes last (Ns = M/R0) and multiplies its
// outputs by post_last[k] (k < M) if given.  Saves one LDS write, one LDS
// read and a barrier pair per transform against two separate FFTs.
// FIRST_DONE: the caller has already written the outputs of the forward
// transform's first stage (radix 2^(log2 M mod 3), Ns = 1) to the slots.
// LAST_REGS (radix-4 last stage only): the backward transform's last stage
// leaves outputs jj + r M/4, r < 2, of its butterflies in yl (butterfly
// b = threadIdx.x + NT q: sequence b mod Tw, jj = b / Tw) instead of the
// slots, without post_last; the slots are not written by that stage.
template <int E, int NT, int DIR, bool FIRST_DONE, bool LAST_REGS>
__device__ __forceinline__ void team_bluestein_x(const FftSpec &s, double2 *buf, int M, int logTw,
                                                 const double2 *__restrict__ post_last, double2 (&yl)[E / 4][2]) {
  constexpr int NB = E / 8, S = team_s<E, NT>();
  const int logM = ilog2_dev(M), nst = logM / 3, r0 = logM % 3;
  const double2 *F = DIR > 0 ? s.filt_b : s.filt_f;
  const int Tw = 1 << logTw, nR = M >> 3;
  TeamTw<E, S> T;
  team_twiddles<E, NT>(T, M, logTw, s.tw);
  double2 pm[NB][8];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int jj = (threadIdx.x + NT * q) >> logTw;
#pragma unroll
    for (int r = 0; r < 8; ++r) pm[q][r] = F[jj + r * nR];
  }
  // forward
  if (!FIRST_DONE) {
    if (r0 == 1) team_stage<2, E, NT, -1>(buf, M, logTw, 1, s.tw, nullptr);
    else if (r0 == 2) team_stage<4, E, NT, -1>(buf, M, logTw, 1, s.tw, nullptr);
  }
  int Ns = 1 << r0;
#pragma unroll
  for (int st = 0; st < S - 1; ++st) {
    if (st < nst - 1) {
      team_stage8<E, NT, -1, false>(buf, M, logTw, Ns, T.w[st], pm);
      Ns *= 8;
    }
  }
  double2 v[NB][8];
  int tt[NB], jjv[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = threadIdx.x + NT * q;
    tt[q] = b & (Tw - 1);
    jjv[q] = b >> logTw;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[q][r] = buf[lpad(((jjv[q] + r * nR) << logTw) + tt[q])];
  }
  // last forward stage's twiddles, then the backward transform's (8-first order)
  double2 wl[NB][3];
#pragma unroll
  for (int q = 0; q < NB; ++q)
#pragma unroll
    for (int i = 0; i < 3; ++i) wl[q][i] = T.w[S - 1][q][i];
#pragma unroll
  for (int st = S - 2; st >= 0; --st) {
    if (nst == st + 1) {
#pragma unroll
      for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int i = 0; i < 3; ++i) wl[q][i] = T.w[st][q][i];
    }
  }
  team_twiddles<E, NT>(T, M, logTw, s.tw, 1);
  lds_barrier();  // every read of the forward transform is done
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    twiddle_powers<8, -1>(v[q], wl[q][0], wl[q][1], wl[q][2], wl[q][2]);
    dft_pow2<8, -1>(v[q]);
#pragma unroll
    for (int r = 0; r < 8; ++r) v[q][r] = cmul(v[q][r], pm[q][r]);
    // backward first stage, Ns = 1: no twiddles, outputs 8 jj + r
    dft_pow2<8, +1>(v[q]);
#pragma unroll
    for (int r = 0; r < 8; ++r) buf[lpad(((jjv[q] * 8 + r) << logTw) + tt[q])] = v[q][r];
  }
  lds_barrier();
  Ns = 8;
#pragma unroll
  for (int st = 1; st < S; ++st) {
    if (st < nst) {
      team_stage8<E, NT, +1, false>(buf, M, logTw, Ns, T.w[st], pm);
      Ns *= 8;
    }
  }
  if constexpr (LAST_REGS) {
    // radix-4 stage at Ns = M/4: butterfly jj reads jj + r M/4 and writes
    // the same positions; twiddle w^r with w = tw[jj]
#pragma unroll
    for (int q = 0; q < E / 4; ++q) {
      const int b = threadIdx.x + NT * q;
      const int t = b & (Tw - 1), jj = b >> logTw, nq = M >> 2;
      double2 x[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) x[r] = buf[lpad(((jj + r * nq) << logTw) + t)];
      twiddle_powers<4, +1>(x, s.tw[jj], s.tw[2 * jj], x[0], x[0]);
      dft_pow2<4, +1>(x);
      yl[q][0] = x[0];
      yl[q][1] = x[1];
    }
    return;
  }
  if (r0 == 1) team_stage<2, E, NT, +1>(buf, M, logTw, Ns, s.tw, post_last);
  else if (r0 == 2) team_stage<4, E, NT, +1>(buf, M, logTw, Ns, s.tw, post_last);
  else if (post_last) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = threadIdx.x + NT * q;
      buf[lpad(i)] = cmul(buf[lpad(i)], post_last[i >> logTw]);
    }
    lds_barrier();
  }
}

template <int E, int NT, int DIR, bool FIRST_DONE = false>
__device__ __forceinline__ void team_bluestein(const FftSpec &s, double2 *buf, int M, int logTw,
                                               const double2 *__restrict__ post_last) {
  double2 yl[E / 4][2];
  team_bluestein_x<E, NT, DIR, FIRST_DONE, false>(s, buf, M, logTw, post_last, yl);
}

// Middle of a length-n transform on the workgroup engine.  Direct power of
// two: the FFT itself.  Bluestein (M >= 2n-1): the caller has loaded
// x_k * chirp_k^{DIR} (k < n, zero up to M); this runs the forward FFT with
// the filter folded into its last stage and the backward FFT; the output k
// is multiplied by post_last[k] if given (the caller multiplies by
// chirp_k^{DIR} otherwise).  M, logTw and bluestein are passed separately
// from s so that kernels specialised on the work length hand down
// compile-time constants (every function here is inlined).
template <int E, int NT, int DIR, bool FIRST_DONE = false>
__device__ __forceinline__ void team_fft_core(const FftSpec &s, double2 *buf, int M, int logTw, bool bluestein,
                                              const double2 *__restrict__ post_last = nullptr) {
  if (bluestein && ilog2_dev(M) >= 3) {
    team_bluestein<E, NT, DIR, FIRST_DONE>(s, buf, M, logTw, post_last);
    return;
  }
  TeamTw<E, team_s<E, NT>()> T;
  team_twiddles<E, NT>(T, M, logTw, s.tw);
  double2 pm[E / 8][8];
  if (!bluestein) {
    team_fft<E, NT, DIR>(buf, M, logTw, T, s.tw, nullptr, pm, false);
    return;
  }
  const double2 *F = DIR > 0 ? s.filt_b : s.filt_f;
  const int nst = ilog2_dev(M) / 3;
  if (nst > 0) {
    // last radix-8 stage: Ns = M/8, output k = jj + r*M/8
#pragma unroll
    for (int q = 0; q < E / 8; ++q) {
      const int jj = (threadIdx.x + NT * q) >> logTw;
#pragma unroll
      for (int r = 0; r < 8; ++r) pm[q][r] = F[jj + r * (M >> 3)];
    }
  }
  team_fft<E, NT, -1>(buf, M, logTw, T, s.tw, F, pm, nst > 0);
  team_fft<E, NT, +1>(buf, M, logTw, T, s.tw, post_last, pm, false);
  if (post_last && nst > 0) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = threadIdx.x + NT * q;
      buf[lpad(i)] = cmul(buf[lpad(i)], post_last[i >> logTw]);
    }
    lds_barrier();
  }
}

// Workgroup id remapped so that consecutive logical tiles run on the same
// XCD (the hardware deals workgroups to the 8 XCDs round-robin), which keeps
// tiles that share cache lines in one L2.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  constexpr int X = 8;
  const int q = nb / X, rm = nb % X;
  const int x = bid % X, i = bid / X;
  return x * q + (x < rm ? x : rm) + i;
}

// ---------------------------------------------------------------------------
// Odd-length workgroup engine: length NN = R0 R1 R2 with odd prime radices,
// T sequences interleaved in LDS (element k of sequence t at slot k T + t),
// NT threads.  Used where Bluestein would double the work length (2hi - 1 =
// 1023 = 31 * 11 * 3 runs here instead of as two length-2048 transforms).
//
// Stockham stages in the order R0, R1, R2 (Ns = 1, R0, R0 R1).  A stage
// reads every butterfly of the thread into registers (and, for Ns > 1, its
// twiddles w^{r jm} from a table of e^{+2 pi i k/NN}, one rounding each),
// waits at a barrier, transforms and writes in place; a second barrier
// closes the stage.  The radix-R butterfly forms the symmetric pairs
// x_r +- x_{R-r} in place and writes each output pair y_q, y_{R-q} as soon
// as it is accumulated, so a thread holds its R inputs and four
// accumulators rather than 2R values.
// ---------------------------------------------------------------------------
template <int R, int DIR, int SPLIT = 1, typename Out>
__device__ __forceinline__ void dft_odd_out(double2 *v, Out out, int part = 0) {
  constexpr int H = (R - 1) / 2, QP = (H + SPLIT - 1) / SPLIT;
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    const double2 a = v[r], b = v[R - r];
    v[r] = cadd(a, b);
    v[R - r] = csub(a, b);
  }
  if (part == 0) {
    double2 y0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; ++r) y0 = cadd(y0, v[r]);
    out(0, y0);
  }
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    if (SPLIT > 1 && (q - 1) / QP != part) continue;  // part is uniform over the wave
    double re = v[0].x, im = v[0].y, ire = 0.0, iim = 0.0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int e = (r * q) % R;
      const double cc = RadixConst<R>::c[e - 1], ss = RadixConst<R>::s[e - 1];
      re += v[r].x * cc;
      im += v[r].y * cc;
      ire += v[R - r].x * ss;
      iim += v[R - r].y * ss;
    }
    const double2 base = make_double2(re, im);
    const double2 rot = cmuli<DIR>(make_double2(ire, iim));
    out(q, cadd(base, rot));
    out(R - q, csub(base, rot));
  }
}

// SPLIT > 1 (first stage only, NS = 1): the threads form SPLIT groups of
// NT / SPLIT (whole waves); every group loads all butterflies and group p
// forms the output pairs q of its share, which divides the stage's latency
// (its longest chain of dependent instructions per lane) by SPLIT at the
// price of SPLIT times the LDS reads.
template <int R, int NN, int NS, int DIR, int NT, int T, int SPLIT = 1>
__device__ __forceinline__ void odd_stage(double2 *buf, const double2 *__restrict__ tw) {
  static_assert(SPLIT == 1 || NS == 1, "split stages carry no twiddles");
  static_assert(NT % SPLIT == 0 && (NT / SPLIT) % 64 == 0, "split groups are whole waves");
  constexpr int NG = NT / SPLIT;
  constexpr int nR = NN / R, nb = nR * T, NB = (nb + NG - 1) / NG, twstep = NN / (NS * R);
  const int tid = stage_tid();
  const int part = SPLIT > 1 ? __builtin_amdgcn_readfirstlane(tid / NG) : 0, lt = SPLIT > 1 ? tid % NG : tid;
  double2 v[NB][R];
  double2 w[NB][R];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = lt + NG * q;
    if (b < nb) {
      const int t = b % T, jj = b / T;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[(jj + r * nR) * T + t];
      if constexpr (NS > 1) {
        const int jm = jj % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) w[q][r] = tw[r * jm * twstep];
      }
    }
  }
  lds_barrier();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = lt + NG * q;
    if (b < nb) {
      const int t = b % T, jj = b / T;
      const int jm = NS > 1 ? jj % NS : 0;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cmulw<DIR>(v[q][r], w[q][r]);
      }
      const int o = ((jj - jm) * R + jm) * T + t;
      dft_odd_out<R, DIR, SPLIT>(v[q], [&](int r, double2 y) { buf[o + r * NS * T] = y; }, part);
    }
  }
  lds_barrier();
}

template <int R0, int R1, int R2, int DIR, int NT, int T, int SPLIT = 1>
__device__ __forceinline__ void odd_fft(double2 *buf, const double2 *__restrict__ tw) {
  constexpr int NN = R0 * R1 * R2;
  odd_stage<R0, NN, 1, DIR, NT, T, SPLIT>(buf, tw);
  odd_stage<R1, NN, R0, DIR, NT, T>(buf, tw);
  odd_stage<R2, NN, R0 * R1, DIR, NT, T>(buf, tw);
}

}  // namespace s2dw
