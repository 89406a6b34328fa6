"""Python access to the CPU oracle (oracle/s2dw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
``s2dw_amd`` never imports this module.

Arrays keep the reference's Fortran layouts (numpy ``order='F'``):
``flm[l, m]`` (B x B complex), ``K[j, l]`` ((J+1) x B), ``S[l, n]`` (B x N
complex), per-scale wavelet coefficients ``w[a, b, g]`` ((2hi-1) x 2hi x N),
static ``wav[j, a, b, g]`` ((J+1) x (2B-1) x 2B x N), ``scoeff[l, m]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, capture_output=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_assimilate_int.argtypes = [C.c_double, _ip]
        L.orc_comp_Jmax.argtypes = [C.c_int, C.c_double]
        L.orc_params_valid.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int]
        L.orc_band_limits.argtypes = [C.c_int, C.c_int, C.c_double, _ip, _ip]
        L.orc_init_kernels.argtypes = [_dp, _dp, _ip, C.c_int, C.c_int, C.c_double]
        L.orc_init_directionality.argtypes = [_dp, C.c_int, C.c_int]
        L.orc_admiss.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int]
        L.orc_quad_weights.argtypes = [C.c_double, C.c_int]
        L.orc_quad_weights.restype = C.c_double
        L.orc_dl_beta_operator.argtypes = [_dp, C.c_double, C.c_int]
        L.orc_fft.argtypes = [C.c_int, _dp, _dp, C.c_int]
        L.orc_analysis_dynamic.argtypes = [C.POINTER(_dp), _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_double]
        L.orc_analysis_static.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_double]
        L.orc_synthesis_dynamic.argtypes = [_dp, C.POINTER(_dp), _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_double]
        L.orc_synthesis_static.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_double]
        L.orc_test_gen_flm.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_stat_moments.argtypes = [C.POINTER(_dp), C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp,
                                       _dp, C.POINTER(_dp), _dp, _dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_dl_precise.argtypes = [C.c_int]
        L.orc_set_exact_phase.argtypes = [C.c_int]
        L.orc_set_exact_beta.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def set_dl_precise(on: bool) -> None:
    """Diagnostic: compute Wigner planes with the same recursion in long double."""
    lib().orc_set_dl_precise(int(bool(on)))


def set_exact_phase(on: bool) -> None:
    """Diagnostic: exact powers of i instead of the reference's rounded exp(I*k*PION2)."""
    lib().orc_set_exact_phase(int(bool(on)))


def set_exact_beta(on: bool) -> None:
    """Diagnostic: exact beta-sum basis instead of cos/sin of the rounded m'*beta_b."""
    lib().orc_set_exact_beta(int(bool(on)))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def assimilate_int(x: float):
    y = C.c_int(0)
    ok = lib().orc_assimilate_int(float(x), C.byref(y))
    return bool(ok), int(y.value)


def comp_Jmax(B: int, alpha: float) -> int:
    return int(lib().orc_comp_Jmax(int(B), float(alpha)))


def band_limits(J: int, B: int, alpha: float):
    hi, lo = [], []
    for jj in range(J + 1):
        h, l_ = C.c_int(), C.c_int()
        lib().orc_band_limits(jj, int(B), float(alpha), C.byref(h), C.byref(l_))
        hi.append(h.value)
        lo.append(l_.value)
    return hi, lo


def init_kernels(B: int, alpha: float, J: int):
    K = np.zeros((J + 1, B), dtype=np.float64, order="F")
    Phi2 = np.zeros(B, dtype=np.float64)
    bl = C.c_int(0)
    err = lib().orc_init_kernels(_p(K), _p(Phi2), C.byref(bl), int(J), int(B), float(alpha))
    if err:
        raise RuntimeError(f"orc_init_kernels error {err}")
    return K, Phi2, int(bl.value)


def init_directionality(B: int, N: int):
    S = np.zeros((B, N), dtype=np.complex128, order="F")
    err = lib().orc_init_directionality(_p(S), int(B), int(N))
    if err:
        raise RuntimeError(f"orc_init_directionality error {err}")
    return S


def admiss(K: np.ndarray, Phi2: np.ndarray, B: int, J: int):
    a = np.zeros(B, dtype=np.float64)
    K = np.asfortranarray(K, dtype=np.float64)
    ok = lib().orc_admiss(_p(a), _p(K), _p(np.ascontiguousarray(Phi2, dtype=np.float64)), int(B), int(J))
    return a, bool(ok)


def quad_weights(theta: float, B: int) -> float:
    return float(lib().orc_quad_weights(float(theta), int(B)))


def dl_beta_operator(l: int, beta: float = np.pi / 2) -> np.ndarray:
    """Full plane dl[a + l, b + l] = d^l_{ab}(beta), a, b in [-l, l]."""
    n = 2 * l + 1
    dl = np.zeros((n, n), dtype=np.float64, order="F")
    lib().orc_dl_beta_operator(_p(dl), float(beta), int(l))
    return dl


def fft(x: np.ndarray, sign: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().orc_fft(int(x.size), _p(x), _p(out), int(sign))
    return out


def test_gen_flm(B: int, seed: int = 1) -> np.ndarray:
    """s2dw_test_gen_flm's ran2 uniform(-1,1) flm (s2dw_test.f90:192-224)."""
    f = np.zeros((B, B), dtype=np.complex128, order="F")
    lib().orc_test_gen_flm(int(B), _p(f), int(seed))
    return f


def analysis_dynamic(flm, K, S, J, B, N, bl_scoeff, alpha):
    hi, _ = band_limits(J, B, alpha)
    flm = np.asfortranarray(flm, dtype=np.complex128)
    K = np.asfortranarray(K, dtype=np.float64)
    S = np.asfortranarray(S, dtype=np.complex128)
    wavs = [np.zeros((2 * h - 1, 2 * h, N), dtype=np.float64, order="F") for h in hi]
    ptrs = (_dp * (J + 1))(*[_p(w) for w in wavs])
    scoeff = np.zeros((bl_scoeff, bl_scoeff), dtype=np.complex128, order="F")
    err = lib().orc_analysis_dynamic(ptrs, _p(scoeff), _p(flm), _p(K), _p(S), int(J), int(B), int(N),
                                     int(bl_scoeff), float(alpha))
    if err:
        raise RuntimeError(f"orc_analysis_dynamic error {err}")
    return wavs, scoeff


def analysis_static(flm, K, S, J, B, N, bl_scoeff, alpha):
    flm = np.asfortranarray(flm, dtype=np.complex128)
    K = np.asfortranarray(K, dtype=np.float64)
    S = np.asfortranarray(S, dtype=np.complex128)
    wav = np.zeros((J + 1, 2 * B - 1, 2 * B, N), dtype=np.float64, order="F")
    scoeff = np.zeros((bl_scoeff, bl_scoeff), dtype=np.complex128, order="F")
    err = lib().orc_analysis_static(_p(wav), _p(scoeff), _p(flm), _p(K), _p(S), int(J), int(B), int(N),
                                    int(bl_scoeff), float(alpha))
    if err:
        raise RuntimeError(f"orc_analysis_static error {err}")
    return wav, scoeff


def synthesis_dynamic(wavs, scoeff, K, Phi2, S, J, B, N, bl_scoeff, alpha):
    wavs = [np.asfortranarray(w, dtype=np.float64) for w in wavs]
    ptrs = (_dp * (J + 1))(*[_p(w) for w in wavs])
    flm = np.zeros((B, B), dtype=np.complex128, order="F")
    scoeff = np.asfortranarray(scoeff, dtype=np.complex128)
    K = np.asfortranarray(K, dtype=np.float64)
    S = np.asfortranarray(S, dtype=np.complex128)
    Phi2 = np.ascontiguousarray(Phi2, dtype=np.float64)
    err = lib().orc_synthesis_dynamic(_p(flm), ptrs, _p(scoeff), _p(K), _p(Phi2), _p(S), int(J), int(B),
                                      int(N), int(bl_scoeff), float(alpha))
    if err:
        raise RuntimeError(f"orc_synthesis_dynamic error {err}")
    return flm


def synthesis_static(wav, scoeff, K, Phi2, S, J, B, N, bl_scoeff, alpha):
    wav = np.asfortranarray(wav, dtype=np.float64)
    flm = np.zeros((B, B), dtype=np.complex128, order="F")
    scoeff = np.asfortranarray(scoeff, dtype=np.complex128)
    K = np.asfortranarray(K, dtype=np.float64)
    S = np.asfortranarray(S, dtype=np.complex128)
    Phi2 = np.ascontiguousarray(Phi2, dtype=np.float64)
    err = lib().orc_synthesis_static(_p(flm), _p(wav), _p(scoeff), _p(K), _p(Phi2), _p(S), int(J), int(B),
                                     int(N), int(bl_scoeff), float(alpha))
    if err:
        raise RuntimeError(f"orc_synthesis_static error {err}")
    return flm


def stat_moments(wavs, J, B, N, alpha, masks=None):
    """s2dw_stat_moments (s2dw_stat_mod.f90:81-275) on per-scale arrays
    ``wavs[jj]`` ((2hi-1) x 2hi x N, order 'F', float64), masked in place when
    ``masks`` is given.  Returns (code, mean, var, skew, kur, five, six)."""
    n = J + 1
    wp = (_dp * n)(*[_p(w) for w in wavs])
    mp = (_dp * n)(*[_p(m) for m in masks]) if masks is not None else None
    out = [np.zeros(n) for _ in range(6)]
    code = lib().orc_stat_moments(wp, J, B, N, alpha, *[_p(o) for o in out[:4]], mp, _p(out[4]), _p(out[5]))
    return (code, *out)


class Setup:
    """Setup bundle for one (B, N, alpha, J): kernels, directionality."""

    def __init__(self, B: int, N: int, alpha: float, J: int | None = None, J_min: int | None = None):
        self.B, self.N, self.alpha = int(B), int(N), float(alpha)
        jmax = comp_Jmax(B, alpha)
        if J is None:
            J = jmax if J_min is None else jmax - int(J_min)
        self.J = int(J)
        self.K, self.Phi2, self.bl_scoeff = init_kernels(B, alpha, self.J)
        self.S = init_directionality(B, N)
        self.hi, self.lo = band_limits(self.J, B, alpha)

    def analysis(self, flm):
        return analysis_dynamic(flm, self.K, self.S, self.J, self.B, self.N, self.bl_scoeff, self.alpha)

    def synthesis(self, wavs, scoeff):
        return synthesis_dynamic(wavs, scoeff, self.K, self.Phi2, self.S, self.J, self.B, self.N,
                                 self.bl_scoeff, self.alpha)
