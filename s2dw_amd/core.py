"""Host-side mirror of the reference's ``s2dw_core_mod`` (src/mod/s2dw_core_mod.f90)
over the C ABI of libs2dw_hip.so (include/s2dw_hip.h).

Procedure names, argument meaning and array layouts follow the Fortran
module: arrays are numpy arrays in Fortran order, ``flm[l, m]`` (B x B
complex), ``K_gamma[j, l]`` ((J+1) x B), ``Phi2[l]``, ``Slm[l, n]`` (B x N
complex), ``scoeff[l, m]``, per-scale wavelet coefficients ``w[a, b, g]``
((2hi-1) x 2hi x N) and the static ``wav[j, a, b, g]``.  Where the Fortran
routines halt through ``s2dw_error`` (s2dw_error_mod.f90:127-191) these raise
:class:`S2dwError` carrying the same numeric code and message text.

Every transform runs on the GPU through the HIP library; there is no CPU
fallback.  Loading fails loudly when libs2dw_hip.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libs2dw_hip.so")

# s2dw_error_mod.f90:37-52 and the comments at 58-76
S2DW_ERROR_NONE = 0
S2DW_ERROR_INIT = 1
S2DW_ERROR_NOT_INIT = 2
S2DW_ERROR_INIT_FAIL = 3
S2DW_ERROR_MEM_ALLOC_FAIL = 4
S2DW_ERROR_ARTH = 5
S2DW_ERROR_ARTH_WARNING = 6
S2DW_ERROR_SIZE_WARNING = 7
S2DW_ERROR_SIZE_INVALID = 8
S2DW_ERROR_SIZE_NOT_DEF = 9
S2DW_ERROR_ARG_INVALID = 10
S2DW_ERROR_ARG_WARNING = 11
S2DW_ERROR_QUAD_STEP_EXCEED = 12
S2DW_ERROR_ADMISS_FAIL = 13
S2DW_ERROR_FILEIO = 14
ERROR_COMMENT = [
    "No error", "Attempt to initialise object that has already been initialised", "Object not initialised",
    "Object initialisation failed", "Memory allocation failed", "Arithmetic exception", "Arithmetic warning",
    "Warning: Sizes not in recommended range", "Invalid sizes", "Sizes not defined", "Arguments invalid",
    "Argument warning", "Exceeded number of steps limit whem computing quadrature",
    "Admissibility test failed (resolution of identity not satisfied)", "File IO error",
]

PREC_FP64 = 0
PREC_SPLIT = 1
# rounding of the FP64 plain form (OR-ed into precision): streamed recursions
# (default), Wigner planes with the reference's rounding, or those planes plus
# the reference's rounded phases and beta-sum basis (the host-buffer entry
# points use the last)
WIGNER_STREAMED = 0
WIGNER_TABLE = 16
ROUNDING_REFERENCE = 48


class S2dwError(RuntimeError):
    """Raised where the reference calls s2dw_error with a halting code."""

    def __init__(self, code: int, procedure: str, detail: str = ""):
        self.code = int(code)
        self.procedure = procedure
        text = ERROR_COMMENT[code] if 0 <= code < len(ERROR_COMMENT) else f"code {code}"
        msg = f"S2DW_ERROR: Error '{text}' occured in procedure '{procedure}'"
        if detail:
            msg += f" - {detail}"
        super().__init__(msg)


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_llp = C.POINTER(C.c_longlong)
_lib = None


def lib() -> C.CDLL:
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m s2dw_amd.build`")
        # torch's wheel carries its own HIP runtime (soname libamdhip64.so.7).
        # Loading torch first makes this library bind to that same runtime by
        # soname; loaded the other way round, the process would hold two HIP
        # runtimes and torch would find no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.s2dw_hip_last_error.restype = C.c_char_p
        L.s2dw_hip_version.restype = C.c_char_p
        L.s2dw_hip_init_kernels.argtypes = [_dp, _dp, _ip, C.c_int, C.c_int, C.c_double]
        L.s2dw_hip_init_directionality.argtypes = [_dp, C.c_int, C.c_int]
        L.s2dw_hip_comp_Jmax.argtypes = [C.c_int, C.c_double]
        L.s2dw_hip_admiss.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _ip]
        L.s2dw_hip_assimilate_int.argtypes = [C.c_double, _ip]
        L.s2dw_hip_band_limits.argtypes = [C.c_int, C.c_int, C.c_double, _ip, _ip]
        L.s2dw_hip_analysis_flm2wav.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.c_double]
        L.s2dw_hip_analysis_flm2wav_dynamic.argtypes = [C.POINTER(_dp), _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                                        C.c_int, C.c_int, C.c_double]
        L.s2dw_hip_synthesis_wav2flm.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                                 C.c_int, C.c_double]
        L.s2dw_hip_synthesis_wav2flm_dynamic.argtypes = [_dp, C.POINTER(_dp), _dp, _dp, _dp, _dp, C.c_int,
                                                         C.c_int, C.c_int, C.c_int, C.c_double]
        L.s2dw_hip_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp,
                                           _dp, C.c_int, C.c_void_p]
        L.s2dw_hip_plan_destroy.argtypes = [C.c_void_p]
        L.s2dw_hip_plan_sizes.argtypes = [C.c_void_p, _llp, _ip, _llp, _ip, _ip]
        L.s2dw_hip_plan_device_bytes.argtypes = [C.c_void_p]
        L.s2dw_hip_plan_device_bytes.restype = C.c_longlong
        L.s2dw_hip_analysis_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.s2dw_hip_synthesis_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.s2dw_hip_plan_sync.argtypes = [C.c_void_p]
        L.s2dw_hip_plan_stream.argtypes = [C.c_void_p]
        L.s2dw_hip_plan_stream.restype = C.c_void_p
        L.s2dw_hip_wigner_plane.argtypes = [C.c_int, _dp, _dp]
        L.s2dw_hip_plan_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.s2dw_hip_plan_set_graphs.argtypes = [C.c_void_p, C.c_int]
        L.s2dw_hip_plan_autotune.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        L.s2dw_hip_plan_stage_times.argtypes = [C.c_void_p, _dp, _ip]
        L.s2dw_hip_stat_moments.argtypes = [C.POINTER(_dp), C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp,
                                            _dp, C.POINTER(_dp), _dp, _dp]
        L.s2dw_hip_stat_moments_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.s2dw_hip_analysis_moments_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.s2dw_hip_plan_moments_fusable.argtypes = [C.c_void_p]
        L.s2dw_hip_plan_wigner_form.argtypes = [C.c_void_p]
        L.s2dw_hip_stat_moments_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                                          C.c_double, C.c_void_p, C.c_void_p]
        L.s2dw_hip_plan_wigner_table.argtypes = [C.c_void_p, C.c_int, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _check(code: int, procedure: str) -> None:
    if code != 0:
        raise S2dwError(code, procedure, lib().s2dw_hip_last_error().decode())


# ---------------------------------------------------------------- setup
def s2dw_core_comp_Jmax(B: int, alpha: float) -> int:
    """s2dw_core_comp_Jmax, s2dw_core_mod.f90:1365-1376."""
    return int(lib().s2dw_hip_comp_Jmax(int(B), float(alpha)))


def s2dw_core_assimilate_int(x: float):
    """s2dw_core_assimilate_int, s2dw_core_mod.f90:1398-1415 -> (assimilated, y)."""
    y = C.c_int(0)
    ok = lib().s2dw_hip_assimilate_int(float(x), C.byref(y))
    return bool(ok), int(y.value)


def s2dw_core_init_kernels(J: int, B: int, alpha: float):
    """s2dw_core_init_kernels, s2dw_core_mod.f90:97-157 -> (K_gamma, Phi2, bl_scoeff)."""
    K = np.zeros((J + 1, B), dtype=np.float64, order="F")
    Phi2 = np.zeros(B, dtype=np.float64)
    bl = C.c_int(0)
    _check(lib().s2dw_hip_init_kernels(_p(K), _p(Phi2), C.byref(bl), int(J), int(B), float(alpha)),
           "s2dw_core_init_kernels")
    return K, Phi2, int(bl.value)


def s2dw_core_init_directionality(B: int, N: int) -> np.ndarray:
    """s2dw_core_init_directionality, s2dw_core_mod.f90:219-250 -> Slm."""
    S = np.zeros((B, N), dtype=np.complex128, order="F")
    _check(lib().s2dw_hip_init_directionality(_p(S), int(B), int(N)), "s2dw_core_init_directionality")
    return S


def s2dw_core_admiss(K_gamma: np.ndarray, Phi2: np.ndarray, B: int, J: int):
    """s2dw_core_admiss, s2dw_core_mod.f90:285-304 -> (admiss, pass)."""
    a = np.zeros(B, dtype=np.float64)
    ok = C.c_int(0)
    K = np.asfortranarray(K_gamma, dtype=np.float64)
    P = np.ascontiguousarray(Phi2, dtype=np.float64)
    _check(lib().s2dw_hip_admiss(_p(a), _p(K), _p(P), int(B), int(J), C.byref(ok)), "s2dw_core_admiss")
    return a, bool(ok.value)


def band_limits(J: int, B: int, alpha: float):
    """(bl_hi, bl_lo) lists per scale, as s2dw_core_mod.f90:606-616 computes them."""
    hi = (C.c_int * (J + 1))()
    lo = (C.c_int * (J + 1))()
    _check(lib().s2dw_hip_band_limits(int(J), int(B), float(alpha), hi, lo), "band_limits")
    return list(hi), list(lo)


def _inputs(flm, K, S):
    return (np.asfortranarray(flm, dtype=np.complex128), np.asfortranarray(K, dtype=np.float64),
            np.asfortranarray(S, dtype=np.complex128))


# ---------------------------------------------------------------- transforms
def s2dw_core_analysis_flm2wav(flm, K_gamma, Slm, J, B, N, bl_scoeff, alpha):
    """s2dw_core_analysis_flm2wav, s2dw_core_mod.f90:342-534 -> (wav, scoeff), static wav."""
    flm, K, S = _inputs(flm, K_gamma, Slm)
    wav = np.zeros((J + 1, 2 * B - 1, 2 * B, N), dtype=np.float64, order="F")
    sc = np.zeros((bl_scoeff, bl_scoeff), dtype=np.complex128, order="F")
    _check(lib().s2dw_hip_analysis_flm2wav(_p(wav), _p(sc), _p(flm), _p(K), _p(S), int(J), int(B), int(N),
                                           int(bl_scoeff), float(alpha)), "s2dw_core_analysis_flm2wav")
    return wav, sc


def s2dw_core_analysis_flm2wav_dynamic(flm, K_gamma, Slm, J, B, N, bl_scoeff, alpha):
    """s2dw_core_analysis_flm2wav_dynamic, s2dw_core_mod.f90:572-775 -> (wavdyn list, scoeff)."""
    flm, K, S = _inputs(flm, K_gamma, Slm)
    hi, _ = band_limits(J, B, alpha)
    wavs = [np.zeros((2 * h - 1, 2 * h, N), dtype=np.float64, order="F") for h in hi]
    ptrs = (_dp * (J + 1))(*[_p(w) for w in wavs])
    sc = np.zeros((bl_scoeff, bl_scoeff), dtype=np.complex128, order="F")
    _check(lib().s2dw_hip_analysis_flm2wav_dynamic(ptrs, _p(sc), _p(flm), _p(K), _p(S), int(J), int(B), int(N),
                                                   int(bl_scoeff), float(alpha)),
           "s2dw_core_analysis_flm2wav_dynamic")
    return wavs, sc


def s2dw_core_synthesis_wav2flm(wav, scoeff, K_gamma, Phi2, Slm, J, B, N, bl_scoeff, alpha):
    """s2dw_core_synthesis_wav2flm, s2dw_core_mod.f90:816-1019 -> flm (static wav input)."""
    wav = np.asfortranarray(wav, dtype=np.float64)
    sc = np.asfortranarray(scoeff, dtype=np.complex128)
    K = np.asfortranarray(K_gamma, dtype=np.float64)
    S = np.asfortranarray(Slm, dtype=np.complex128)
    P = np.ascontiguousarray(Phi2, dtype=np.float64)
    flm = np.zeros((B, B), dtype=np.complex128, order="F")
    _check(lib().s2dw_hip_synthesis_wav2flm(_p(flm), _p(wav), _p(sc), _p(K), _p(P), _p(S), int(J), int(B), int(N),
                                            int(bl_scoeff), float(alpha)), "s2dw_core_synthesis_wav2flm")
    return flm


def s2dw_core_synthesis_wav2flm_dynamic(wavdyn, scoeff, K_gamma, Phi2, Slm, J, B, N, bl_scoeff, alpha):
    """s2dw_core_synthesis_wav2flm_dynamic, s2dw_core_mod.f90:1060-1266 -> flm.

    The Fortran routine deallocates wavdyn (1130, 1264); Python frees it when
    the caller drops its references."""
    wavs = [np.asfortranarray(w, dtype=np.float64) for w in wavdyn]
    ptrs = (_dp * (J + 1))(*[_p(w) for w in wavs])
    sc = np.asfortranarray(scoeff, dtype=np.complex128)
    K = np.asfortranarray(K_gamma, dtype=np.float64)
    S = np.asfortranarray(Slm, dtype=np.complex128)
    P = np.ascontiguousarray(Phi2, dtype=np.float64)
    flm = np.zeros((B, B), dtype=np.complex128, order="F")
    _check(lib().s2dw_hip_synthesis_wav2flm_dynamic(_p(flm), ptrs, _p(sc), _p(K), _p(P), _p(S), int(J), int(B),
                                                    int(N), int(bl_scoeff), float(alpha)),
           "s2dw_core_synthesis_wav2flm_dynamic")
    return flm


def s2dw_stat_moments(wavdyn, J, B, N, alpha, wavdyn_mask=None):
    """s2dw_stat_moments, s2dw_stat_mod.f90:81-275 -> (mean, var, skew, kur, five, six),
    each of length J+1.

    With ``wavdyn_mask`` (per-scale arrays shaped like ``wavdyn``) the
    coefficients are multiplied by the mask in place, as the Fortran routine
    does with its intent(inout) argument.  nj < 2 on a scale raises
    :class:`S2dwError` (a halting error in the reference); a variance below
    TOL_ZERO emits a warning (the reference's non-halting
    S2DW_ERROR_ARTH_WARNING) and returns zero skewness and higher moments for
    that scale."""
    n = int(J) + 1
    if len(wavdyn) != n or (wavdyn_mask is not None and len(wavdyn_mask) != n):
        raise S2dwError(S2DW_ERROR_ARG_INVALID, "s2dw_stat_moments", "one array per scale 0..J is required")
    hi, _ = band_limits(J, B, alpha)
    wavs, masks = [], []
    for j in range(n):
        shape = (2 * hi[j] - 1, 2 * hi[j], N)
        w = np.asfortranarray(wavdyn[j], dtype=np.float64)
        if w.shape != shape:
            raise S2dwError(S2DW_ERROR_SIZE_INVALID, "s2dw_stat_moments", f"scale {j}: shape {w.shape} != {shape}")
        wavs.append(w)
        if wavdyn_mask is not None:
            m = np.asfortranarray(wavdyn_mask[j], dtype=np.float64)
            if m.shape != shape:
                raise S2dwError(S2DW_ERROR_SIZE_INVALID, "s2dw_stat_moments", f"mask {j}: shape {m.shape}")
            masks.append(m)
    wp = (_dp * n)(*[_p(w) for w in wavs])
    mp = (_dp * n)(*[_p(m) for m in masks]) if masks else None
    out = [np.zeros(n) for _ in range(6)]
    code = lib().s2dw_hip_stat_moments(wp, n - 1, int(B), int(N), float(alpha), *[_p(o) for o in out[:4]], mp,
                                       _p(out[4]), _p(out[5]))
    if masks:
        for w, w0 in zip(wavs, wavdyn):
            if w is not w0:
                np.copyto(w0, w)
    if code == S2DW_ERROR_ARTH_WARNING:
        import warnings

        warnings.warn(str(S2dwError(code, "s2dw_stat_moments", lib().s2dw_hip_last_error().decode())),
                      RuntimeWarning, stacklevel=2)
    else:
        _check(code, "s2dw_stat_moments")
    return tuple(out)


def wigner_plane(l: int, which: str = "both"):
    """Delta^l_{m'm}(pi/2), 0 <= m', m <= l, from the device recursions:
    (l-recursion plane, m'-recursion plane), each indexed [m', m]."""
    n = l + 1
    a = np.zeros((n, n), dtype=np.float64, order="F")
    b = np.zeros((n, n), dtype=np.float64, order="F")
    _check(lib().s2dw_hip_wigner_plane(int(l), _p(a) if which in ("both", "l") else None,
                                       _p(b) if which in ("both", "m") else None), "wigner_plane")
    return a, b


# ---------------------------------------------------------------- device plans
class Plan:
    """Device-resident transform plan for one (B, N, J, alpha) configuration.

    Device buffers are torch tensors on the plan's device (torch provides the
    memory and the stream; the work runs in the HIP library).  The plan runs
    on the torch stream current at creation unless ``stream`` is given.
    """

    def __init__(self, B, N, J, alpha, K_gamma, Phi2, Slm, precision=PREC_FP64, stream=None):
        import torch

        self.B, self.N, self.J, self.alpha = int(B), int(N), int(J), float(alpha)
        K = np.asfortranarray(K_gamma, dtype=np.float64)
        P = np.ascontiguousarray(Phi2, dtype=np.float64)
        S = np.asfortranarray(Slm, dtype=np.complex128)
        if stream is None:
            stream = torch.cuda.current_stream()
        self.torch_stream = stream
        h = C.c_void_p()
        _check(lib().s2dw_hip_plan_create(C.byref(h), self.B, self.N, self.J, self.alpha, _p(K), _p(P), _p(S),
                                          int(precision), C.c_void_p(stream.cuda_stream)), "s2dw_hip_plan_create")
        self.h = h
        wd = C.c_longlong()
        bl = C.c_int()
        offs = (C.c_longlong * (J + 1))()
        hi = (C.c_int * (J + 1))()
        lo = (C.c_int * (J + 1))()
        _check(lib().s2dw_hip_plan_sizes(h, C.byref(wd), C.byref(bl), offs, hi, lo), "s2dw_hip_plan_sizes")
        self.wav_doubles = int(wd.value)
        self.bl_scoeff = int(bl.value)
        self.wav_offsets = list(offs)
        self.bl_hi = list(hi)
        self.bl_lo = list(lo)

    @property
    def device_bytes(self) -> int:
        return int(lib().s2dw_hip_plan_device_bytes(self.h))

    @property
    def wigner_form(self) -> int:
        """WIGNER_TABLE or WIGNER_STREAMED: where the contractions take Delta from."""
        return int(lib().s2dw_hip_plan_wigner_form(self.h))

    def wigner_table(self, l: int) -> np.ndarray:
        """Plane l of the plan's reference-rounding table, indexed [m', m]."""
        out = np.zeros((l + 1, l + 1), dtype=np.float64)  # row-major: out[m', m]
        _check(lib().s2dw_hip_plan_wigner_table(self.h, int(l), _p(out)), "s2dw_hip_plan_wigner_table")
        return out

    def empty_wav(self, device="cuda"):
        import torch

        return torch.empty(self.wav_doubles, dtype=torch.float64, device=device)

    def empty_flm(self, device="cuda"):
        import torch

        return torch.empty((self.B, self.B), dtype=torch.complex128, device=device)

    def empty_scoeff(self, device="cuda"):
        import torch

        return torch.empty((self.bl_scoeff, self.bl_scoeff), dtype=torch.complex128, device=device)

    def scale_view(self, wav, j: int):
        """View of scale j of a packed wav tensor as [g][b][a] (== Fortran (a,b,g))."""
        h = self.bl_hi[j]
        n = (2 * h - 1) * 2 * h * self.N
        o = self.wav_offsets[j]
        return wav[o:o + n].view(self.N, 2 * h, 2 * h - 1)

    def analysis(self, d_flm, d_wav, d_scoeff):
        """Enqueue analysis: d_flm (B x B complex, Fortran order i.e. flm[m][l] in torch),
        d_wav (packed), d_scoeff (bl x bl complex, Fortran order)."""
        _check(lib().s2dw_hip_analysis_device(self.h, C.c_void_p(d_wav.data_ptr()), C.c_void_p(d_scoeff.data_ptr()),
                                              C.c_void_p(d_flm.data_ptr())), "s2dw_hip_analysis_device")

    def synthesis(self, d_wav, d_scoeff, d_flm):
        _check(lib().s2dw_hip_synthesis_device(self.h, C.c_void_p(d_flm.data_ptr()), C.c_void_p(d_wav.data_ptr()),
                                               C.c_void_p(d_scoeff.data_ptr())), "s2dw_hip_synthesis_device")

    def sync(self):
        _check(lib().s2dw_hip_plan_sync(self.h), "s2dw_hip_plan_sync")

    MOMENT_ROWS = ["mean", "var", "skew", "kur", "five", "six", "nj", "flag"]

    def stat_moments(self, d_wav, d_mask=None, d_out=None):
        """Enqueue the per-scale moments of a packed wav (s2dw_stat_moments) on
        the plan's stream; returns d_out, a (8, J+1) float64 device tensor with
        rows MOMENT_ROWS (flag 1: variance below TOL_ZERO, 2: nj < 2).  With
        d_mask the coefficients of d_wav are masked in place."""
        import torch

        if d_out is None:
            d_out = torch.empty((8, self.J + 1), dtype=torch.float64, device=d_wav.device)
        mask = C.c_void_p(d_mask.data_ptr()) if d_mask is not None else None
        _check(lib().s2dw_hip_stat_moments_device(self.h, C.c_void_p(d_wav.data_ptr()), mask,
                                                  C.c_void_p(d_out.data_ptr())), "s2dw_hip_stat_moments_device")
        return d_out

    STAGES = ["ana_relayout", "ana_contract", "ana_beta_fft", "ana_alpha_c2r", "ana_scoeff", "syn_alpha_r2c",
              "syn_beta_fft", "syn_contract", "syn_assemble", "ana_small_ffts", "syn_small_ffts"]

    @property
    def moments_fusable(self) -> bool:
        """Whether analysis_moments can run on this plan."""
        return bool(lib().s2dw_hip_plan_moments_fusable(self.h))

    def analysis_moments(self, d_flm, d_scoeff, d_out=None):
        """Enqueue an analysis that reduces the wavelet coefficients to their
        per-scale moments without storing them; returns d_out ((8, J+1)
        float64, rows MOMENT_ROWS) and fills d_scoeff."""
        import torch

        if d_out is None:
            d_out = torch.empty((8, self.J + 1), dtype=torch.float64, device=d_flm.device)
        _check(lib().s2dw_hip_analysis_moments_device(self.h, C.c_void_p(d_out.data_ptr()),
                                                      C.c_void_p(d_scoeff.data_ptr()),
                                                      C.c_void_p(d_flm.data_ptr())),
               "s2dw_hip_analysis_moments_device")
        return d_out

    def set_timing(self, enable: bool):
        _check(lib().s2dw_hip_plan_set_timing(self.h, int(bool(enable))), "s2dw_hip_plan_set_timing")

    def autotune(self, reps: int = 3) -> str:
        """Time the switchable kernel paths on the device and keep the fastest;
        returns the measured times and the choices."""
        buf = C.create_string_buffer(1024)
        _check(lib().s2dw_hip_plan_autotune(self.h, int(reps), buf, len(buf)), "s2dw_hip_plan_autotune")
        return buf.value.decode()

    def set_graphs(self, enable: bool):
        """Replay each (direction, buffers) combination as a captured HIP graph."""
        _check(lib().s2dw_hip_plan_set_graphs(self.h, int(bool(enable))), "s2dw_hip_plan_set_graphs")

    def stage_times(self):
        """{stage: (total ms, launches)} for everything recorded since the last call."""
        ms = np.zeros(len(self.STAGES), dtype=np.float64)
        cnt = np.zeros(len(self.STAGES), dtype=np.int32)
        _check(lib().s2dw_hip_plan_stage_times(self.h, _p(ms), cnt.ctypes.data_as(_ip)), "s2dw_hip_plan_stage_times")
        return {s: (float(ms[i]), int(cnt[i])) for i, s in enumerate(self.STAGES)}

    def close(self):
        if getattr(self, "h", None):
            lib().s2dw_hip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stat_moments_device_noalloc(d_wav, J, B, N, alpha, d_out, d_mask=None, stream=None):
    """Per-scale moments of packed device coefficients without any allocation
    (s2dw_hip_stat_moments_device_noalloc): d_wav, d_mask (or None) and d_out
    (8 x (J+1) float64) are torch tensors on the current device; enqueued on
    `stream` (default: torch's current stream)."""
    import torch

    if stream is None:
        stream = torch.cuda.current_stream()
    _check(lib().s2dw_hip_stat_moments_device_noalloc(
        C.c_void_p(d_wav.data_ptr()), C.c_void_p(d_mask.data_ptr()) if d_mask is not None else None, int(J), int(B),
        int(N), float(alpha), C.c_void_p(d_out.data_ptr()), C.c_void_p(stream.cuda_stream)),
        "s2dw_hip_stat_moments_device_noalloc")
    return d_out
