"""Python side of the request loop in libs2dw_hip.so (s2dw_hip_server_*).

    srv = Server(hbm_budget_bytes=64 << 30, max_inflight=8)
    ta = srv.submit_analysis(B, N, J, alpha, flm, keep_resident=True)
    ts = srv.submit_synthesis(B, N, J, alpha, wav_from=ta)
    flm2 = srv.result(ts)             # waits; raises S2dwError on failure
    wavs, scoeff = srv.result(ta)
    srv.release(ta)                   # frees the resident coefficients

Arrays use the layouts of the dynamic Fortran routines: flm (B, B) complex
(Fortran order, flm[l, m]), one (2hi-1, 2hi, N) real array per scale,
scoeff (bl, bl) complex.  A ticket object keeps its host arrays alive until
its result has been taken.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .core import S2dwError, _check, band_limits, lib

_dp = C.POINTER(C.c_double)
_ll = C.c_longlong
_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    L.s2dw_hip_server_create.argtypes = [C.POINTER(C.c_void_p), _ll, C.c_int]
    L.s2dw_hip_server_create_ex.argtypes = [C.POINTER(C.c_void_p), _ll, C.c_int, C.c_int]
    L.s2dw_hip_server_destroy.argtypes = [C.c_void_p]
    L.s2dw_hip_server_submit_analysis.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp,
                                                  C.c_int, C.POINTER(_ll)]
    L.s2dw_hip_server_submit_synthesis.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, _ll, _dp, _dp,
                                                   _dp, C.POINTER(_ll)]
    L.s2dw_hip_server_wait.argtypes = [C.c_void_p, _ll, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.s2dw_hip_server_release.argtypes = [C.c_void_p, _ll]
    L.s2dw_hip_server_stats.argtypes = [C.c_void_p, C.POINTER(_ll), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(_ll)]
    _bound = True
    return L


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def bl_scoeff_of(J, B, alpha):
    hi, _ = band_limits(J + 1, B, alpha)
    return hi[-1]


@dataclass
class Ticket:
    id: int
    kind: str
    B: int
    N: int
    J: int
    alpha: float
    keep: list = field(default_factory=list)  # host arrays in use by the request
    out: object = None
    latency_ms: float = float("nan")
    run_ms: float = float("nan")
    done: bool = False


class Server:
    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 8, precision: int = 0):
        """precision: core.PREC_FP64 (default) or core.PREC_SPLIT, OR-ed with a
        rounding form (core.WIGNER_TABLE, core.ROUNDING_REFERENCE)."""
        self._L = _bind()
        h = C.c_void_p()
        _check(self._L.s2dw_hip_server_create_ex(C.byref(h), int(hbm_budget_bytes), int(max_inflight),
                                                 int(precision)), "s2dw_hip_server_create_ex")
        self._h = h

    def submit_analysis(self, B, N, J, alpha, flm, keep_resident=False, want_wav=True) -> Ticket:
        flm = np.asfortranarray(flm, dtype=np.complex128)
        hi, _ = band_limits(J, B, alpha)
        packed = np.zeros(sum((2 * h - 1) * 2 * h * N for h in hi), dtype=np.float64) if want_wav else None
        bl = bl_scoeff_of(J, B, alpha)
        sc = np.zeros((bl, bl), dtype=np.complex128, order="F")
        t = _ll()
        _check(self._L.s2dw_hip_server_submit_analysis(self._h, int(B), int(N), int(J), float(alpha), _p(flm),
                                                       _p(packed), _p(sc), int(bool(keep_resident)), C.byref(t)),
               "s2dw_hip_server_submit_analysis")
        tk = Ticket(t.value, "analysis", B, N, J, alpha, keep=[flm, packed, sc])
        tk.out = (packed, sc, hi)
        return tk

    def submit_synthesis(self, B, N, J, alpha, wav_from: Ticket | None = None, wavs=None, scoeff=None) -> Ticket:
        flm = np.zeros((B, B), dtype=np.complex128, order="F")
        keep = [flm]
        if wav_from is not None:
            src, wp, sp = int(wav_from.id), None, None
        else:
            src = -1
            wp = np.concatenate([np.asfortranarray(w, dtype=np.float64).ravel(order="F") for w in wavs])
            sp = np.asfortranarray(scoeff, dtype=np.complex128)
            keep += [wp, sp]
        t = _ll()
        _check(self._L.s2dw_hip_server_submit_synthesis(self._h, int(B), int(N), int(J), float(alpha), src, _p(wp),
                                                        _p(sp), _p(flm), C.byref(t)),
               "s2dw_hip_server_submit_synthesis")
        tk = Ticket(t.value, "synthesis", B, N, J, alpha, keep=keep)
        tk.out = flm
        return tk

    def wait(self, tk: Ticket) -> Ticket:
        if not tk.done:
            lat, run = C.c_double(), C.c_double()
            code = self._L.s2dw_hip_server_wait(self._h, int(tk.id), C.byref(lat), C.byref(run))
            tk.done = True
            tk.latency_ms, tk.run_ms = lat.value, run.value
            _check(code, f"s2dw_hip_server_wait ({tk.kind}, ticket {tk.id})")
        return tk

    def result(self, tk: Ticket):
        self.wait(tk)
        if tk.kind == "synthesis":
            return tk.out
        packed, sc, hi = tk.out
        if packed is None:
            return None, sc
        wavs, off = [], 0
        for h in hi:
            n = (2 * h - 1) * 2 * h * tk.N
            wavs.append(packed[off:off + n].reshape((2 * h - 1, 2 * h, tk.N), order="F"))
            off += n
        return wavs, sc

    def release(self, tk: Ticket) -> None:
        _check(self._L.s2dw_hip_server_release(self._h, int(tk.id)), "s2dw_hip_server_release")

    def stats(self) -> dict:
        b, i, q, c = _ll(), C.c_int(), C.c_int(), _ll()
        _check(self._L.s2dw_hip_server_stats(self._h, C.byref(b), C.byref(i), C.byref(q), C.byref(c)),
               "s2dw_hip_server_stats")
        return {"in_use_bytes": b.value, "inflight": i.value, "queued": q.value, "completed": c.value}

    def close(self) -> None:
        if self._h:
            _check(self._L.s2dw_hip_server_destroy(self._h), "s2dw_hip_server_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ = ["Server", "Ticket", "S2dwError", "bl_scoeff_of"]
