"""Write golden vectors of the reference's own setup routines and Wigner planes.

Runs oracle/_ref/ref_dump (the reference's s2dw_core_mod / s2dw_dl_mod /
s2dw_stat_mod compiled from its Fortran sources by ``make -C oracle ref``) and
stores its outputs as
small .npz fixtures beside this script.  Run in a container that holds the
reference sources; the fixtures are committed, the GPU box only reads them.

    python tests/golden/gen_ref_fixtures.py [--moments-only | --fits-only]

The FITS and MATLAB fixtures are wavelet files written by the reference's
s2dw_fileio_fits_wav_write / s2dw_fileio_matlab_wav_write (oracle/_ref/
ref_fits: s2dw_amd/fortran/s2dw_hip_fits_test.f90 linked against the
reference's s2dw_fileio_mod); the MATLAB ones are written from this folder
with relative names, which the .m file embeds.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
REF_FITS = os.path.join(ROOT, "oracle", "_ref", "ref_fits")
# (layout, B, N, alpha): the wavelet files written by the reference
FITS_CASES = [("dynamic", 8, 2, 2.0), ("static", 12, 3, 3.0)]
# (layout, B, N, alpha, J): MATLAB .m/.dat files written by the reference; J
# below J_max so that there is more than one scaling coefficient
MATLAB_CASES = [("dynamic", 8, 2, 2.0, 1), ("static", 6, 2, 3.0, 1)]

# (B, alpha, J); J=None means J_max.  Covers the test config (64,2,full J),
# L=256 N=5 full J, L=512 J_min=2 (J=7), alpha=3 and non-power-of-alpha B.
KERNEL_CASES = [(8, 2.0, None), (16, 2.0, None), (64, 2.0, None), (64, 2.0, 3), (64, 3.0, None),
                (100, 2.0, None), (37, 3.0, 2), (81, 3.0, None), (256, 2.0, None), (512, 2.0, 7),
                (1024, 2.0, None), (200, 3.0, None)]
DIREC_CASES = [(16, 1), (16, 2), (16, 3), (16, 4), (16, 5), (16, 8), (32, 7), (12, 12)]
DL_FULL = [1, 2, 3, 4, 5, 7, 8, 15, 16, 31, 32, 64]
DL_BIG = [100, 127, 255, 256, 511, 1023]


def run(*args) -> bytes:
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "o.bin")
        subprocess.run([REF_DUMP, *map(str, args), out], check=True, capture_output=True)
        with open(out, "rb") as fh:
            return fh.read()


def jmax_of(B, alpha):
    b = run("bands", B, alpha, 1)
    return int(np.frombuffer(b[:4], dtype="<i4")[0])


# s2dw_stat_moments cases: (name, B, N, alpha, data, mask); J = J_max.
MOMENT_CASES = [("plain", 16, 3, 2.0, "normal", None), ("mask", 16, 3, 2.0, "normal", "binary"),
                ("frac", 16, 3, 2.0, "normal", "fractional"), ("const", 8, 2, 2.0, "const1", None),
                ("offset", 27, 1, 3.0, "offset_exp", None), ("gamma", 16, 2, 2.0, "gamma", "binary")]


def scale_his(B, alpha, J):
    bb = np.frombuffer(run("bands", B, alpha, J), dtype="<i4")
    return [int(h) for h in bb[1:].reshape(J + 1, 2)[:, 0]]


def moments_fixtures():
    out = {}
    rng = np.random.default_rng(2012)
    for name, B, N, alpha, data, mask in MOMENT_CASES:
        J = jmax_of(B, alpha)
        sizes = [(2 * h - 1) * 2 * h * N for h in scale_his(B, alpha, J)]
        tot = sum(sizes)
        if data == "normal":
            x = rng.standard_normal(tot)
        elif data == "offset_exp":
            x = 1000.0 + rng.exponential(1.0, tot)
        elif data == "gamma":
            x = rng.gamma(2.0, 0.5, tot) - 1.0
        else:  # "const1": scale 1 constant (zero variance), the others normal
            x = rng.standard_normal(tot)
            x[sizes[0]:sizes[0] + sizes[1]] = 0.25
        m = None
        if mask == "binary":
            m = (rng.random(tot) < 0.75).astype(np.float64)
        elif mask == "fractional":
            m = rng.choice(np.array([0.0, 0.5, 1.0, 1.0 + 5e-13, 1.0]), size=tot)
        with tempfile.TemporaryDirectory() as td:
            inp = os.path.join(td, "in.bin")
            with open(inp, "wb") as fh:
                fh.write(x.astype("<f8").tobytes())
                if m is not None:
                    fh.write(m.astype("<f8").tobytes())
            b = run("moments", B, N, alpha, J, 1 if m is not None else 0, inp)
        d = np.frombuffer(b, dtype="<f8")
        k = f"{name}_"
        out[k + "meta"] = np.array([B, N, J], dtype=np.int64)
        out[k + "alpha"] = np.array([alpha])
        out[k + "x"] = x
        if m is not None:
            out[k + "mask"] = m
        out[k + "moments"] = d[:6 * (J + 1)].reshape(6, J + 1)
        after = d[6 * (J + 1):]
        if m is None:
            assert np.array_equal(after, x), name  # unmasked: the input is left alone
        else:
            out[k + "x_after"] = after
    np.savez_compressed(os.path.join(HERE, "ref_moments.npz"), **out)
    # s2dw_stat_moments_write: the reference's text format, byte for byte
    for nopt in (0, 1, 2):
        with open(os.path.join(HERE, f"ref_moments_write_{nopt}.txt"), "wb") as fh:
            fh.write(run("momwrite", 3, 4, nopt))


def fits_name(layout, B, N, alpha):
    return f"ref_wav_{layout}_B{B}_N{N}_a{int(alpha)}.fits"


def matlab_name(layout, B, N, alpha, J):
    return f"ref_mat_{layout}_B{B}_N{N}_a{int(alpha)}_J{J}"


def fits_fixtures():
    for layout, B, N, alpha, J in MATLAB_CASES:
        base = matlab_name(layout, B, N, alpha, J)
        for suffix in (".m", "_scoeff.dat", "_wcoeff.dat"):
            if os.path.exists(os.path.join(HERE, base + suffix)):
                os.remove(os.path.join(HERE, base + suffix))  # the writer refuses existing files
        subprocess.run([REF_FITS, "mwrite_" + layout, base + ".m", str(B), str(N), str(alpha), str(J)], check=True,
                       capture_output=True, cwd=HERE)
    for layout, B, N, alpha in FITS_CASES:
        out = os.path.join(HERE, fits_name(layout, B, N, alpha))
        if os.path.exists(out):
            os.remove(out)  # the writer refuses an existing file
        subprocess.run([REF_FITS, "write_" + layout, out, str(B), str(N), str(alpha)], check=True,
                       capture_output=True)


def main():
    if not os.path.exists(REF_DUMP) or not os.path.exists(REF_FITS):
        raise SystemExit("build the reference drivers first: make -C oracle ref")
    if "--fits-only" in sys.argv:
        fits_fixtures()
        print("fixtures written to", HERE)
        return
    if "--moments-only" in sys.argv:
        moments_fixtures()
        print("fixtures written to", HERE)
        return
    fits_fixtures()
    moments_fixtures()
    # --- kernels / admissibility / band limits
    kern = {}
    for B, alpha, J in KERNEL_CASES:
        Jm = jmax_of(B, alpha)
        J = Jm if J is None else J
        b = run("kernels", B, alpha, J)
        jmax, bl, ok = np.frombuffer(b[:12], dtype="<i4")
        d = np.frombuffer(b[12:], dtype="<f8")
        nK = (J + 1) * B
        key = f"B{B}_a{int(alpha)}_J{J}"
        kern[key + "_K"] = d[:nK].reshape((J + 1, B), order="F")
        kern[key + "_Phi2"] = d[nK:nK + B]
        kern[key + "_admiss"] = d[nK + B:nK + 2 * B]
        bb = np.frombuffer(run("bands", B, alpha, J), dtype="<i4")
        kern[key + "_meta"] = np.array([B, J, int(jmax), int(bl), int(ok)], dtype=np.int64)
        kern[key + "_alpha"] = np.array([alpha])
        kern[key + "_bands"] = bb[1:].reshape(J + 1, 2).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "ref_kernels.npz"), **kern)
    # --- J_max over a sweep
    sweep = []
    for alpha in (2.0, 3.0, 1.5, 2.5):
        for B in (2, 3, 4, 5, 7, 8, 9, 16, 27, 31, 32, 33, 63, 64, 65, 81, 100, 128, 243, 255, 256, 512,
                  729, 1000, 1024):
            sweep.append((B, alpha, jmax_of(B, alpha)))
    np.save(os.path.join(HERE, "ref_jmax.npy"), np.array(sweep, dtype=np.float64))
    # --- directionality
    direc = {}
    for B, N in DIREC_CASES:
        b = run("direc", B, N)
        direc[f"B{B}_N{N}"] = np.frombuffer(b, dtype="<c16").reshape((B, N), order="F")
    np.savez_compressed(os.path.join(HERE, "ref_direc.npz"), **direc)
    # --- Wigner planes at beta = pi/2
    dls = {}
    for l in DL_FULL:
        b = run("dl", l)
        dls[f"l{l}"] = np.frombuffer(b, dtype="<f8").reshape((2 * l + 1, 2 * l + 1), order="F")
    rng = np.random.default_rng(7)
    for l in DL_BIG:
        b = run("dl", l)
        P = np.frombuffer(b, dtype="<f8").reshape((2 * l + 1, 2 * l + 1), order="F")
        Q = P[l:, l:]  # quadrant m', m >= 0: Q[m', m] = d^l_{m' m}
        dls[f"big{l}_cols"] = np.ascontiguousarray(Q[:, :9])          # m < 9, all m'
        rows = sorted({0, 1, 2, 3, l // 3, l // 2, (2 * l) // 3, l - 1, l})
        dls[f"big{l}_rowidx"] = np.array(rows, dtype=np.int64)
        dls[f"big{l}_rows"] = np.ascontiguousarray(Q[rows, :])
        idx = rng.integers(0, l + 1, size=(3000, 2))
        dls[f"big{l}_sidx"] = idx.astype(np.int64)
        dls[f"big{l}_svals"] = Q[idx[:, 0], idx[:, 1]]
        dls[f"big{l}_neg"] = np.ascontiguousarray(P[l:, :l][:, ::-1][:, :8])  # d^l_{m', -k}, k=1..8
    np.savez_compressed(os.path.join(HERE, "ref_dl.npz"), **dls)
    # --- a Wigner plane at a generic beta (the recursion's general branch)
    gen = {}
    for l, beta in ((5, 0.3), (12, 1.1), (20, 2.5)):
        b = run("dlbeta", l, beta)
        gen[f"l{l}_b{beta}"] = np.frombuffer(b, dtype="<f8").reshape((2 * l + 1, 2 * l + 1), order="F")
    np.savez_compressed(os.path.join(HERE, "ref_dl_beta.npz"), **gen)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
