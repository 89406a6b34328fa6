"""FITS wavelet files (s2dw_amd/fortran/s2dw_fileio_mod.f90) against files the
reference's own writer produced.

tests/golden/ref_wav_*.fits were written by the reference's
s2dw_fileio_fits_wav_write (src/mod/s2dw_fileio_mod.f90:796-926 static,
956-1077 dynamic) through s2dw_hip_fits_test; tests/golden/gen_ref_fixtures.py
made them.  The module here must write the same file card for card and byte
for byte (the DATE card aside), and read the reference's files back.  No GPU
work is involved: the programs only move coefficients through CFITSIO.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FDIR = os.path.join(ROOT, "s2dw_amd", "fortran")
FITS_TEST = os.path.join(FDIR, "_build", "s2dw_hip_fits_test")
REF_FITS = os.path.join(ROOT, "oracle", "_ref", "ref_fits")
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = [("dynamic", 8, 2, 2.0), ("static", 12, 3, 3.0)]

pytestmark = pytest.mark.skipif(shutil.which("amdflang") is None, reason="no Fortran compiler")


def golden(layout, B, N, alpha):
    return os.path.join(GOLDEN, f"ref_wav_{layout}_B{B}_N{N}_a{int(alpha)}.fits")


@pytest.fixture(scope="module", autouse=True)
def built():
    r = subprocess.run(["make", "-C", FDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(FITS_TEST)


def hdus(path):
    """(cards, data bytes) per HDU of a FITS file: 2880-byte blocks of 80-byte
    header cards up to END, then the padded data array."""
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    out, i = [], 0
    while i < len(raw):
        cards = []
        while True:
            blk = raw[i:i + 2880]
            i += 2880
            cards += [blk[k:k + 80].decode("ascii") for k in range(0, 2880, 80)]
            if any(c.rstrip() == "END" for c in cards[-36:]):
                break
        cards = [c for c in cards if c.strip()]
        kv = {c[:8].strip(): c[10:].split("/")[0].strip() for c in cards if c[8:10] == "= "}
        n = int(kv["NAXIS"]) and int(np.prod([int(kv[f"NAXIS{a + 1}"]) for a in range(int(kv["NAXIS"]))]))
        nbytes = n * abs(int(kv["BITPIX"])) // 8
        out.append((cards, kv, raw[i:i + nbytes]))
        i += (nbytes + 2879) // 2880 * 2880
    return out


def run(*args):
    cmd = [str(a) for a in args]
    print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=120)


def fits_value(jj, a, b, g):
    # the coefficients s2dw_hip_fits_test writes
    return np.sin(0.1 * a + 0.2 * b + 0.3 * g + jj)


@pytest.mark.parametrize("layout,B,N,alpha", CASES)
def test_reference_file_layout(layout, B, N, alpha):
    # the reference's header keys (s2dw_fileio_mod.f90:1018-1022), scoeff
    # images (1027-1041) and one WAV cube per scale (1045-1064)
    import oracle as O

    J = O.comp_Jmax(B, alpha)
    his, _ = O.band_limits(J, B, alpha)
    h = hdus(golden(layout, B, N, alpha))
    assert len(h) == 3 + J + 1
    kv = h[0][1]
    assert (int(kv["J"]), int(kv["B"]), int(kv["N"])) == (J, B, N)
    bl = int(kv["BLSCOEFF"])
    assert float(kv["ALPHA"]) == alpha
    for k, part in ((1, "real"), (2, "imag")):
        assert h[k][1]["EXTNAME"].strip("' ") == f"SCOEFF_{part.upper()}"
        img = np.frombuffer(h[k][2], dtype=">f8").reshape(bl, bl)  # FITS: first axis fastest
        l, m = np.meshgrid(np.arange(bl), np.arange(bl), indexing="xy")
        want = (l + 0.5 * m) if part == "real" else -0.25 * l
        np.testing.assert_array_equal(img, want)
    for jj, hi in enumerate(his):
        _, kvj, data = h[3 + jj]
        assert kvj["EXTNAME"].strip("' ") == "WAV"
        shape = tuple(int(kvj[f"NAXIS{a}"]) for a in (1, 2, 3))
        assert shape == (2 * hi - 1, 2 * hi, N)
        cube = np.frombuffer(data, dtype=">f8").reshape(shape[::-1])  # [g][b][a]
        g, b, a = np.meshgrid(np.arange(N), np.arange(2 * hi), np.arange(2 * hi - 1), indexing="ij")
        np.testing.assert_allclose(cube, fits_value(jj, a, b, g), rtol=0, atol=2e-16)


@pytest.mark.parametrize("layout,B,N,alpha", CASES)
@pytest.mark.parametrize("ours", ["static", "dynamic"])
def test_writer_matches_reference_file(tmp_path, layout, B, N, alpha, ours):
    out = tmp_path / "w.fits"
    r = run(FITS_TEST, "write_" + ours, out, B, N, alpha)
    assert r.returncode == 0, r.stdout + r.stderr
    mine, ref = hdus(str(out)), hdus(golden(layout, B, N, alpha))
    assert len(mine) == len(ref)
    for k, ((c1, _, d1), (c2, _, d2)) in enumerate(zip(mine, ref)):
        c1 = [c for c in c1 if not c.startswith("DATE    =")]
        c2 = [c for c in c2 if not c.startswith("DATE    =")]
        assert c1 == c2, f"HDU {k} header"
        assert d1 == d2, f"HDU {k} data"
    assert os.path.getsize(out) == os.path.getsize(golden(layout, B, N, alpha))


def parse_read(stdout):
    head = re.search(r"^header\s+(.*)$", stdout, re.M).group(1).split()
    scales = [line.split() for line in stdout.splitlines() if line.startswith("scale")]
    sc = float(re.search(r"scoeff\s+([0-9.Ee+-]+)", stdout).group(1))
    return head, scales, sc


@pytest.mark.parametrize("layout,B,N,alpha", CASES)
@pytest.mark.parametrize("ours", ["static", "dynamic"])
def test_reader_reads_reference_file(layout, B, N, alpha, ours):
    import oracle as O

    r = run(FITS_TEST, "read_" + ours, golden(layout, B, N, alpha))
    assert r.returncode == 0, r.stdout + r.stderr
    head, scales, sc = parse_read(r.stdout)
    J = O.comp_Jmax(B, alpha)
    his, _ = O.band_limits(J, B, alpha)
    assert [int(v) for v in head[:3]] == [J, B, N] and float(head[4]) == alpha
    assert len(scales) == J + 1
    for jj, (row, hi) in enumerate(zip(scales, his)):
        assert [int(v) for v in row[1:5]] == [jj, 2 * hi - 1, 2 * hi, N]
        assert float(row[5]) == 0.0
    assert sc == 0.0


@pytest.mark.skipif(not os.path.exists(REF_FITS), reason="reference FITS module not built (needs reference sources)")
@pytest.mark.parametrize("ours", ["static", "dynamic"])
def test_reference_reader_reads_our_file(tmp_path, ours):
    out = tmp_path / "w.fits"
    assert run(FITS_TEST, "write_" + ours, out, 16, 3, 2.0).returncode == 0
    for reader in ("read_static", "read_dynamic"):
        r = run(REF_FITS, reader, out)
        assert r.returncode == 0, r.stdout + r.stderr
        _, scales, sc = parse_read(r.stdout)
        assert all(float(row[5]) == 0.0 for row in scales) and sc == 0.0


def test_existing_file_refused(tmp_path):
    # s2dw_fileio_mod.f90:824-830: an existing file halts through s2dw_error
    out = tmp_path / "w.fits"
    assert run(FITS_TEST, "write_dynamic", out, 8, 2, 2.0).returncode == 0
    before = out.read_bytes()
    r = run(FITS_TEST, "write_static", out, 8, 2, 2.0)
    assert ("S2DW_ERROR: Error 'File IO error' occured in procedure 's2dw_fileio_fits_wav_write'"
            in r.stdout)
    assert "File already exists" in r.stdout and "Halting program execution" in r.stdout
    assert out.read_bytes() == before


@pytest.mark.parametrize("ours", ["static", "dynamic"])
def test_missing_file_refused(tmp_path, ours):
    # s2dw_fileio_mod.f90:1134-1140
    r = run(FITS_TEST, "read_" + ours, tmp_path / "none.fits")
    assert "S2DW_ERROR: Error 'File IO error' occured in procedure 's2dw_fileio_fits_wav_read'" in r.stdout
    assert "File does not exist" in r.stdout and "Halting program execution" in r.stdout
